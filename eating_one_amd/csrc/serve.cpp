// Request loop: continuous batching of a timed request stream.
//
// The reference has no serving path: every join is a separate process run
// (joins/hashJoin_gpu/main_street/hj.cc:1189-1269).  eo_serve_run replays a
// stream of requests with given arrival times against one device:
//   - a request becomes eligible at its arrival time (host steady clock);
//   - admission: a request's bytes (its inputs, partition scratch and output
//     buffer) count against the HBM budget while it is in flight; a batch
//     holds at most budget / slots, so the slots' scratch arenas never
//     exceed the budget together; a request larger than that runs alone once
//     the device is idle, and one larger than the whole budget is refused
//     (EO_EBUDGET);
//   - whenever a batch slot is free and requests are waiting, all waiting
//     requests that fit are submitted as one batch (continuous batching);
//     `slots` batches run at once on their own streams;
//   - a waiting request that is cancelled or outlives its admission deadline
//     leaves the queue without running (the policy is eo::ServeQueue);
//   - latency = completion (event observed by the loop) - arrival.
// Mode EO_SERVE_PER_JOIN serves the same stream one eo_hash_join call per
// request, the per-join launch sequence the batched mode is measured against.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <thread>
#include <vector>

#include "ctx.hpp"
#include "eojoin.h"
#include "kernels.hpp"
#include "plan.hpp"
#include "serve_queue.hpp"
#include "table.hpp"


namespace eo {
namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// Admission size of one request: inputs, partition copies + starts, output.
uint64_t request_bytes(const eo_request &q, int cus) {
  const uint64_t nR = q.table ? q.table->n : q.nR;
  uint64_t b = 8 * q.nS + 8 * q.out_cap + 256;
  if (!q.table) b += 8 * nR;
  if (nR == 0 || q.nS == 0) return b;
  HashPlan hp;
  if (q.table) {
    eo_join_opts o{};
    o.radix_bits = q.table->hp.bits ? q.table->hp.bits : -1;
    o.passes = q.table->hp.passes;
    if (plan_hash_join(nR, q.nS, &o, &hp, cus) != EO_OK) return b;
  } else if (plan_hash_join(nR, q.nS, nullptr, &hp, cus) != EO_OK) {
    return b;
  }
  const uint64_t P = 1ull << hp.bits;
  // probe tile descriptors (as the batch estimates them)
  b += sizeof(BatchTile) * P * std::max<uint64_t>(1, (nR / P + kBatchChunk - 1) / kBatchChunk) *
       ((q.nS / P + hp.s_span - 1) / hp.s_span + 1);
  if (hp.bits == 0) return b;
  const uint64_t len = batch_stripe_len(hp.bits);
  b += 8 * q.nS + 4 * (P + 1) + 8 * (((q.nS + len - 1) / len) << hp.bits) + 8 * P;
  if (!q.table) b += 8 * nR + 4 * (P + 1) + 8 * (((nR + len - 1) / len) << hp.bits);
  if (hp.passes == 2) b += 8 * std::max(nR, q.nS);
  return b;
}

struct Slot {
  eo_batch *b = nullptr;
  bool busy = false;
  bool exclusive = false;
  uint64_t bytes = 0;
  std::vector<uint32_t> idx;
  std::vector<eo_request> reqs;
  std::vector<eo_result> res;
};

void percentiles(std::vector<double> v, eo_serve_stats *st) {
  if (v.empty()) return;
  std::sort(v.begin(), v.end());
  auto pct = [&](double p) {
    // nearest-rank percentile
    size_t r = (size_t)std::ceil(p / 100.0 * (double)v.size());
    if (r < 1) r = 1;
    return v[std::min(r, v.size()) - 1];
  };
  st->p50_ms = pct(50);
  st->p99_ms = pct(99);
  st->max_ms = v.back();
  double s = 0;
  for (double x : v) s += x;
  st->mean_ms = s / (double)v.size();
}

}  // namespace
}  // namespace eo

using namespace eo;

extern "C" eo_status eo_serve_run(eo_ctx *c, const eo_request *reqs, const double *arrival_ms, uint32_t n,
                                  const eo_serve_opts *o, eo_result *results, double *latency_ms,
                                  eo_serve_stats *stats) {
  if (!c || (n && (!reqs || !results))) return EO_EINVAL;
  // a caller built against the struct without deadline_ms / cancel passes the shorter size
  const size_t v1_size = offsetof(eo_serve_opts, deadline_ms);
  if (o && o->struct_size != 0 && o->struct_size < v1_size) return EO_EINVAL;
  const bool has_v2 = o && (o->struct_size == 0 || o->struct_size >= sizeof(eo_serve_opts));
  const double *deadline = has_v2 ? o->deadline_ms : nullptr;
  const volatile uint32_t *cancel = has_v2 ? o->cancel : nullptr;
  const int mode = o ? o->mode : EO_SERVE_BATCHED;
  if (mode != EO_SERVE_BATCHED && mode != EO_SERVE_PER_JOIN) return EO_EINVAL;
  for (uint32_t i = 1; arrival_ms && i < n; ++i)
    if (!(arrival_ms[i] >= arrival_ms[i - 1])) return EO_EINVAL;
  for (uint32_t i = 0; deadline && i < n; ++i)
    if (deadline[i] != deadline[i]) return EO_EINVAL;  // NaN
  const uint32_t max_batch = (o && o->max_batch) ? o->max_batch : 512u;
  const uint32_t nslots = (o && o->slots) ? std::min<uint32_t>(o->slots, 8u) : 2u;
  const uint64_t budget = (o && o->hbm_budget) ? o->hbm_budget : c->hbm_budget;
  DeviceGuard g(c->device);
  std::vector<double> lat(n, 0.0);
  std::vector<uint8_t> ran(n, 0);
  eo_serve_stats st{};

  // batch objects (each with its stream and grow-only arena) persist in the
  // ctx between calls, so a warmed-up loop allocates nothing
  std::vector<Slot> slots(mode == EO_SERVE_BATCHED ? nslots : 0);
  while (c->serve_pool.size() < slots.size()) {
    eo_batch *b = nullptr;
    eo_status e = eo_batch_create(c, &b);
    if (e != EO_OK) return e;
    c->serve_pool.push_back(b);
  }
  for (size_t k = 0; k < slots.size(); ++k) slots[k].b = c->serve_pool[k];
  auto cleanup = [&]() {
    for (size_t k = 0; k < slots.size(); ++k) c->serve_pool[k] = slots[k].b;
    c->serve_pool.erase(std::remove(c->serve_pool.begin(), c->serve_pool.end(), nullptr), c->serve_pool.end());
  };
  std::vector<uint64_t> rbytes(n);
  for (uint32_t i = 0; i < n; ++i) rbytes[i] = request_bytes(reqs[i], c->props.cus);

  ServeQueueParams qp;
  qp.budget = budget;
  qp.batch_cap = mode == EO_SERVE_BATCHED ? budget / nslots : budget;
  qp.max_batch = mode == EO_SERVE_BATCHED ? max_batch : 1u;
  ServeQueue queue(n, arrival_ms, rbytes.data(), deadline, cancel, qp);
  std::vector<std::pair<uint32_t, eo_status>> gone;
  uint32_t done = 0;
  // requests that leave the queue without running
  auto retire = [&](double now) {
    gone.clear();
    queue.retire(now, &gone);
    for (const auto &r : gone) {
      eo_result res{};
      res.tag = reqs[r.first].tag;
      res.status = r.second;
      results[r.first] = res;
      lat[r.first] = std::max(0.0, now - queue.arrival(r.first));  // 0: cancelled before it arrived
      st.refused += r.second == EO_EBUDGET;
      st.expired += r.second == EO_ETIMEDOUT;
      st.cancelled += r.second == EO_ECANCELED;
    }
    done += (uint32_t)gone.size();
    return !gone.empty();
  };

  const Clock::time_point t0 = Clock::now();
  uint64_t in_flight_bytes = 0;
  eo_status fail = EO_OK;
  std::vector<uint32_t> idx;

  if (mode == EO_SERVE_PER_JOIN) {
    eo_join_opts jo{};
    jo.struct_size = sizeof(jo);
    jo.where = EO_DEVICE;
    while (done < n && fail == EO_OK) {
      retire(ms_since(t0));
      uint64_t bytes = 0;
      bool exclusive = false;
      if (!queue.take(ms_since(t0), false, 0, &idx, &bytes, &exclusive)) {
        if (queue.drained()) break;
        std::this_thread::yield();
        continue;
      }
      const uint32_t i = idx[0];
      const eo_request &q = reqs[i];
      eo_result r{};
      r.tag = q.tag;
      const eo_tuple *R = q.table ? q.table->part : q.R;
      const uint64_t nR = q.table ? q.table->n : q.nR;
      uint64_t cnt = 0;
      eo_status e = eo_hash_join(c, R, nR, q.S, q.nS, q.out, q.out ? q.out_cap : 0, &cnt, &jo);
      if (e != EO_OK && e != EO_ECAPACITY) fail = e;
      r.n_out = cnt;
      r.status = e == EO_ECAPACITY ? EO_ECAPACITY : EO_OK;
      lat[i] = ms_since(t0) - queue.arrival(i);
      ran[i] = 1;
      results[i] = r;
      st.n_out += r.n_out;
      st.batches += 1;
      st.largest_batch = 1;
      st.peak_admitted = std::max(st.peak_admitted, bytes);
      ++done;
    }
  }

  while (mode == EO_SERVE_BATCHED && done < n && fail == EO_OK) {
    bool progress = false;
    // completions
    for (auto &s : slots) {
      if (!s.busy) continue;
      eo_status p = eo_batch_poll(s.b);
      if (p == EO_EAGAIN) continue;
      if (p == EO_OK) {
        s.res.resize(s.idx.size());
        p = eo_batch_wait(s.b, s.res.data(), (uint32_t)s.idx.size());
      }
      const double now = ms_since(t0);
      s.busy = false;
      in_flight_bytes -= s.bytes;
      if (p != EO_OK) {
        fail = p;
        break;
      }
      for (size_t k = 0; k < s.idx.size(); ++k) {
        const uint32_t i = s.idx[k];
        results[i] = s.res[k];
        lat[i] = now - queue.arrival(i);
        ran[i] = 1;
        st.n_out += s.res[k].n_out;
      }
      done += (uint32_t)s.idx.size();
      if (s.exclusive) {  // give back an arena grown past the per-batch share
        eo_batch *nb = nullptr;
        eo_batch_free(s.b);
        s.b = nullptr;
        fail = eo_batch_create(c, &nb);
        s.b = nb;
        s.exclusive = false;
      }
      progress = true;
    }
    if (fail != EO_OK) break;
    // cancelled, expired and too large requests leave the queue
    progress |= retire(ms_since(t0));
    // admission into free slots
    bool any_busy = false;
    for (auto &s : slots) any_busy |= s.busy;
    for (auto &s : slots) {
      if (s.busy) continue;
      uint64_t bytes = 0;
      bool exclusive = false;
      if (!queue.take(ms_since(t0), any_busy, in_flight_bytes, &s.idx, &bytes, &exclusive)) break;
      s.reqs.clear();
      for (uint32_t i : s.idx) s.reqs.push_back(reqs[i]);
      s.bytes = bytes;
      s.exclusive = exclusive;
      eo_status e = eo_batch_submit(s.b, s.reqs.data(), (uint32_t)s.reqs.size(), nullptr);
      if (e != EO_OK) {
        fail = e;
        break;
      }
      s.busy = true;
      any_busy = true;
      in_flight_bytes += s.bytes;
      st.peak_admitted = std::max(st.peak_admitted, in_flight_bytes);
      st.batches += 1;
      st.largest_batch = std::max<uint32_t>(st.largest_batch, (uint32_t)s.idx.size());
      progress = true;
    }
    if (!progress) std::this_thread::yield();
  }
  if (fail != EO_OK) {
    for (auto &s : slots)
      if (s.busy) {
        s.res.resize(s.idx.size());
        (void)eo_batch_wait(s.b, s.res.data(), (uint32_t)s.idx.size());
      }
    cleanup();
    return fail;
  }
  st.wall_ms = ms_since(t0);
  std::vector<double> ran_lat;
  for (uint32_t i = 0; i < n; ++i) {
    const uint64_t nR = reqs[i].table ? reqs[i].table->n : reqs[i].nR;
    st.tuples += nR + reqs[i].nS;
    if (ran[i]) ran_lat.push_back(lat[i]);
  }
  percentiles(ran_lat, &st);
  if (latency_ms)
    for (uint32_t i = 0; i < n; ++i) latency_ms[i] = lat[i];
  if (stats) *stats = st;
  cleanup();
  return EO_OK;
}
