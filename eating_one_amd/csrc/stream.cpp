// Out-of-core streaming join (eo_stream_join): host-resident R and S of any
// size pass through the device, R in tiles that stay resident while every S
// chunk streams past them.  The reference does this for its NLJ
// (joins/blockJoin_gpu/main_street/nestLoopJoin.cc:323-519): for every pair
// of PART-sized tiles a synchronous HtoD of both, count, presum, alloc, join,
// DtoH and a host append.  Here each R tile crosses PCIe once (and is
// partitioned once for the hash join), S chunks are double-buffered, the
// copies in both directions run on their own streams beside the joins on the
// context stream, and each chunk is joined in one fused pass that reserves
// its output with an atomic (count and rows together).
//
// Chunk i, buffers b = i & 1:
//   host   : stage S chunk i+1 into pinned hS[b^1] (pageable S only), queue its
//            H2D, wait for chunk i's row count, queue chunk i's D2H, copy the
//            rows of chunk i-1 out of pinned hOut[b^1] (pageable out only)
//   up     : H2D S chunk i+1 -> dS[b^1]                         (ev_up)
//   compute: waits ev_up[b], partitions chunk i (hash), probes into dOut[b]
//            after the D2H of chunk i-2 left it                  (ev_done)
//   down   : waits ev_done[b], D2H dOut[b] -> out or hOut[b]     (ev_down)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "ctx.hpp"
#include "eojoin.h"
#include "kernels.hpp"
#include "plan.hpp"

namespace eo {
namespace {

constexpr uint64_t kDefaultChunk = 1ull << 24;

// memcpy split over up to 8 host threads: pageable <-> pinned staging runs at
// host memory speed instead of one core's copy rate.
void host_copy(void *dst, const void *src, size_t bytes) {
  constexpr size_t kPerThread = (size_t)16 << 20;
  const size_t nt = std::min<size_t>(8, bytes / kPerThread);
  if (nt <= 1) {
    if (bytes) std::memcpy(dst, src, bytes);
    return;
  }
  const size_t part = (bytes / nt + 4095) & ~(size_t)4095;
  std::vector<std::thread> pool;
  size_t done_by_pool = part;  // the calling thread copies [0, part)
  try {
    for (size_t b = part; b < bytes; b += part) {
      const size_t e = std::min(bytes, b + part);
      pool.emplace_back([=] { std::memcpy((char *)dst + b, (const char *)src + b, e - b); });
      done_by_pool = e;
    }
  } catch (...) {  // no thread available: copy the rest here
  }
  std::memcpy(dst, src, std::min(bytes, part));
  if (done_by_pool < bytes) std::memcpy((char *)dst + done_by_pool, (const char *)src + done_by_pool, bytes - done_by_pool);
  for (auto &t : pool) t.join();
}

struct Events {
  hipEvent_t up[2] = {}, done[2] = {}, down[2] = {};
  hipError_t create() {
    for (int b = 0; b < 2; ++b)
      for (hipEvent_t *e : {&up[b], &done[b], &down[b]}) {
        hipError_t r = hipEventCreateWithFlags(e, hipEventDisableTiming);
        if (r != hipSuccess) return r;
      }
    return hipSuccess;
  }
  ~Events() {
    for (int b = 0; b < 2; ++b)
      for (hipEvent_t e : {up[b], done[b], down[b]})
        if (e) (void)hipEventDestroy(e);
  }
};

struct Pending {
  uint64_t off = 0, rows = 0;  // rows of a chunk parked in hOut[b], destined for out + off
};

struct Streamer {
  eo_ctx *c = nullptr;
  hipStream_t cs = nullptr, up = nullptr, down = nullptr;
  bool hash = true, wide = false, outer = false;
  size_t row = sizeof(eo_pair);
  HashPlan hp;
  uint32_t P = 1;
  // device (arena)
  eo_tuple *dR = nullptr, *rp = nullptr, *tmp = nullptr, *sp = nullptr, *dS[2] = {};
  uint32_t *r_start = nullptr, *s_start = nullptr, *tile_start = nullptr, *matched = nullptr;
  uint64_t *small = nullptr;  // [0] tile count, [1] row cursor
  void *pscr = nullptr, *jscr = nullptr;
  // device output staging (own allocations: they grow inside a call)
  eo_pair *dOut[2] = {};
  uint64_t out_rows = 0;
  // pinned host staging
  char *hS[2] = {};
  PinnedBuf hout_buf;
  char *hOut[2] = {};
  Events ev;
  Pending pend[2];
  bool s_pinned = false, out_pinned = false;
  eo_pair *out = nullptr;
  uint64_t out_cap = 0, written = 0, total = 0;
  eo_stream_stats st{};

  ~Streamer() {
    if (cs) (void)hipStreamSynchronize(cs);
    if (up) (void)hipStreamSynchronize(up);
    if (down) (void)hipStreamSynchronize(down);
    for (eo_pair *&d : dOut)
      if (d) (void)hipFree(d);
    hout_buf.release();
  }

  eo_status grow_out(uint64_t rows) {
    // nothing may be in flight on the buffers being replaced
    EO_TRY_HIP(hipStreamSynchronize(down), "sync");
    EO_TRY_HIP(hipStreamSynchronize(cs), "sync");
    for (eo_pair *&d : dOut)
      if (d) {
        (void)hipFree(d);
        d = nullptr;
      }
    for (eo_pair *&d : dOut) EO_TRY_HIP(hipMalloc((void **)&d, std::max<uint64_t>(rows, 1) * row), "output staging");
    if (!out_pinned && out) {
      const size_t half = (std::max<uint64_t>(rows, 1) * row + 4095) & ~(size_t)4095;
      hout_buf.release();
      EO_TRY_HIP(hout_buf.reserve(2 * half), "pinned output staging");
      hOut[0] = hout_buf.base;
      hOut[1] = hout_buf.base + half;
    }
    out_rows = rows;
    return EO_OK;
  }

  // Copies bytes of host memory to device memory on the up stream, through
  // the S staging slots (starting with `slot`, alternating) when the source
  // is pageable, and records ev.up[slot] after the copy (pinned source) or
  // each staged piece.  Returns with the copies queued.
  eo_status upload(void *dst, const void *src, size_t bytes, bool pinned, int slot, size_t slot_bytes) {
    if (pinned) {
      if (bytes) EO_TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, up), "h2d");
      EO_TRY_HIP(hipEventRecord(ev.up[slot], up), "event");
    } else {
      for (size_t off = 0; off < bytes; off += slot_bytes, slot ^= 1) {
        const size_t n = std::min(slot_bytes, bytes - off);
        EO_TRY_HIP(hipEventSynchronize(ev.up[slot]), "staging wait");  // the slot's last H2D has left it
        host_copy(hS[slot], (const char *)src + off, n);
        EO_TRY_HIP(hipMemcpyAsync((char *)dst + off, hS[slot], n, hipMemcpyHostToDevice, up), "h2d");
        EO_TRY_HIP(hipEventRecord(ev.up[slot], up), "event");
      }
    }
    st.h2d_bytes += bytes;
    return EO_OK;
  }

  // Copies the parked rows of slot b to the caller's buffer.
  eo_status drain(int b) {
    if (!pend[b].rows) return EO_OK;
    EO_TRY_HIP(hipEventSynchronize(ev.down[b]), "d2h wait");
    host_copy((char *)out + pend[b].off * row, hOut[b], pend[b].rows * row);
    pend[b] = Pending{};
    return EO_OK;
  }

  // Joins S chunk (device dS[b], n tuples) against the resident R tile;
  // returns its match count (rows beyond out_rows are counted, not written).
  eo_status probe(int b, uint64_t nr, uint64_t n, bool write, uint64_t *count) {
    eo_pair *dst = write ? dOut[b] : nullptr;
    const uint64_t cap = write ? out_rows : 0;
    EO_TRY_HIP(fill_u64(small + 1, 0, 1, cs), "cursor reset");
    if (outer) EO_TRY_HIP(fill_u32(matched, 0u, matched_words(n), cs), "match bitmap reset");
    const eo_tuple *probe_side = dS[b];
    hipError_t e = hipSuccess;
    if (hash) {
      if (hp.passes) {
        eo_status s = partition_relation(c, dS[b], n, hp, tmp, sp, s_start, pscr, cs);
        if (s != EO_OK) return s;
        probe_side = sp;
      } else {
        EO_TRY_HIP(fill_u32(s_start, 0u, 1, cs), "fill");
        EO_TRY_HIP(fill_u32(s_start + 1, (uint32_t)n, 1, cs), "fill");
      }
      const uint32_t cap_lds = join_fused_cap(hp.lds_cap);
      EO_TRY_HIP(join_fused_plan(P, hp.s_span, cap_lds, r_start, r_start + 1, s_start, s_start + 1, tile_start,
                                 small, jscr, c->d_err, cs),
                 "tile plan");
      // the probe's grid does not need the tile count: no readback before it
      const uint64_t tiles_est = (uint64_t)P + nr / cap_lds + n / hp.s_span + 1024;
      e = join_fused(P, hp.s_span, cap_lds, join_fused_grid(P, nr, n, tiles_est, c->props.cus), small, hp.passes ? rp : dR,
                     r_start, r_start + 1, probe_side, s_start, s_start + 1, tile_start, dst, cap, small + 1, wide,
                     outer ? matched : nullptr, c->d_err, cs);
    } else {
      e = nlj_fused(plan_nlj(nr, n), dR, dS[b], small + 1, dst, cap, wide, outer ? matched : nullptr, cs);
    }
    if (e == hipSuccess && outer) e = emit_unmatched(probe_side, n, matched, small + 1, dst, cap, wide, cs);
    if (e != hipSuccess) return hip_fail(e, hash ? "stream hash probe" : "stream nlj");
    EO_TRY_HIP(hipEventRecord(ev.done[b], cs), "event");
    return read_u64(c, small + 1, count, 1, cs);
  }

  // Streams every S chunk past the resident R tile of nr tuples.
  eo_status run_tile(const eo_tuple *S, uint64_t nS, uint64_t sc, uint64_t nr) {
    const uint64_t chunks = (nS + sc - 1) / sc;
    auto chunk_len = [&](uint64_t i) { return std::min(sc, nS - i * sc); };
    auto stage = [&](uint64_t i) {  // a chunk is one staged piece: slot b only
      const int b = (int)(i & 1);
      return upload(dS[b], S + i * sc, chunk_len(i) * sizeof(eo_tuple), s_pinned, b, sc * sizeof(eo_tuple));
    };
    eo_status s = stage(0);
    if (s != EO_OK) return s;
    for (uint64_t i = 0; i < chunks; ++i) {
      const int b = (int)(i & 1);
      const uint64_t n = chunk_len(i);
      EO_TRY_HIP(hipStreamWaitEvent(cs, ev.up[b], 0), "wait");    // chunk i is in dS[b]
      EO_TRY_HIP(hipStreamWaitEvent(cs, ev.down[b], 0), "wait");  // dOut[b] left by chunk i-2's D2H
      if (i + 1 < chunks && (s = stage(i + 1)) != EO_OK) return s;
      const bool write = out && written < out_cap;
      uint64_t cnt = 0;
      if ((s = probe(b, nr, n, write, &cnt)) != EO_OK) return s;
      const uint64_t want = write ? std::min(cnt, out_cap - written) : 0;
      if (want > out_rows) {  // rows outgrew the staging: grow it and join this chunk again
        if ((s = drain(b ^ 1)) != EO_OK) return s;
        if ((s = grow_out(std::max(want, 2 * out_rows))) != EO_OK) return s;
        if ((s = probe(b, nr, n, true, &cnt)) != EO_OK) return s;
        ++st.regrown;
      }
      if (want) {
        EO_TRY_HIP(hipStreamWaitEvent(down, ev.done[b], 0), "wait");
        if (out_pinned) {
          EO_TRY_HIP(hipMemcpyAsync((char *)out + written * row, dOut[b], want * row, hipMemcpyDeviceToHost, down),
                     "d2h");
        } else {
          EO_TRY_HIP(hipMemcpyAsync(hOut[b], dOut[b], want * row, hipMemcpyDeviceToHost, down), "d2h");
          pend[b] = Pending{written, want};
        }
        EO_TRY_HIP(hipEventRecord(ev.down[b], down), "event");
        st.d2h_bytes += want * row;
      }
      written += want;
      total += cnt;
      ++st.s_chunks;
      if ((s = drain(b ^ 1)) != EO_OK) return s;
    }
    for (int b = 0; b < 2; ++b)
      if ((s = drain(b)) != EO_OK) return s;
    EO_TRY_HIP(hipStreamSynchronize(down), "sync");
    return EO_OK;
  }
};

bool stream_opts_ok(const eo_stream_opts *o) {
  if (!o) return true;
  if (o->struct_size != 0 && o->struct_size < sizeof(eo_stream_opts)) return false;
  if (o->op != EO_STREAM_HASH && o->op != EO_STREAM_NLJ) return false;
  if (o->r_tile > kMaxRows || o->s_chunk > kMaxRows || o->out_chunk > ((uint64_t)1 << 40)) return false;
  if (o->flags & ~(EO_F_RESULT16 | EO_F_LEFT_OUTER)) return false;
  return !o->join || opts_ok(o->join);
}

}  // namespace
}  // namespace eo

using namespace eo;

extern "C" eo_status eo_stream_join(eo_ctx *c, const eo_tuple *R, uint64_t nR, const eo_tuple *S, uint64_t nS,
                                    eo_pair *out, uint64_t out_cap, uint64_t *n_out, const eo_stream_opts *o) {
  if (!c || !n_out || !stream_opts_ok(o)) return EO_EINVAL;
  if ((nR && !R) || (nS && !S) || (out_cap && !out)) return EO_EINVAL;
  const auto t0 = std::chrono::steady_clock::now();
  *n_out = 0;
  eo_stream_stats *stats = o ? o->stats : nullptr;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  const bool outer = o && (o->flags & EO_F_LEFT_OUTER);
  const uint64_t rt = std::max<uint64_t>(1, std::min<uint64_t>(o && o->r_tile ? o->r_tile : kMaxRows, nR));
  const uint64_t sc = std::max<uint64_t>(1, std::min<uint64_t>(o && o->s_chunk ? o->s_chunk : kDefaultChunk, nS));
  const uint64_t r_tiles = nR ? (nR + rt - 1) / rt : (outer ? 1 : 0);
  if (outer && r_tiles > 1) {
    set_last_error("EO_F_LEFT_OUTER needs R in one tile (r_tile >= nR)");
    return EO_EINVAL;
  }
  if (nS == 0 || r_tiles == 0) return EO_OK;

  DeviceGuard g(c->device);
  Streamer x;
  x.c = c;
  x.hash = !o || o->op == EO_STREAM_HASH;
  x.wide = o && (o->flags & EO_F_RESULT16);
  x.outer = outer;
  x.row = x.wide ? 16 : sizeof(eo_pair);
  x.out = out;
  x.out_cap = out ? out_cap : 0;
  x.s_pinned = is_pinned(c, S, nS * sizeof(eo_tuple));
  x.out_pinned = out && is_pinned(c, out, out_cap * x.row);
  const bool r_pinned = is_pinned(c, R, nR * sizeof(eo_tuple));
  x.st.pinned = (r_pinned ? 1u : 0u) | (x.s_pinned ? 2u : 0u) | (x.out_pinned ? 4u : 0u);
  x.cs = c->stream;
  for (hipStream_t &s : c->xfer)
    if (!s) EO_TRY_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "copy stream");
  x.up = c->xfer[0];
  x.down = c->xfer[1];
  EO_TRY_HIP(x.ev.create(), "events");

  // one plan for every R tile (sized for the largest) and every S chunk
  const uint64_t r_last = nR ? nR - (r_tiles - 1) * rt : 0, s_last = nS - ((nS - 1) / sc) * sc;
  if (x.hash) {
    eo_join_opts jo{};
    if (o && o->join) jo = *o->join;
    jo.where = EO_DEVICE;
    jo.flags = 0;
    eo_status s = plan_hash_join(std::min(rt, nR), sc, &jo, &x.hp, c->props.cus);
    if (s != EO_OK) return s;
    x.P = 1u << x.hp.bits;
  }
  const int cus = c->props.cus;
  Layout L;
  const size_t o_rin = L.add(std::min(rt, nR) * sizeof(eo_tuple));
  const bool parted = x.hash && x.hp.passes;
  const size_t o_rp = L.add(parted ? std::min(rt, nR) * sizeof(eo_tuple) : 0);
  const size_t o_tmp = L.add(x.hash && x.hp.passes == 2 ? std::max(std::min(rt, nR), sc) * sizeof(eo_tuple) : 0);
  const size_t o_rs = L.add((x.P + 1) * sizeof(uint32_t));
  const size_t o_s0 = L.add(sc * sizeof(eo_tuple));
  const size_t o_s1 = L.add(sc * sizeof(eo_tuple));
  const size_t o_sp = L.add(parted ? sc * sizeof(eo_tuple) : 0);
  const size_t o_ss = L.add((x.P + 1) * sizeof(uint32_t));
  const size_t o_ts = L.add((x.P + 1) * sizeof(uint32_t));
  const size_t o_small = L.add(8 * sizeof(uint64_t));
  const size_t o_bits = L.add(outer ? matched_words(sc) * sizeof(uint32_t) : 0);
  size_t pscr = 0;
  if (parted)
    for (uint64_t n : {std::min(rt, nR), r_last, sc, s_last}) pscr = std::max(pscr, partition_scratch(n, x.hp, cus));
  const size_t o_pscr = L.add(pscr);
  const size_t o_jscr = L.add(x.hash ? join_plan_scratch_bytes(x.P) : 0);
  EO_TRY_HIP(hipStreamSynchronize(x.cs), "sync");
  EO_TRY_HIP(c->arena.reserve(L.total), "arena");
  char *A = c->arena.base;
  x.dR = (eo_tuple *)(A + o_rin);
  x.rp = (eo_tuple *)(A + o_rp);
  x.tmp = (eo_tuple *)(A + o_tmp);
  x.r_start = (uint32_t *)(A + o_rs);
  x.dS[0] = (eo_tuple *)(A + o_s0);
  x.dS[1] = (eo_tuple *)(A + o_s1);
  x.sp = (eo_tuple *)(A + o_sp);
  x.s_start = (uint32_t *)(A + o_ss);
  x.tile_start = (uint32_t *)(A + o_ts);
  x.small = (uint64_t *)(A + o_small);
  x.matched = (uint32_t *)(A + o_bits);
  x.pscr = A + o_pscr;
  x.jscr = A + o_jscr;
  if (!x.s_pinned || !r_pinned) {
    EO_TRY_HIP(c->stage.reserve(2 * sc * sizeof(eo_tuple)), "pinned staging");
    x.hS[0] = c->stage.base;
    x.hS[1] = c->stage.base + sc * sizeof(eo_tuple);
  }
  for (int b = 0; b < 2; ++b) {  // every event starts out complete
    EO_TRY_HIP(hipEventRecord(x.ev.up[b], x.up), "event");
    EO_TRY_HIP(hipEventRecord(x.ev.down[b], x.down), "event");
    EO_TRY_HIP(hipEventRecord(x.ev.done[b], x.cs), "event");
  }
  if (out) {
    const uint64_t rows = o && o->out_chunk ? o->out_chunk : sc;
    eo_status s = x.grow_out(std::min(rows, out_cap));
    if (s != EO_OK) return s;
  }

  for (uint64_t t = 0; t < r_tiles; ++t) {
    const uint64_t r0 = t * rt, nr = nR ? std::min(rt, nR - r0) : 0;
    // the previous tile's joins are done (their counts were read), so dR is free
    eo_status s = x.upload(x.dR, R + r0, nr * sizeof(eo_tuple), r_pinned, 0, sc * sizeof(eo_tuple));
    if (s != EO_OK) return s;
    EO_TRY_HIP(hipEventRecord(x.ev.up[0], x.up), "event");  // after every piece of the tile
    EO_TRY_HIP(hipStreamWaitEvent(x.cs, x.ev.up[0], 0), "wait");
    if (x.hash) {
      if (parted) {
        s = partition_relation(c, x.dR, nr, x.hp, x.tmp, x.rp, x.r_start, x.pscr, x.cs);
        if (s != EO_OK) return s;
      } else {
        EO_TRY_HIP(fill_u32(x.r_start, 0u, 1, x.cs), "fill");
        EO_TRY_HIP(fill_u32(x.r_start + 1, (uint32_t)nr, 1, x.cs), "fill");
      }
    }
    // S staging slot 0 may still feed the R upload: the first chunk waits for it
    EO_TRY_HIP(hipStreamSynchronize(x.up), "sync");
    s = x.run_tile(S, nS, sc, nr);
    if (s != EO_OK) return s;
    ++x.st.r_tiles;
  }
  EO_TRY_HIP(hipStreamSynchronize(x.cs), "sync");
  eo_status s = check_device_errors(c, x.cs);
  if (s != EO_OK) return s;
  *n_out = x.total;
  if (stats) {
    *stats = x.st;
    stats->n_out = x.total;
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return (out && x.total > out_cap) ? EO_ECAPACITY : EO_OK;
}
