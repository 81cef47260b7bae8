// Autotuner for the hash join plan.
//
// The reference fixes its plan with compile-time tables (PART_STANDARD,
// per-size `#define`s; joins/hashJoin_gpu/main_street/hj.cc:378-436 derives
// p_num from |R| / 64 and 6 bits per pass).  Here the planner's choice
// (plan.cpp: partitions sized to one LDS table) is a starting point: for one
// size class (ceil log2 |R|, ceil log2 |S|) eo_tune_hash_join times every
// candidate on the device — the LDS table size (a full 17408-tuple table,
// one 1024-thread workgroup per CU, or a half table with two workgroups per
// CU), radix bits around the planner's for that table, one or two passes,
// three S spans per probe tile — and keeps the fastest in the context's
// table.  Joins and batches that leave radix_bits, passes,
// lds_tuples and s_tile at 0 then use the table's entry for their class.
// Every candidate computes the same result (tested), so tuning only changes
// speed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "eojoin.h"
#include "plan.hpp"

namespace eo {

namespace {
int ceil_log2(uint64_t x) {
  int b = 0;
  while ((1ull << b) < x) ++b;
  return b;
}
}  // namespace

}  // namespace eo

using namespace eo;

extern "C" {

eo_status eo_tune_hash_join(eo_ctx *c, const eo_tuple *R, uint64_t nR, const eo_tuple *S, uint64_t nS, eo_pair *out,
                            uint64_t out_cap, int reps, eo_tune_result *res) {
  if (!c || !res || nR == 0 || nS == 0 || !R || !S || (out_cap && !out) || nR > kMaxRows || nS > kMaxRows)
    return EO_EINVAL;
  reps = std::max(1, std::min(reps, 16));
  std::memset(res, 0, sizeof(*res));
  eo_status st = EO_OK;
  // candidates per LDS table (16384: the full table, 8192: the half table):
  // bits b0-1 .. b0+2 around the planner's for that table, passes 1
  // (bits <= 11) or 2, span auto / x2 / x0.5
  struct Cand {
    int bits, passes, lds, s_tile;
  };
  std::vector<Cand> cands;
  for (int lds : {16384, 8192}) {
    eo_join_opts lo{};
    lo.lds_tuples = lds;
    HashPlan base;
    if (plan_hash_join(nR, nS, &lo, &base, c->props.cus) != EO_OK) continue;
    for (int b = std::max(0, base.bits - 1); b <= std::min(kMaxJoinBits, base.bits + 2); ++b) {
      for (int p = 1; p <= 2; ++p) {
        if (b == 0 && p == 2) continue;
        if (b > 0 && p == 1 && b > kMaxPassBits) continue;
        if (p == 2 && (b < 2 || b > 2 * kMaxPassBits)) continue;
        HashPlan hp;
        eo_join_opts o{};
        o.radix_bits = b ? b : -1;
        o.passes = b ? p : 0;
        o.lds_tuples = lds;
        if (plan_hash_join(nR, nS, &o, &hp, c->props.cus) != EO_OK) continue;
        for (int k = 0; k < 3; ++k) {
          uint32_t span = hp.s_span;
          if (k == 1) span = std::min<uint32_t>(span * 2, 1u << 18);
          if (k == 2) span = std::max<uint32_t>(span / 2, 8192u);
          if (k > 0 && span == hp.s_span) continue;
          cands.push_back({b, b ? p : 0, lds, k == 0 ? 0 : (int)span});
        }
        if (b == 0) break;
      }
    }
  }
  c->tuned.erase(tune_key(nR, nS));  // the baseline below is the untuned planner
  eo_join_stats stats;
  auto time_plan = [&](const eo_join_opts &o, double *best) -> eo_status {
    *best = 1e30;
    for (int r = 0; r <= reps; ++r) {  // first run warms the arena
      uint64_t n = 0;
      eo_status s2 = eo_hash_join(c, R, nR, S, nS, out, out_cap, &n, &o);
      if (s2 != EO_OK && s2 != EO_ECAPACITY) return s2;
      if (r > 0) *best = std::min(*best, stats.ms_total);
    }
    return EO_OK;
  };
  eo_join_opts o{};
  o.struct_size = sizeof(o);
  o.where = EO_DEVICE;
  o.stats = &stats;
  // the untuned planner
  double def_ms = 0;
  st = time_plan(o, &def_ms);
  if (st != EO_OK) return st;
  TunedPlan best{0, 0, 0, 0, 1e30};
  for (const Cand &cd : cands) {
    o.radix_bits = cd.bits ? cd.bits : -1;
    o.passes = cd.passes;
    o.lds_tuples = cd.lds;
    o.s_tile = cd.s_tile;
    double ms = 0;
    st = time_plan(o, &ms);
    if (st != EO_OK) return st;
    if (ms < best.ms) best = {cd.bits, cd.passes, cd.lds, cd.s_tile, ms};
  }
  c->tuned[tune_key(nR, nS)] = best;
  res->radix_bits = best.radix_bits;
  res->passes = best.passes;
  res->lds_tuples = best.lds_tuples;
  res->s_tile = best.s_tile;
  res->ms = best.ms;
  res->default_ms = def_ms;
  res->candidates = (uint32_t)cands.size();
  return EO_OK;
}

}  // extern "C"
