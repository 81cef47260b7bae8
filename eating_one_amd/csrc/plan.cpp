#include "plan.hpp"

#include <algorithm>

#include "kernels.hpp"

namespace eo {

namespace {
int ceil_log2(uint64_t x) {
  int b = 0;
  while ((1ull << b) < x) ++b;
  return b;
}
bool valid_cap(int c) { return c == 4096 || c == 8192 || c == 16384; }
}  // namespace

eo_status plan_hash_join(uint64_t nR, uint64_t nS, const eo_join_opts *o, HashPlan *out, int cus) {
  HashPlan p;
  const int cap_opt = o ? o->lds_tuples : 0;
  if (cap_opt != 0 && !valid_cap(cap_opt)) return EO_EINVAL;
  p.lds_cap = cap_opt ? (uint32_t)cap_opt : 16384u;

  const int bits_opt = o ? o->radix_bits : 0;
  if (bits_opt < -1 || bits_opt > kMaxJoinBits) return EO_EINVAL;
  // A join that needs two partition passes anyway takes half tables: the
  // extra radix bit lands in the second pass (which moves the same bytes),
  // and two probe workgroups per CU hide each other's barrier and atomic
  // round trips (C2, 2^27 x 2^28: probe 1.32 -> 1.08 ms, join 5.09 -> 4.93 ms;
  // deterministic order 7.35 -> 6.86 ms).
  if (!cap_opt && !bits_opt && nR > 16384 && ceil_log2((nR + 16383) / 16384) > kMaxPassBits) p.lds_cap = 8192u;
  if (bits_opt == -1) {
    p.bits = 0;
  } else if (bits_opt > 0) {
    p.bits = bits_opt;
  } else if (nR <= p.lds_cap) {
    p.bits = 0;
  } else if (!cap_opt && (nR + fused_table_tuples(p.lds_cap) - 1) / fused_table_tuples(p.lds_cap) <=
                             kUnpartitionedChunks &&
             (nR + fused_table_tuples(p.lds_cap) - 1) / fused_table_tuples(p.lds_cap) * nS <=
                 kUnpartitionedSReads) {
    // A build side of a few LDS tables beside a small S: no partitioning;
    // every S tile streams past each R chunk (S read up to four times), which
    // costs less than partitioning both sides (measured: 2^16 x 2^20 0.071 ms
    // against 0.089 partitioned, 2^15 x 2^20 0.064 against 0.096; 2^17 x
    // 2^20, 8 chunks, and 2^16 x 2^22, 2^24 S reads, are faster partitioned)
    p.bits = 0;
  } else {
    // Mean R partition at most lds_cap tuples, so that each partition is one
    // LDS table (the table holds 17408: a full 16K partition plus binomial
    // spread).  Up to 2^11 digits in one pass (each digit's run per 16K-tuple
    // staging tile stays near 64 bytes, so writes coalesce); beyond that two
    // passes, high digit first.
    p.bits = ceil_log2((nR + p.lds_cap - 1) / p.lds_cap);
    if (p.bits > kMaxJoinBits) p.bits = kMaxJoinBits;
  }

  const int passes_opt = o ? o->passes : 0;
  if (passes_opt < 0 || passes_opt > 2) return EO_EINVAL;
  if (p.bits == 0) {
    p.passes = 0;
  } else if (passes_opt) {
    p.passes = passes_opt;
  } else {
    p.passes = p.bits <= kMaxPassBits ? 1 : 2;
  }
  if (p.passes == 1) {
    if (p.bits > kMaxPassBits) return EO_EINVAL;
    p.bits_pass1 = p.bits;
  } else if (p.passes == 2) {
    if (p.bits < 2 || p.bits > 2 * kMaxPassBits) return EO_EINVAL;
    p.bits_pass1 = p.bits / 2;
    p.bits_pass2 = p.bits - p.bits_pass1;
  }

  const int st_opt = o ? o->s_tile : 0;
  if (st_opt < 0) return EO_EINVAL;
  p.s_tile = st_opt ? (uint32_t)st_opt : 16384u;
  p.s_span = st_opt ? std::max<uint32_t>(8192, (uint32_t)st_opt) : plan_span(nR, nS, p.lds_cap, p.bits, nS, cus);
  *out = p;
  return EO_OK;
}

uint32_t plan_span(uint64_t nR, uint64_t nS, uint32_t lds_cap, int bits, uint64_t s_pool, int cus) {
  // Fused probe: S spans short enough to give every CU a few tiles of the
  // s_pool S tuples that share the GPU (this join's, or a whole batch's), and
  // at most twice the mean S partition, so that a partition grown by skewed
  // keys splits into tiles no longer than the others (a long tile late in the
  // grid would hold the whole join up); long enough to amortise the LDS build.
  const uint64_t parts = 1ull << bits;
  const uint64_t want_tiles = (uint64_t)std::max(cus, 8) * 4;
  const uint64_t chunks = (nR + lds_cap - 1) / std::max<uint32_t>(lds_cap, 1);
  uint64_t span = std::max(s_pool, nS) * chunks / want_tiles;
  span = std::min<uint64_t>(span, 2 * ((nS + parts - 1) / parts));
  uint32_t s2 = 8192;
  while (s2 < span && s2 < (1u << 18)) s2 <<= 1;
  return s2;
}

}  // namespace eo
