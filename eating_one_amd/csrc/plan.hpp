// Host-only planning for the hash join: radix bits, passes, LDS table size
// and S tile, per input size.  Pure functions (no HIP calls), so they run and
// are tested without a GPU.
//
// The reference plans at joins/hashJoin_gpu/main_street/hj.cc:378-436:
// p_num = |R| / PART_STANDARD(64) rounded to a power of 64 or of 2, then
// 6 bits per pass.  Here partitions are sized to the LDS table instead, so
// that each R partition is built once in LDS and probed by its S partition.
#pragma once

#include <stdint.h>

#include "eojoin.h"

namespace eo {

struct HashPlan {
  int bits = 0;        // total radix bits (0 = single partition)
  int passes = 0;      // 0 (no partitioning), 1 or 2
  int bits_pass1 = 0;  // LSD: first pass takes the low bits
  int bits_pass2 = 0;
  uint32_t lds_cap = 0;
  uint32_t s_tile = 0;  // S tuples per tile of the deterministic count/write path
  uint32_t s_span = 0;  // S tuples per tile of the fused probe
};

constexpr int kMaxJoinBits = 22;

// R tuples one fused-probe LDS table holds for a planned lds_cap: the full
// table (lds_cap 16384) 17408, the half table (lds_cap <= 8192) 8704.
constexpr uint32_t kFusedTableFull = 17408, kFusedTableHalf = 8704;
inline uint32_t fused_table_tuples(uint32_t lds_cap) { return lds_cap <= 8192 ? kFusedTableHalf : kFusedTableFull; }
// An unpartitioned join (every S tile probes each R chunk of one LDS table)
// when R fills at most this many tables and the S tuples read across them stay
// within kUnpartitionedSReads.
constexpr uint64_t kUnpartitionedChunks = 4;
constexpr uint64_t kUnpartitionedSReads = 1ull << 22;

// Returns EO_OK and fills *out, or EO_EINVAL for inconsistent overrides.
eo_status plan_hash_join(uint64_t nR, uint64_t nS, const eo_join_opts *opts, HashPlan *out, int cus = 256);

// S tuples per fused-probe tile of a join of nR x nS planned with lds_cap and
// bits, when s_pool S tuples (>= nS: the join's own, or a batch's) share the
// GPU: a power of two in [8192, 2^18].
uint32_t plan_span(uint64_t nR, uint64_t nS, uint32_t lds_cap, int bits, uint64_t s_pool, int cus);

}  // namespace eo
