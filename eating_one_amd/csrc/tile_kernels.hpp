// Device bodies shared by the single-join kernels and the batched persistent
// kernels: the fused probe of one (R chunk, S span) tile, the fast radix
// scatter of one stripe and the LDS histogram of one stripe.  Each works on a
// caller-declared __shared__ struct so that a kernel can run several tiles or
// stripes in a loop.
#pragma once

#include "device.hpp"

namespace eo {
namespace tile {

// ---- fused probe tile ----------------------------------------------------------
// ProbeCfg<CAP, LOG2NB, SPT>: an LDS table of up to CAP R tuples in 2^LOG2NB
// hash buckets, built and probed by a 1024-thread workgroup (SPT S tuples per
// thread per batch).  ProbeBig (17408 tuples, ~156 KB) runs one workgroup per
// CU with up to 128 VGPRs; ProbeSmall (8704 tuples, ~78 KB) lets two share a
// CU, which leaves 64 VGPRs: 4 S tuples per thread keep its build (9 R tuples
// per thread plus the first S batch in flight) within them.
template <int CAP, int LOG2NB, int S_PER_THREAD>
struct ProbeCfg {
  static constexpr int NT = 1024;
  static constexpr int Cap = CAP;
  static constexpr int Log2NB = LOG2NB;
  static constexpr int NB = 1 << LOG2NB;
  static constexpr int SPT = S_PER_THREAD;
  static constexpr int Batch = NT * SPT;
  static constexpr int RPT = (CAP + NT - 1) / NT;
  static constexpr int BPT = NB / NT;
};
using ProbeBig = ProbeCfg<17408, 13, 8>;   // a 16K-tuple partition plus binomial spread
using ProbeSmall = ProbeCfg<8704, 12, 4>;  // an 8K-tuple partition plus spread
constexpr int kFNT = 1024;
constexpr int kFCap = ProbeBig::Cap;

struct SpanRef {
  uint32_t r_beg, r_n, s_beg, s_n;
};

template <typename C>
struct ProbeShared {
  uint32_t keys[C::Cap];
  uint32_t pays[C::Cap];
  uint16_t off[C::NB + 2];
  uint32_t scan_tmp[C::NT / kWave + 1];
  // per S batch, by batch parity: match totals of each (tuple slot, wave),
  // then their exclusive prefix; the batch's output base
  alignas(8) uint32_t wtot[2][C::SPT * (C::NT / kWave)];  // also the u64 scratch of a count's sum
  uint64_t base_sh[2];
};

// What probe_tile does with the matches of a tile.
constexpr int kProbeFused = 0;    // write rows at ranges reserved from *cursor, one atomic per batch
constexpr int kProbeCount = 1;    // count only: return the tile's match count
constexpr int kProbeOrdered = 2;  // write rows from position base on, in an order fixed by the inputs

// Builds the LDS table of R[ref.r_beg, +r_n) (r_n <= C::Cap) and probes the S
// span S[ref.s_beg, +s_n) against it.  kProbeFused writes (R payload, S
// payload) pairs at positions reserved from *cursor; kProbeOrdered writes them
// from `base` on: the table's bucket order is fixed (lanes sharing a bucket
// ranked by lane, waves taking turns) and batches take consecutive ranges, so
// every run writes the same rows at the same positions; kProbeCount writes
// nothing and returns the number of matches.  Positions >= out_cap are
// counted only.  wide: 16-byte RESULT rows instead of pairs (see emit_pair).
// matched (optional bitmap over S positions): sets the bit of every S tuple
// that matched, for a left outer join.  swapped: the table side is the
// caller's S and the streamed side its R (an inner join built on its smaller
// relation), so every row is written with the two payloads exchanged.
// All threads of the workgroup call it; it leaves LDS free for reuse.
template <typename C, int MODE = kProbeFused>
__device__ __forceinline__ uint64_t probe_tile(ProbeShared<C> &sh, const uint2 *__restrict__ R,
                                               const uint2 *__restrict__ S, const SpanRef &ref,
                                               uint2 *__restrict__ out, uint64_t out_cap, uint64_t *cursor,
                                               bool wide = false, uint32_t *matched = nullptr, uint64_t base = 0,
                                               bool swapped = false) {
  constexpr int NW = C::NT / kWave;
  auto load_batch = [&](uint32_t b0, uint2 *dst) {
#pragma unroll
    for (int u = 0; u < C::SPT; ++u) {
      const uint32_t i = b0 + u * C::NT + threadIdx.x;
      dst[u] = i < ref.s_n ? S[ref.s_beg + i] : make_uint2(0u, 0u);
    }
  };
  uint2 cur[C::SPT];
  load_batch(0, cur);

  // ---- build: rank by LDS atomics on counters aliased onto sh.keys[] ----
  uint32_t *cnt = sh.keys;  // C::NB u32 counters, dead once the offsets exist
  for (int h = threadIdx.x; h < C::NB; h += C::NT) cnt[h] = 0;
  __syncthreads();
  uint2 r[C::RPT];
  uint32_t hh[C::RPT], rk[C::RPT];
#pragma unroll
  for (int k = 0; k < C::RPT; ++k) {
    const uint32_t i = k * C::NT + threadIdx.x;
    if (i < ref.r_n) r[k] = R[ref.r_beg + i];
  }
  if (MODE != kProbeOrdered) {
#pragma unroll
    for (int k = 0; k < C::RPT; ++k) {
      const uint32_t i = k * C::NT + threadIdx.x;
      if (i < ref.r_n) {
        hh[k] = bucket_hash((int32_t)r[k].x, C::Log2NB);
        rk[k] = atomicAdd(&cnt[hh[k]], 1u);
      }
    }
  } else {
    // ranks fixed by the input: within a wave, lanes whose tuples share a
    // bucket are ranked by lane and the group's first lane takes the group's
    // ranks with one atomic; the waves take turns
    // (packed per tuple in rk until the turn: rank in group | leader << 8 | size << 16)
    const int wv = threadIdx.x / kWave;
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int k = 0; k < C::RPT; ++k) {
      const bool valid = k * C::NT + threadIdx.x < ref.r_n;
      hh[k] = valid ? bucket_hash((int32_t)r[k].x, C::Log2NB) : 0u;
      const uint64_t peers = match_peers<C::Log2NB>(hh[k], __ballot(valid));
      const uint32_t lead = valid ? (uint32_t)(__ffsll((unsigned long long)peers) - 1) : (uint32_t)lane_id();
      rk[k] = (uint32_t)__popcll(peers & lt) | lead << 8 | (uint32_t)__popcll(peers) << 16;
    }
    for (int turn = 0; turn < C::NT / kWave; ++turn) {
      if (wv == turn) {
#pragma unroll
        for (int k = 0; k < C::RPT; ++k) {
          const bool valid = k * C::NT + threadIdx.x < ref.r_n;
          const uint32_t before = rk[k] & 0xffu;
          uint32_t first = 0;
          if (valid && before == 0) first = atomicAdd(&cnt[hh[k]], rk[k] >> 16);
          first = __shfl(first, (int)((rk[k] >> 8) & 0xffu));
          rk[k] = first + before;
        }
      }
      __syncthreads();
    }
  }
  __syncthreads();
  {
    uint32_t loc[C::BPT];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < C::BPT; ++k) {
      loc[k] = cnt[threadIdx.x * C::BPT + k];
      sum += loc[k];
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan<C::NT, uint32_t>(sum, sh.scan_tmp, &tot);
#pragma unroll
    for (int k = 0; k < C::BPT; ++k) {
      sh.off[threadIdx.x * C::BPT + k] = (uint16_t)ex;
      ex += loc[k];
    }
    if (threadIdx.x == 0) sh.off[C::NB] = (uint16_t)ref.r_n;  // r_n <= Cap fits u16
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < C::RPT; ++k) {
    const uint32_t i = k * C::NT + threadIdx.x;
    if (i < ref.r_n) {
      const uint32_t pos = sh.off[hh[k]] + rk[k];
      sh.keys[pos] = r[k].x;
      sh.pays[pos] = r[k].y;
    }
  }
  __syncthreads();

  // ---- probe the S span batch by batch ----
  const int w = threadIdx.x / kWave, lane = lane_id();
  constexpr int NV = C::SPT * NW;             // match totals per batch
  constexpr int VPL = (NV + kWave - 1) / kWave;  // of them per lane of wave 0
  uint32_t par = 0;
  uint64_t mine = 0;        // kProbeCount: this thread's matches
  uint64_t running = base;  // kProbeOrdered, wave 0: the next batch's first position
  for (uint32_t b0 = 0; b0 < ref.s_n; b0 += C::Batch, par ^= 1u) {
    uint2 nxt[C::SPT];
    if (b0 + C::Batch < ref.s_n) load_batch(b0 + C::Batch, nxt);
    uint32_t m[C::SPT], fp[C::SPT];
#pragma unroll
    for (int u = 0; u < C::SPT; ++u) {
      m[u] = 0;
      fp[u] = 0;
      if (b0 + u * C::NT + threadIdx.x < ref.s_n) {
        const uint32_t h = bucket_hash((int32_t)cur[u].x, C::Log2NB);
        const uint32_t e1 = sh.off[h + 1];
        for (uint32_t e = sh.off[h]; e < e1; ++e) {
          if (sh.keys[e] == cur[u].x) {
            if (MODE != kProbeCount && m[u] == 0) fp[u] = sh.pays[e];
            ++m[u];
          }
        }
      }
    }
    if (matched) {
#pragma unroll
      for (int u = 0; u < C::SPT; ++u)
        mark_bits(matched, (uint64_t)ref.s_beg + b0 + (uint64_t)u * C::NT + (uint64_t)w * kWave, __ballot(m[u] != 0));
    }
    if (MODE == kProbeCount) {
#pragma unroll
      for (int u = 0; u < C::SPT; ++u) {
        mine += m[u];
        cur[u] = nxt[u];
      }
      continue;
    }
    uint32_t wex[C::SPT];
#pragma unroll
    for (int u = 0; u < C::SPT; ++u) {
      const uint32_t inc = wave_incl_scan(m[u]);
      wex[u] = inc - m[u];
      if (lane == kWave - 1) sh.wtot[par][u * NW + w] = inc;
    }
    __syncthreads();
    // wave 0 alone turns the NV totals into exclusive offsets (VPL consecutive
    // ones per lane) and reserves the batch's output; one barrier publishes it
    if (w == 0) {
      uint32_t a[VPL], sum = 0;
#pragma unroll
      for (int k = 0; k < VPL; ++k) {
        const int i = lane * VPL + k;
        a[k] = i < NV ? sh.wtot[par][i] : 0u;
        sum += a[k];
      }
      const uint32_t inc = wave_incl_scan(sum);
      uint32_t ex = inc - sum;
#pragma unroll
      for (int k = 0; k < VPL; ++k) {
        const int i = lane * VPL + k;
        if (i < NV) sh.wtot[par][i] = ex;
        ex += a[k];
      }
      const uint32_t batch_total = read_lane(inc, kWave - 1);
      if (MODE == kProbeOrdered) {
        if (lane == 0) sh.base_sh[par] = running;
        running += batch_total;
      } else if (lane == 0) {
        sh.base_sh[par] = batch_total ? atomic_add_agent(cursor, (uint64_t)batch_total) : 0ull;
      }
    }
    __syncthreads();
    // no barrier after the writes: the next batch uses the other parity's
    // slots, and its first barrier orders everything after this batch's reads
    const uint64_t base = sh.base_sh[par];
#pragma unroll
    for (int u = 0; u < C::SPT; ++u) {
      if (m[u] == 0) continue;
      uint64_t pos = base + sh.wtot[par][u * NW + w] + wex[u];
      if (m[u] == 1) {
        if (pos < out_cap) {
          if (swapped) emit_pair<true>(out, pos, cur[u].y, fp[u], cur[u].x, wide);
          else emit_pair<true>(out, pos, fp[u], cur[u].y, cur[u].x, wide);
        }
      } else {
        const uint32_t h = bucket_hash((int32_t)cur[u].x, C::Log2NB);
        const uint32_t e1 = sh.off[h + 1];
        for (uint32_t e = sh.off[h]; e < e1; ++e) {
          if (sh.keys[e] == cur[u].x) {
            if (pos < out_cap) {
              if (swapped) emit_pair<true>(out, pos, cur[u].y, sh.pays[e], cur[u].x, wide);
              else emit_pair<true>(out, pos, sh.pays[e], cur[u].y, cur[u].x, wide);
            }
            ++pos;
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < C::SPT; ++u) cur[u] = nxt[u];
  }
  __syncthreads();
  if (MODE == kProbeCount) return block_sum<C::NT, uint64_t>(mine, reinterpret_cast<uint64_t *>(sh.wtot[0]));
  return 0;
}

// ---- fast radix scatter of one stripe ---------------------------------------------
// Unstable scatter: a workgroup of 1024 threads stages 16K tuples per tile;
// ranks come from LDS atomics (the histogram and the within-tile ranks in one
// pass), the tile is placed digit-sorted in LDS and flushed in digit runs; the
// next tile's loads are issued before the flush so reads overlap writes.
constexpr int kFastNT = 1024;
constexpr int kFastPairs = 8;                         // uint4 loads (2 tuples) per thread per tile
constexpr int kFastTile = kFastNT * kFastPairs * 2;   // 16384 tuples: a 128 KB staging tile
// gfx9 buffer resource word 3 for untyped loads (32-bit data format, no swizzle)
constexpr int kBufferRsrcWord3 = 0x00020000;
constexpr int kFastWgPerCu = 1;                       // one workgroup per CU (LDS)

__device__ __forceinline__ uint32_t digit_of(uint32_t key, uint32_t xm, int shift, uint32_t mask) {
  return ((key ^ xm) >> shift) & mask;
}

// ORDERED: ranks fixed by the input (per-wave digit counters, no LDS
// atomics), so every run writes every tuple at the same position; up to 7
// bits (the counters take 16 waves x 2^BITS words).
template <int BITS, bool ORDERED>
struct ScatterWaveCounts {  // per-wave digit counts, then each wave's offset within its digit
  uint32_t wc[kFastNT / kWave][1 << BITS];
};
template <int BITS>
struct ScatterWaveCounts<BITS, false> {};

template <int BITS, bool ORDERED = false>
struct ScatterShared : ScatterWaveCounts<BITS, ORDERED> {
  static_assert(!ORDERED || BITS <= 7, "ordered scatter: per-wave counters fit up to 7 bits");
  uint2 stage[kFastTile];
  uint32_t cnt[1 << BITS];     // this tile's digit counts (LDS atomics)
  uint32_t tbase[1 << BITS];   // this tile's digit starts in stage
  uint32_t cursor[1 << BITS];  // next output position of each digit
  uint32_t pbase[1 << BITS];   // this tile: output position of stage slot i (digit d) = pbase[d] + i (mod 2^32)
};  // 160 KB at 11 bits: the whole LDS of a gfx950 workgroup

// Set by wave 0 of a reserving scatter when a run of the tile passes its
// region's end (an LDS word only the reserving kernels use).
__device__ __forceinline__ uint32_t &scatter_dirty_flag() {
  __shared__ uint32_t flag;
  return flag;
}

// Scatters in[sbeg, send) (in + sbeg 16-byte aligned, or sbeg > 0) by digit
// into out, digit d starting at sh.cursor[d] (set by the caller, followed by a
// barrier).  Positions >= nbound are never written (kErrOverflow is raised).
// RES (reserving mode): every tile takes the base of its run of digit d from
// the global cursor res[d] with one atomic add, and sh.cursor[d] holds the end
// of digit d's region instead: a run that would pass it is not written past
// it and raises kErrCapacity (the regions were sized from an estimate).  The
// atomics are issued as soon as the tile's counts are known and their results
// are needed only after staging, so their round trip overlaps it.
template <int BITS, bool ORDERED = false, bool RES = false>
__device__ __forceinline__ void scatter_stripe(ScatterShared<BITS, ORDERED> &sh, const uint2 *__restrict__ in, uint64_t sbeg,
                                               uint64_t send, uint64_t nbound, int shift, uint32_t xm,
                                               uint2 *__restrict__ out, uint32_t *err, uint32_t *res = nullptr,
                                               const uint32_t *abort_flag = nullptr) {
  static_assert(!(RES && ORDERED), "reserved runs land in atomic order");
  constexpr int F = 1 << BITS;
  constexpr int DPL = (F + kWave - 1) / kWave;  // digits per lane of wave 0
  // a stripe whose first tuple is only 8-byte aligned is read from one tuple
  // earlier (16-byte pairs); that tuple is not part of the stripe: skip it
  const uint64_t skip = ((uintptr_t)(in + sbeg) & 15) ? 1 : 0;
  const uint64_t beg = sbeg - skip;
  const uint64_t end = send;
  if (sbeg >= end) return;
  const uint4 *in4 = reinterpret_cast<const uint4 *>(in + beg);
  const uint64_t nfull = (end - beg) / 2;              // complete 16-byte pairs
  const uint64_t npairs = (end - beg + 1) / 2;         // incl. a trailing single tuple

  // Tiles are read with 16-byte loads only: a load of another width into the
  // same registers would make the compiler wait for every earlier load (and
  // store) before each prefetch load.  The stripe's odd last tuple, if any, is
  // read once here and substituted when its tile is consumed.
  const bool odd_tail = nfull < npairs;
  uint2 last = make_uint2(0u, 0u);
  if (odd_tail) last = in[beg + 2 * nfull];
  // Loads j in [j0, j1) of the tile at pair t0pair, through a buffer resource
  // over that tile's complete pairs: lanes past the end read zero in hardware
  // (no memory access), and all loads share one 32-bit offset register
  // instead of a 64-bit address each.
  // The next tile's first kPre loads are issued at the top of each tile when
  // their 4 kPre registers fit beside the rest (1024 threads leave 128 VGPRs:
  // all 8 spill, 6 fit but time the same as 4, and the ordered ranks and the
  // 10- and 11-bit scans leave no room), the rest once this tile is staged.
  constexpr int kPre = (!ORDERED && BITS <= 9) ? kFastPairs / 2 : 0;
  auto load_part = [&](uint64_t t0pair, uint4 *dst, int j0, int j1) {
    if constexpr (BITS >= 10) {  // plain loads: the resource's registers tip these scans into spills
#pragma unroll
      for (int j = j0; j < j1; ++j) {
        const uint64_t pi = t0pair + (uint64_t)j * kFastNT + threadIdx.x;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (pi < nfull) v = in4[pi];
        dst[j] = v;
      }
      return;
    }
    const uint64_t left = t0pair < nfull ? nfull - t0pair : 0;
    const uint32_t nrec = (uint32_t)min<uint64_t>(left, (uint64_t)kFastNT * kFastPairs) * 16u;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)(in4 + min(t0pair, nfull)), 0, (int)nrec, kBufferRsrcWord3);
#pragma unroll
    for (int j = j0; j < j1; ++j) {
      typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
      const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)((j * kFastNT + threadIdx.x) * 16u), 0, 0);
      dst[j] = make_uint4(v.x, v.y, v.z, v.w);
    }
  };
  uint4 cur[kFastPairs];
  uint4 pre[kPre > 0 ? kPre : 1];  // the next tile's first loads, in flight while this tile is ranked and staged
  load_part(0, cur, 0, kFastPairs);
  __syncthreads();
  // abort_flag: an LDS word the caller wrote before this barrier; nonzero
  // moves nothing (every thread reads the same word after the barrier)
  if (abort_flag && *abort_flag) return;

  for (uint64_t t0 = 0; t0 < npairs; t0 += (uint64_t)kFastNT * kFastPairs) {
    const uint64_t tnext = t0 + (uint64_t)kFastNT * kFastPairs;
    load_part(tnext, pre, 0, kPre);
    const uint32_t tn = (uint32_t)min((uint64_t)kFastTile, (end - beg) - 2 * t0);  // tuples read in this tile
    const uint32_t tskip = t0 == 0 ? (uint32_t)skip : 0u;                            // lead-in tuple, not ours
    if (odd_tail) {
#pragma unroll
      for (int j = 0; j < kFastPairs; ++j)
        if (t0 + (uint64_t)j * kFastNT + threadIdx.x == nfull) cur[j] = make_uint4(last.x, last.y, 0u, 0u);
    }
    uint32_t rk[2 * kFastPairs];
    if constexpr (!ORDERED && RES) {
#pragma unroll
      for (int j = 0; j < kFastPairs; ++j) {
        const uint32_t e = 2u * (j * kFastNT + threadIdx.x);
        if (e >= tskip && e < tn) rk[2 * j] = atomicAdd(&sh.cnt[digit_of(cur[j].x, xm, shift, F - 1)], 1u);
        if (e + 1 < tn) rk[2 * j + 1] = atomicAdd(&sh.cnt[digit_of(cur[j].z, xm, shift, F - 1)], 1u);
      }
      __syncthreads();
    } else if constexpr (!ORDERED) {
      // as above, but a wave whose valid tuples all share one digit (sorted
      // input: 128 consecutive keys) ranks them with one atomic instead of
      // 128 on one word.  (The reserving scatter above serves sampled
      // partitions, which hand sorted input to this one.)
      const uint64_t lt = lanemask_lt();
#pragma unroll
      for (int j = 0; j < kFastPairs; ++j) {
        const uint32_t e = 2u * (j * kFastNT + threadIdx.x);
        const bool v0 = e >= tskip && e < tn, v1 = e + 1 < tn;
        const uint32_t d0 = digit_of(cur[j].x, xm, shift, F - 1), d1 = digit_of(cur[j].z, xm, shift, F - 1);
        const uint32_t dr = __builtin_amdgcn_readfirstlane(d0);
        if (__ballot((v0 && d0 != dr) || (v1 && d1 != dr)) == 0) {  // wave-uniform
          const uint64_t m0 = __ballot(v0), m1 = __ballot(v1);
          const uint32_t n0 = (uint32_t)__popcll(m0), nv = n0 + (uint32_t)__popcll(m1);
          uint32_t base = 0;
          if (lane_id() == 0 && nv) base = atomicAdd(&sh.cnt[dr], nv);
          base = __builtin_amdgcn_readfirstlane(base);
          rk[2 * j] = base + (uint32_t)__popcll(m0 & lt);
          rk[2 * j + 1] = base + n0 + (uint32_t)__popcll(m1 & lt);
        } else {
          if (v0) rk[2 * j] = atomicAdd(&sh.cnt[d0], 1u);
          if (v1) rk[2 * j + 1] = atomicAdd(&sh.cnt[d1], 1u);
        }
      }
      __syncthreads();
    } else {
      // each wave ranks its own tuples in (register, lane) order: the lanes
      // that share a digit are ranked by lane, the group's first lane moves
      // the wave's counter of that digit (distinct digits, distinct words)
      const int w = threadIdx.x / kWave;
      const uint64_t lt = lanemask_lt();
      for (int d = lane_id(); d < F; d += kWave) sh.wc[w][d] = 0;
#pragma unroll
      for (int q = 0; q < 2 * kFastPairs; ++q) {
        const int j = q >> 1;
        const uint32_t e = 2u * (j * kFastNT + threadIdx.x) + (q & 1);
        const bool valid = e >= tskip && e < tn;
        const uint32_t d = digit_of((q & 1) ? cur[j].z : cur[j].x, xm, shift, F - 1);
        const uint64_t peers = match_peers<BITS>(d, __ballot(valid));
        const uint32_t before = (uint32_t)__popcll(peers & lt);
        const int lead = valid ? __ffsll((unsigned long long)peers) - 1 : lane_id();
        uint32_t first = 0;
        if (valid && before == 0) {
          first = sh.wc[w][d];
          sh.wc[w][d] = first + (uint32_t)__popcll(peers);
        }
        rk[q] = __shfl(first, lead) + before;
      }
      __syncthreads();
      // per digit: the tile's count, and each wave's offset within the digit
      for (int d = threadIdx.x; d < F; d += kFastNT) {
        uint32_t run = 0;
        for (int v = 0; v < kFastNT / kWave; ++v) {
          const uint32_t c = sh.wc[v][d];
          sh.wc[v][d] = run;
          run += c;
        }
        sh.cnt[d] = run;
      }
      __syncthreads();
    }
    // wave 0: the tile's digit starts in staging (exclusive scan of the counts,
    // DPL consecutive digits per lane), each digit's output base for this tile,
    // the cursors advanced past it and the counts cleared for the next tile
    uint32_t rbase[RES ? DPL : 1];  // RES: the reserved run bases, in flight during staging
    if (threadIdx.x < kWave) {
      uint32_t loc[DPL];
      uint32_t sum = 0;
#pragma unroll
      for (int k = 0; k < DPL; ++k) {
        const int d = threadIdx.x * DPL + k;
        loc[k] = d < F ? sh.cnt[d] : 0u;
        sum += loc[k];
      }
      const uint32_t inc = wave_incl_scan(sum);
      uint32_t ex = inc - sum;
#pragma unroll
      for (int k = 0; k < DPL; ++k) {
        const int d = threadIdx.x * DPL + k;
        if (d < F) {
          if constexpr (RES) {
            rbase[k] = loc[k] ? atomic_add_agent(res + d, loc[k]) : 0u;
          } else {
            const uint32_t c = sh.cursor[d];
            sh.pbase[d] = c - ex;
            sh.cursor[d] = c + loc[k];
          }
          sh.tbase[d] = ex;
          sh.cnt[d] = 0;
        }
        ex += loc[k];
      }
    }
    __syncthreads();
    // a tuple's stage slot: its digit's start, (ORDERED) its wave's offset within the digit, its rank
    auto slot = [&](uint32_t d, uint32_t r) -> uint32_t {
      if constexpr (ORDERED) return sh.tbase[d] + sh.wc[threadIdx.x / kWave][d] + r;
      else return sh.tbase[d] + r;
    };
#pragma unroll
    for (int j = 0; j < kFastPairs; ++j) {
      const uint32_t e = 2u * (j * kFastNT + threadIdx.x);
      if (e >= tskip && e < tn) sh.stage[slot(digit_of(cur[j].x, xm, shift, F - 1), rk[2 * j])] = make_uint2(cur[j].x, cur[j].y);
      if (e + 1 < tn) sh.stage[slot(digit_of(cur[j].z, xm, shift, F - 1), rk[2 * j + 1])] = make_uint2(cur[j].z, cur[j].w);
    }
    // RES: the run bases arrive; a tile all of whose runs fit their regions
    // (the usual case) writes without a per-tuple bound check
    bool dirty = false;
    if constexpr (RES) {
      if (threadIdx.x < kWave) {
        bool over = false;
#pragma unroll
        for (int k = 0; k < DPL; ++k) {
          const int d = threadIdx.x * DPL + k;
          if (d < F) {
            const uint32_t tb = sh.tbase[d];
            const uint32_t cnt = (d + 1 < F ? sh.tbase[d + 1] : tn - tskip) - tb;
            sh.pbase[d] = rbase[k] - tb;
            over |= cnt != 0 && (uint64_t)rbase[k] + cnt > sh.cursor[d];
          }
        }
        if constexpr (BITS <= 10) {
          const uint64_t any = __ballot(over);
          if (threadIdx.x == 0) scatter_dirty_flag() = any != 0;
        } else {
          (void)over;
        }
      }
      __syncthreads();
      // (11 bits: the staging tile and the digit arrays fill the whole LDS,
      // so every tuple is checked)
      if constexpr (BITS <= 10) dirty = __builtin_amdgcn_readfirstlane((int)scatter_dirty_flag()) != 0;
      else dirty = true;
    } else {
      __syncthreads();
    }
    // the next tile: its first kPre loads arrived during this tile's rank
    // phase, the rest are read while this tile's runs are written
#pragma unroll
    for (int j = 0; j < kPre; ++j) cur[j] = pre[j];
    load_part(tnext, cur, kPre, kFastPairs);
    // no barrier after the flush: the next tile touches stage, tbase and pbase
    // only after its first barrier, which every wave reaches after its flush
    for (uint32_t i = threadIdx.x; i < tn - tskip; i += kFastNT) {
      const uint2 t = sh.stage[i];
      const uint32_t d = digit_of(t.x, xm, shift, F - 1);
      const uint64_t pos = (uint32_t)(sh.pbase[d] + i);
      if constexpr (RES) {
        // (the bound is read only when some run of the tile passes its region's end)
        if (!dirty || pos < sh.cursor[d]) out[pos] = t;
        else atomic_or_agent(err, kErrCapacity);  // the region's estimate was too small
      } else {
        if (pos < nbound) out[pos] = t;
        else atomic_or_agent(err, kErrOverflow);  // inconsistent counts: report, never write outside
      }
    }
  }
  __syncthreads();
}

// ---- sampled partition sizes ------------------------------------------------------
// A relation's partition sizes estimated from a sample: of every kFastTile-
// tuple tile's 1024 lines of 128 bytes one line in each group of 16 is read,
// at an offset hashed from (tile, group), so 1/16 of the bytes, spread so that
// no arrangement that repeats with the tile or the line skews it.
constexpr uint32_t kGapLineTuples = 16;                                        // 128-byte lines
constexpr uint32_t kGapGroup = 16;                                             // lines per sampled line
constexpr uint32_t kGapLinesPerTile = kFastTile / kGapLineTuples / kGapGroup;  // 64
constexpr int kGapSampleNT = kGapLinesPerTile * 8;  // 8 threads (16-byte pairs) per line
constexpr int kGapSampleUnroll = 8;                 // tiles in flight per workgroup
constexpr int kGapClumpEvery = 8;                   // the clumping sums count one sampled line in 8
constexpr double kGapKappaMax = 2.0;                // clumping the single-join buffers are sized for
constexpr double kGapKappaNoise = 0.25;             // taken off a measured clumping ratio (its noise)
constexpr double kGapZ = 6.0;                       // margin in standard deviations
constexpr uint32_t kGapPad = 16;                    // plus a fixed margin, in tuples

// v from lane (lane ^ R) of the same 32-lane half (R < 8: the 8 lanes of one
// sampled line), by ds_swizzle's bit-mask mode (and 0x1f, or 0, xor R)
template <int R>
__device__ __forceinline__ uint32_t swizzle_xor(uint32_t v) {
  static_assert(R > 0 && R < 32, "xor within 32 lanes");
  return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1f | (R << 10));
}

// For the two digits a and b (< 2^15) a lane holds of a 16-tuple line spread
// over 8 lanes (two tuples each), packed p = a | b << 16 (0xffffffff for a
// lane holding none): ca and cb are how many of the line's tuples have digit a
// and b, and (second form) ha and hb how many have its high part (the digit
// >> LO; 0xffff >> LO is no high part of a digit under 2^15).  Every lane of
// the wave executes it.
__device__ __forceinline__ void line_same_counts(uint32_t p, uint32_t &ca, uint32_t &cb) {
  const uint32_t a = p & 0xffffu, b = p >> 16;
  ca = 1u + (a == b);
  cb = ca;
  auto add = [&](uint32_t o) {
    const uint32_t oa = o & 0xffffu, ob = o >> 16;
    ca += (uint32_t)(a == oa) + (uint32_t)(a == ob);
    cb += (uint32_t)(b == oa) + (uint32_t)(b == ob);
  };
  add(swizzle_xor<1>(p));
  add(swizzle_xor<2>(p));
  add(swizzle_xor<3>(p));
  add(swizzle_xor<4>(p));
  add(swizzle_xor<5>(p));
  add(swizzle_xor<6>(p));
  add(swizzle_xor<7>(p));
}
template <int LO>
__device__ __forceinline__ void line_same_counts(uint32_t p, uint32_t &ca, uint32_t &cb, uint32_t &ha,
                                                 uint32_t &hb) {
  const uint32_t a = p & 0xffffu, b = p >> 16, a1 = a >> LO, b1 = b >> LO;
  ca = 1u + (a == b);
  cb = ca;
  ha = 1u + (a1 == b1);
  hb = ha;
  auto add = [&](uint32_t o) {
    const uint32_t oa = o & 0xffffu, ob = o >> 16, oa1 = oa >> LO, ob1 = ob >> LO;
    ca += (uint32_t)(a == oa) + (uint32_t)(a == ob);
    cb += (uint32_t)(b == oa) + (uint32_t)(b == ob);
    ha += (uint32_t)(a1 == oa1) + (uint32_t)(a1 == ob1);
    hb += (uint32_t)(b1 == oa1) + (uint32_t)(b1 == ob1);
  };
  add(swizzle_xor<1>(p));
  add(swizzle_xor<2>(p));
  add(swizzle_xor<3>(p));
  add(swizzle_xor<4>(p));
  add(swizzle_xor<5>(p));
  add(swizzle_xor<6>(p));
  add(swizzle_xor<7>(p));
}

__host__ __device__ __forceinline__ uint32_t gap_line(uint32_t t, uint32_t j) {
  uint32_t h = (t * kGapLinesPerTile + j) * 0x9e3779b1u;
  h ^= h >> 15;
  h *= 0x85ebca77u;
  h ^= h >> 13;
  return j * kGapGroup + (h >> 28);  // line of the tile, in [16 j, 16 j + 16)
}

// Region capacity for a sampled count s at scale f (tuples per sampled
// tuple): f times the upper confidence bound of a Poisson mean given s
// observations, chi^2(2(s + 1)) / 2 at kGapZ standard deviations by the
// Wilson-Hilferty approximation, (s + 1) (1 - 1/(9 (s + 1)) + z/(3 sqrt(s +
// 1)))^3 — about s + z sqrt(s) for large s and still a safe bound where a
// partition was sampled a few times or never (thinning a count by 1/f varies
// less than a Poisson draw) — plus kGapPad, rounded up to a whole line.
// (w in single precision, raised by 2^-20 to cover its rounding; the
// product in double)
//
// kappa >= 1 is the sampled count's variance over its mean.  The sample takes
// whole lines, so a line whose tuples share a digit (sorted or grouped input)
// adds to that digit's count in steps of more than one: clumps of about kappa
// tuples, of which s / kappa were seen, stand in for the Poisson observation
// and the bound is kappa times theirs.  kappa = 1 is the bound above.
__device__ __forceinline__ uint64_t gap_cap(uint32_t s, float kappa, double f) {
  const float mf = (float)s / kappa + 1.0f;
  const float wf = (1.0f - 1.0f / (9.0f * mf) + (float)kGapZ / (3.0f * sqrtf(mf))) * (1.0f + 0x1p-20f);
  const double m = (double)s / (double)kappa + 1.0, w = (double)wf;
  const uint64_t k = (uint64_t)ceil(f * (double)kappa * m * w * w * w) + kGapPad;
  return (k + kGapLineTuples - 1) / kGapLineTuples * kGapLineTuples;
}

// kappa of sampled counts from q, the sum over their sampled tuples of how
// many tuples of their line have the same digit, and g, the number of lines
// the sample could have taken each count from.  With one line taken from each
// group of kGapGroup, a count's variance is the sum over the groups of E[c^2]
// - E[c]^2 over the group's lines (c: the line's tuples of the digit); q
// estimates the sum of E[c^2], and s^2 / g is at most the sum of E[c]^2
// (Cauchy-Schwarz over the g groups), so k = (q - s^2 / g) / s bounds the
// variance over the mean from above in expectation; pooled over several
// digits (s = sum of the counts, s2 = sum of their squares) it is their
// count-weighted mean.  Independent tuples give about 1 - 2^-bits (less
// where a digit fills much of a line), every line alike (all tuples of one
// digit) gives 0: no sampling error; sorted input gives about 16, keys in
// adjacent pairs about 2.  k is itself an estimate (pooled over a range's
// digits, random keys stay within 1.03 at 2^20 tuples), so kappa = k -
// kGapKappaNoise, never below 1: no estimate of independent tuples inflates a
// margin, and a real ratio K keeps a margin of at least z sqrt((K - 0.25) /
// K) standard deviations (5.6 at K = 2).
__device__ __forceinline__ float gap_kappa(double s, double s2, double q, double g) {
  if (s <= 0.0 || g <= 0.0) return 1.0f;
  const double k = (q - s2 / g) / s - kGapKappaNoise;
  return k > 1.0 ? (float)k : 1.0f;
}


// ---- LDS histogram of one stripe --------------------------------------------------
// Adds the digit counts of in[beg, end) to the LDS array hist[2^BITS] with LDS
// atomics (the caller zeroes it).  V16: 16-byte loads from the first 16-byte
// aligned tuple on.
template <int BITS, int NT, bool V16>
__device__ __forceinline__ void hist_stripe(uint32_t *hist, const uint2 *__restrict__ in, uint64_t beg, uint64_t end,
                                            int shift, uint32_t xm) {
  constexpr int F = 1 << BITS;
  if (beg < end) {
    // 16-byte loads from the first 16-byte aligned tuple on
    const uint64_t a0 = (V16 && ((uintptr_t)(in + beg) & 15)) ? beg + 1 : beg;
    if (V16 && a0 < end) {
      if (a0 != beg && threadIdx.x == 0) atomicAdd(&hist[digit_of(in[beg].x, xm, shift, F - 1)], 1u);
      const uint4 *in4 = reinterpret_cast<const uint4 *>(in + a0);
      const uint64_t npairs = (end - a0) / 2;
      uint64_t i = threadIdx.x;
      for (; i + 3 * NT < npairs; i += 4 * NT) {
        uint4 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = ld_stream(in4 + i + u * NT);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          atomicAdd(&hist[digit_of(q[u].x, xm, shift, F - 1)], 1u);
          atomicAdd(&hist[digit_of(q[u].z, xm, shift, F - 1)], 1u);
        }
      }
      for (; i < npairs; i += NT) {
        const uint4 q = in4[i];
        atomicAdd(&hist[digit_of(q.x, xm, shift, F - 1)], 1u);
        atomicAdd(&hist[digit_of(q.z, xm, shift, F - 1)], 1u);
      }
      if (((end - a0) & 1) && threadIdx.x == 0) atomicAdd(&hist[digit_of(in[end - 1].x, xm, shift, F - 1)], 1u);
    } else {
      for (uint64_t i = beg + threadIdx.x; i < end; i += NT) atomicAdd(&hist[digit_of(in[i].x, xm, shift, F - 1)], 1u);
    }
  }
}

}  // namespace tile
}  // namespace eo
