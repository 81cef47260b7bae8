// Batched execution of many joins of uneven sizes.
//
// The reference runs one join per process, with a host synchronisation after
// every launch (joins/hashJoin_gpu/main_street/hj.cc:454-1105).  For a stream
// of small and medium joins those launch and sync gaps dominate.  Here a batch
// of joins runs as a fixed, host-sync-free sequence:
//   1. partition: the relations of all joins that need the same number of
//      radix bits form one group; one histogram launch, one u64 scan and one
//      scatter launch partition every relation of the group (each stripe
//      finds its relation by binary search in a descriptor table);
//   2. plan: one thread per (join, partition) pair counts its probe tiles,
//      a scan numbers them, and one thread per tile (up to an estimate of
//      the count) writes the tile's descriptor;
//   3. probe: a persistent kernel (one workgroup per CU) drains a device work
//      queue of (join, partition, R chunk, S span) tiles with an atomic
//      ticket and runs the fused LDS probe on each, writing into the join's
//      own output with the join's own output cursor.
// The only host interaction is reading the per-join counts at the end, so
// the whole sequence can also be captured once into a hipGraph and replayed.
#include <algorithm>

#include "device.hpp"
#include "kernels.hpp"
#include "tile_kernels.hpp"

namespace eo {
namespace {

using tile::kFastNT;
using BProbe = tile::ProbeSmall;  // the batch probe's LDS table: two workgroups per CU
constexpr uint32_t kFCap = BProbe::Cap;
using tile::kFNT;
using tile::SpanRef;

constexpr int kBHistNT = 512;
static_assert(kFCap == kBatchChunk, "batch tile chunk is the probe table's capacity");
static_assert(tile::kFastTile == kPartTile, "the host's tile size is the scatter's");

__device__ __forceinline__ uint32_t find_seg(const BatchSeg *segs, uint32_t nseg, uint32_t g) {
  uint32_t lo = 0, hi = nseg;  // segs[lo].stripe_base <= g < segs[hi].stripe_base
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (segs[mid].stripe_base <= g) lo = mid; else hi = mid;
  }
  return lo;
}

// Descriptors read from the device tables, their pointers marked global (see
// as_global: flat accesses would make every LDS wait also wait for them).
__device__ __forceinline__ BatchSeg load_seg(const BatchSeg *segs, uint32_t i) {
  BatchSeg sg = segs[i];
  sg.in = as_global(sg.in);
  sg.out = as_global(sg.out);
  sg.part_start = as_global(sg.part_start);
  sg.hs = as_global(sg.hs);
  sg.re = as_global(sg.re);
  sg.cur = as_global(sg.cur);
  sg.part_end = as_global(sg.part_end);
  return sg;
}
__device__ __forceinline__ BatchJoin load_join(const BatchJoin *joins, uint32_t j) {
  BatchJoin J = joins[j];
  J.R = as_global(J.R);
  J.S = as_global(J.S);
  J.r_start = as_global(J.r_start);
  J.s_start = as_global(J.s_start);
  J.r_end = as_global(J.r_end);
  J.s_end = as_global(J.s_end);
  J.out = as_global(J.out);
  J.cursor = as_global(J.cursor);
  J.matched = as_global(J.matched);
  return J;
}

template <int BITS>
__global__ __launch_bounds__(kBHistNT) void k_bseg_hist(const BatchSeg *__restrict__ segs, uint32_t nseg,
                                                        uint32_t total_stripes, uint64_t len,
                                                        uint64_t *__restrict__ counts) {
  constexpr int F = 1 << BITS;
  __shared__ uint32_t hist[F];
  const uint32_t g = xcd_local_index(blockIdx.x, gridDim.x);
  if (g >= total_stripes) return;
  const BatchSeg sg = load_seg(segs, find_seg(segs, nseg, g));
  const uint32_t s = g - sg.stripe_base;
  const uint64_t beg = (uint64_t)s * len;
  const uint64_t end = beg + len < sg.n ? beg + len : sg.n;
  for (int d = threadIdx.x; d < F; d += kBHistNT) hist[d] = 0;
  __syncthreads();
  tile::hist_stripe<BITS, kBHistNT, true>(hist, sg.in, beg, end, 0, 0u);
  __syncthreads();
  for (int d = threadIdx.x; d < F; d += kBHistNT) counts[sg.count_base + (uint64_t)d * sg.nstripes + s] = hist[d];
}

// RES (sampled groups): the stripe reserves its runs from the relation's
// region cursors; otherwise its offsets come from the scanned counts.
template <int BITS, bool RES = false>
__global__ __launch_bounds__(kFastNT) void k_bseg_scatter(const BatchSeg *__restrict__ segs, uint32_t nseg,
                                                          uint32_t total_stripes, uint64_t len,
                                                          const uint64_t *__restrict__ offsets, uint32_t *err) {
  __shared__ tile::ScatterShared<BITS> sh;
  const uint32_t g = xcd_local_index(blockIdx.x, gridDim.x);
  if (g >= total_stripes) return;
  const BatchSeg sg = load_seg(segs, find_seg(segs, nseg, g));
  const uint32_t s = g - sg.stripe_base;
  const uint64_t beg = (uint64_t)s * len;
  const uint64_t end = beg + len < sg.n ? beg + len : sg.n;
  // RES (sampled groups; err is the batch's own capacity flag): once a layout
  // has found a relation's regions too large for its buffer, or a region
  // overflowed, the batch is redone exactly, so the stripes still to come
  // move nothing (read beside the cursors, tested after the first barrier)
  __shared__ uint32_t s_err;
  if constexpr (RES) {
    if (threadIdx.x == 0) s_err = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int d = threadIdx.x; d < (1 << BITS); d += kFastNT) {
      sh.cursor[d] = sg.re[d];
      sh.cnt[d] = 0;
    }
  } else {
    const uint64_t base = offsets[sg.count_base];  // tuples of the group's earlier relations
    for (int d = threadIdx.x; d < (1 << BITS); d += kFastNT) {
      const uint32_t c = (uint32_t)(offsets[sg.count_base + (uint64_t)d * sg.nstripes + s] - base);
      sh.cursor[d] = c;
      sh.cnt[d] = 0;
      // the relation's first stripe starts every digit: its offsets are the
      // partition starts (relative to the relation)
      if (s == 0) sg.part_start[d] = c;
    }
    if (s == 0 && threadIdx.x == 0) sg.part_start[1 << BITS] = (uint32_t)sg.n;
  }
  tile::scatter_stripe<BITS, false, RES>(sh, sg.in, beg, end, sg.n, 0, 0u, sg.out, err, RES ? sg.cur : nullptr,
                                         RES ? &s_err : nullptr);
}

// ---- sampled groups ----------------------------------------------------------------
// Workgroup u samples kBatchSampleTiles tiles of one relation (the tile's
// sampled lines, as the single-join sampled partition reads them) into an LDS
// histogram and adds it to the relation's counts hs with one atomic per
// non-empty digit.  It also measures the lines' clumping as the single-join
// sample does: wave w takes the w-th of every kGapSampleUnroll tiles and sums,
// for each of its tuples, the tuples of its line with the same digit, added
// with the number of lines measured to hs[2^bits] and hs[2^bits + 1].
__device__ __forceinline__ uint32_t find_samp(const BatchSeg *segs, uint32_t nseg, uint32_t u) {
  uint32_t lo = 0, hi = nseg;  // segs[lo].samp_base <= u < segs[hi].samp_base
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (segs[mid].samp_base <= u) lo = mid; else hi = mid;
  }
  return lo;
}

template <int BITS>
__global__ __launch_bounds__(tile::kGapSampleNT) void k_bsamp(const BatchSeg *__restrict__ segs, uint32_t nseg,
                                                              uint32_t total_samp) {
  constexpr uint32_t F = 1u << BITS;
  constexpr int NT = tile::kGapSampleNT, U = tile::kGapSampleUnroll;
  __shared__ uint32_t hist[F];
  const uint32_t u = blockIdx.x;
  if (u >= total_samp) return;
  const BatchSeg sg = load_seg(segs, find_samp(segs, nseg, u));
  for (uint32_t d = threadIdx.x; d < F; d += NT) hist[d] = 0;
  __syncthreads();
  const uint32_t T = (uint32_t)((sg.n + tile::kFastTile - 1) / tile::kFastTile);
  const uint32_t t0 = (u - sg.samp_base) * kBatchSampleTiles, t1 = min(T, t0 + kBatchSampleTiles);
  const uint64_t npairs = sg.n / 2;
  const uint4 *in4 = reinterpret_cast<const uint4 *>(sg.in);
  const uint32_t j = threadIdx.x >> 3, q = threadIdx.x & 7;
  static_assert(NT / kWave == U && F <= 0x8000u, "one wave per unrolled tile; digits packed two to a word");
  uint32_t qsum = 0, lines = 0;  // clumping sum (per lane), lines measured (per wave)
  for (uint32_t tb = t0; tb < t1; tb += U) {
    uint4 v[U];
    bool ok[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint32_t t = tb + (uint32_t)k;
      const uint64_t p = (uint64_t)t * (tile::kFastTile / 2) +
                         (uint64_t)tile::gap_line(t, j) * (tile::kGapLineTuples / 2) + q;
      ok[k] = t < t1 && p < npairs;
      v[k] = ok[k] ? ld_stream(in4 + p) : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint32_t d0 = v[k].x & (F - 1), d1 = v[k].z & (F - 1);
      if (k == (int)(threadIdx.x / kWave)) {  // wave-uniform: this wave's tile of the step
        uint32_t c0, c1;
        tile::line_same_counts(ok[k] ? d0 | d1 << 16 : ~0u, c0, c1);
        if (ok[k]) qsum += c0 + c1;
        lines += (uint32_t)__popcll(__ballot(ok[k] && q == 0));  // a line's first pair is in
      }
      if (!ok[k]) continue;
      atomicAdd(&hist[d0], 1u);
      atomicAdd(&hist[d1], 1u);
    }
  }
  qsum = wave_sum(qsum);
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < F; d += NT)
    if (hist[d]) atomic_add_agent(sg.hs + d, hist[d]);
  if ((threadIdx.x & (kWave - 1)) == 0 && lines) {
    atomic_add_agent(sg.hs + F, qsum);
    atomic_add_agent(sg.hs + F + 1, lines);
  }
}

// One workgroup per relation: region d = [part_start[d], re[d]) sized by
// gap_cap from the sampled counts and the relation's clumping ratio
// (gap_kappa), cursor cur[d] = part_start[d]; a total beyond the relation's
// buffer (clumped input, or an empty sample) raises kErrCapacity and every
// region is clipped to the buffer.
template <int BITS>
__global__ __launch_bounds__(1024) void k_blayout(const BatchSeg *__restrict__ segs, uint32_t *err) {
  constexpr uint32_t F = 1u << BITS, NT = 1024, BPT = F > NT ? F / NT : 1;
  __shared__ uint64_t red[NT / kWave + 1];
  const BatchSeg sg = load_seg(segs, blockIdx.x);
  const uint32_t b0 = threadIdx.x * BPT;
  uint32_t sv[BPT];
  uint64_t loc = 0, sq = 0;
#pragma unroll
  for (uint32_t k = 0; k < BPT; ++k) {
    sv[k] = b0 + k < F ? sg.hs[b0 + k] : 0u;
    loc += sv[k];
    sq += (uint64_t)sv[k] * sv[k];
  }
  const uint64_t ns = block_sum<NT, uint64_t>(loc, red);
  const uint64_t ns2 = block_sum<NT, uint64_t>(sq, red);
  const double f = ns ? (double)sg.n / (double)ns : 0.0;
  // groups of lines the sample draws from: 64 per tile, those of the tail
  // tuples in the last; the clumping sum scaled from the lines measured
  const uint64_t T = (sg.n + tile::kFastTile - 1) / tile::kFastTile;
  const uint64_t tail = sg.n - (T - 1) * tile::kFastTile;
  const double g = (double)(T - 1) * tile::kGapLinesPerTile +
                   (double)((tail + tile::kGapLineTuples * tile::kGapGroup - 1) / (tile::kGapLineTuples * tile::kGapGroup));
  const uint32_t qs = sg.hs[F], m = sg.hs[F + 1];
  const float kap = (ns && m) ? tile::gap_kappa((double)ns, (double)ns2, (double)qs * g / m, g) : 1.0f;
  uint64_t c[BPT];
  uint64_t sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < BPT; ++k) {
    c[k] = (ns && b0 + k < F) ? tile::gap_cap(sv[k], kap, f) : 0;
    sum += c[k];
  }
  uint64_t tot;
  uint64_t at = block_excl_scan<NT, uint64_t>(sum, red, &tot);
  if (threadIdx.x == 0 && (ns == 0 || tot > sg.budget)) atomic_or_agent(err, kErrCapacity);
#pragma unroll
  for (uint32_t k = 0; k < BPT; ++k) {
    if (b0 + k < F) {
      const uint64_t a = min(at, sg.budget), e = min(at + c[k], sg.budget);
      sg.part_start[b0 + k] = (uint32_t)a;
      sg.cur[b0 + k] = (uint32_t)a;
      sg.re[b0 + k] = (uint32_t)e;
    }
    at += c[k];
  }
  if (threadIdx.x == 0) sg.part_start[F] = (uint32_t)min(tot, sg.budget);
}

// part_end[d] = the filled end of every partition of the group's relations.
template <int BITS>
__global__ void k_bends(const BatchSeg *__restrict__ segs, uint32_t nseg) {
  constexpr uint32_t F = 1u << BITS;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)nseg * F) return;
  const BatchSeg sg = load_seg(segs, (uint32_t)(i / F));
  const uint32_t d = (uint32_t)(i % F);
  sg.part_end[d] = min(sg.cur[d], sg.re[d]);
}

// pair_join[pair_base + p] = j for every partition p of join j (one workgroup per join).
__global__ void k_bpairs(const BatchJoin *__restrict__ joins, uint32_t *__restrict__ pair_join) {
  const uint32_t j = blockIdx.x;
  const BatchJoin J = load_join(joins, j);
  for (uint32_t p = threadIdx.x; p < J.P; p += blockDim.x) pair_join[J.pair_base + p] = j;
}

// tiles[q] for every (join, partition) pair q; tiles[npairs] = 0.
__global__ void k_bplan(const BatchJoin *__restrict__ joins, const uint32_t *__restrict__ pair_join, uint32_t npairs,
                        uint32_t *tiles) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > npairs) return;
  if (q == npairs) {
    tiles[npairs] = 0;
    return;
  }
  const BatchJoin J = load_join(joins, pair_join[q]);
  const uint32_t p = q - J.pair_base;
  const uint32_t rn = J.r_end[p] - J.r_start[p];
  const uint32_t sn = J.s_end[p] - J.s_start[p];
  tiles[q] = (rn && sn) ? ((rn + kFCap - 1) / kFCap) * ((sn + J.s_span - 1) / J.s_span) : 0u;
}

// Descriptor of probe tile t of pair q (tile_start[q] <= t < tile_start[q + 1]):
// its R chunk and S span, numbered chunk-minor.
__device__ __forceinline__ BatchTile tile_of_pair(const BatchJoin *joins, const uint32_t *pair_join,
                                                  const uint32_t *tile_start, uint32_t q, uint32_t t) {
  const uint32_t jn = pair_join[q];
  const BatchJoin J = load_join(joins, jn);
  const uint32_t p = q - J.pair_base;
  const uint32_t local = t - tile_start[q];
  const uint32_t rb = J.r_start[p], rn = J.r_end[p] - rb;
  const uint32_t sb = J.s_start[p], sn = J.s_end[p] - sb;
  const uint32_t nc = (rn + kFCap - 1) / kFCap;
  const uint32_t c = local % nc, k = local / nc;
  BatchTile T;
  T.R = J.R + rb + c * kFCap;
  T.S = J.S;
  T.j = jn;
  T.r_n = min((uint32_t)kFCap, rn - c * kFCap);
  T.s_beg = sb + k * J.s_span;
  T.s_n = min(J.s_span, sn - k * J.s_span);
  return T;
}

// The descriptor of every probe tile t < min(total, cap).
__global__ void k_btiles(const BatchJoin *__restrict__ joins, const uint32_t *__restrict__ pair_join,
                         const uint32_t *__restrict__ tile_start, uint32_t npairs, BatchTile *__restrict__ tiles,
                         uint32_t cap) {
  const uint32_t total = tile_start[npairs];
  const uint32_t n = min(total, cap);
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = npairs;  // tile_start[lo] <= t < tile_start[hi]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (tile_start[mid] <= t) lo = mid; else hi = mid;
    }
    tiles[t] = tile_of_pair(joins, pair_join, tile_start, lo, t);
  }
}

__device__ __forceinline__ const uint2 *uniform_ptr(const uint2 *p) {
  const uint64_t v = (uint64_t)(uintptr_t)p;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return as_global((const uint2 *)(uintptr_t)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// Persistent probe: every workgroup takes tiles from the queue until it is
// empty; every workgroup leaves the loop once its ticket passes the total.
// A workgroup takes its next ticket as it starts a tile, so that the atomic's
// round trip overlaps that tile's loads instead of delaying the next tile;
// every ticket below the total that a workgroup takes it also runs.  Wave 0
// puts the tile's descriptor in LDS: read from the table written by the plan
// (one load before the tile's data loads) or, past the table, found by search.
// (A grid of one workgroup per estimated tile, taking only the overflow from
// the queue, measured slower: C5 batched 99.9 -> 104.9 ms.)
__global__ __launch_bounds__(kFNT) __attribute__((amdgpu_waves_per_eu(8))) void k_bprobe(const BatchJoin *__restrict__ joins,
                                                 const uint32_t *__restrict__ pair_join,
                                                 const uint32_t *__restrict__ tile_start, uint32_t npairs,
                                                 const BatchTile *__restrict__ tiles, uint32_t tiles_cap,
                                                 uint32_t *queue_head) {
  __shared__ tile::ProbeShared<BProbe> sh;
  __shared__ BatchTile cur;
  __shared__ uint32_t item;
  const uint32_t total = tile_start[npairs];
  uint32_t ticket = 0;  // thread 0: the ticket of the next tile
  if (threadIdx.x == 0) ticket = ticket_next(queue_head);
  for (;;) {
    if (threadIdx.x < kWave) {
      const uint32_t t = __shfl(ticket, 0);
      if (t < total) {
        BatchTile T;
        if (t < tiles_cap) T = tiles[t];
        else T = tile_of_pair(joins, pair_join, tile_start, wave_find(tile_start, npairs, t), t);
        if (threadIdx.x == 0) cur = T;
      }
      if (threadIdx.x == 0) item = t;
    }
    __syncthreads();
    const uint32_t t = item;
    const BatchTile T = cur;
    __syncthreads();
    if (t >= total) break;
    const BatchJoin J = load_join(joins, uniform_u32(T.j));
    SpanRef ref;
    ref.r_beg = 0;
    ref.r_n = uniform_u32(T.r_n);
    ref.s_beg = uniform_u32(T.s_beg);
    ref.s_n = uniform_u32(T.s_n);
    if (threadIdx.x == 0) ticket = ticket_next(queue_head);
    tile::probe_tile<BProbe>(sh, uniform_ptr(T.R), uniform_ptr(T.S), ref, J.out, J.out_cap, J.cursor,
                                     J.wide != 0, J.matched, 0, uniform_u32(J.swapped) != 0);
  }
}

// Null rows of the batch's left outer joins: block b belongs to the join j
// with emit_start[j] <= b < emit_start[j + 1] and covers 256 bitmap words
// (8192 S positions) of it; one atomic per block on the join's cursor.
__global__ __launch_bounds__(256) void k_bemit(const BatchJoin *__restrict__ joins, uint32_t njoins,
                                               const uint32_t *__restrict__ emit_start) {
  __shared__ uint64_t red[256 / kWave + 1];
  __shared__ uint64_t base_sh;
  const uint32_t j = wave_find(emit_start, njoins, blockIdx.x);
  const BatchJoin J = load_join(joins, j);
  const uint64_t wd = (uint64_t)(blockIdx.x - emit_start[j]) * 256 + threadIdx.x;
  uint32_t unm = 0;
  if (wd * 32 < J.nS) {
    const uint64_t left = J.nS - wd * 32;
    const uint32_t valid = left >= 32 ? ~0u : ((1u << left) - 1u);
    unm = ~J.matched[wd] & valid;
  }
  uint64_t tot;
  const uint64_t excl = block_excl_scan<256, uint64_t>((uint64_t)__popc(unm), red, &tot);
  if (threadIdx.x == 0) base_sh = tot ? atomic_add_agent(J.cursor, tot) : 0ull;
  __syncthreads();
  if (!J.out) return;
  uint64_t pos = base_sh + excl;
  while (unm) {
    const uint32_t b = (uint32_t)__ffs(unm) - 1u;
    unm &= unm - 1u;
    const uint2 t = J.S[wd * 32 + b];
    if (pos < J.out_cap) emit_null(J.out, pos, t.y, t.x, J.wide != 0);
    ++pos;
  }
}

template <int BITS>
hipError_t launch_group(const BatchGroup &g, uint64_t *counts, void *scan_scratch, uint32_t *err, hipStream_t stream) {
  const uint32_t grid = (g.total_stripes + kXcds - 1) / kXcds * kXcds;
  if (g.sampled) {
    // sample -> region layout -> scatter with reserved runs -> partition ends
    // (planned up to kBatchSampleMaxBits; wider kernels are not built)
    if constexpr (BITS > kBatchSampleMaxBits) {
      return hipErrorInvalidValue;
    } else {
      hipError_t e = fill_u32(g.hs_all, 0u, (uint64_t)g.nseg * batch_hs_words(BITS), stream);
      if (e != hipSuccess) return e;
      EO_TIMED("batch_sample", stream,
               hipLaunchKernelGGL(k_bsamp<BITS>, dim3(g.total_samp), dim3(tile::kGapSampleNT), 0, stream, g.segs,
                                  g.nseg, g.total_samp));
      if ((e = hipGetLastError()) != hipSuccess) return e;
      hipLaunchKernelGGL(k_blayout<BITS>, dim3(g.nseg), dim3(1024), 0, stream, g.segs, err);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      EO_TIMED("batch_scatter", stream,
               hipLaunchKernelGGL((k_bseg_scatter<BITS, true>), dim3(grid), dim3(kFastNT), 0, stream, g.segs, g.nseg,
                                  g.total_stripes, g.stripe_len, nullptr, err));
      if ((e = hipGetLastError()) != hipSuccess) return e;
      const uint64_t ends = (uint64_t)g.nseg << BITS;
      hipLaunchKernelGGL(k_bends<BITS>, dim3((uint32_t)((ends + 255) / 256)), dim3(256), 0, stream, g.segs, g.nseg);
      return hipGetLastError();
    }
  }
  EO_TIMED("batch_hist", stream,
           hipLaunchKernelGGL(k_bseg_hist<BITS>, dim3(grid), dim3(kBHistNT), 0, stream, g.segs, g.nseg,
                              g.total_stripes, g.stripe_len, counts));
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = scan_u64(counts, counts, g.ncounts, nullptr, scan_scratch, err, stream);
  if (e != hipSuccess) return e;
  EO_TIMED("batch_scatter", stream,
           hipLaunchKernelGGL(k_bseg_scatter<BITS>, dim3(grid), dim3(kFastNT), 0, stream, g.segs, g.nseg,
                              g.total_stripes, g.stripe_len, counts, err));
  return hipGetLastError();
}

}  // namespace

uint64_t batch_stripe_len(int bits) {
  // One 16K-tuple tile per partition workgroup up to 8 bits (workgroups
  // handed out in address order keep the scatter's writes near each other);
  // longer stripes for more digits, whose counts (2^bits u64 per stripe,
  // written, scanned and read) would otherwise pass 6 % of the bytes moved.
  return bits <= 8 ? (uint64_t)tile::kFastTile : bits == 9 ? 2ull * tile::kFastTile : 4ull * tile::kFastTile;
}

hipError_t batch_partition_group(const BatchGroup &g, uint64_t *counts, void *scan_scratch, uint32_t *err,
                                 hipStream_t stream) {
  if (g.nseg == 0) return hipSuccess;
  switch (g.bits) {
#define EO_BG_CASE(B) \
  case B:             \
    return launch_group<B>(g, counts, scan_scratch, err, stream);
    EO_BG_CASE(1) EO_BG_CASE(2) EO_BG_CASE(3) EO_BG_CASE(4) EO_BG_CASE(5) EO_BG_CASE(6)
    EO_BG_CASE(7) EO_BG_CASE(8) EO_BG_CASE(9) EO_BG_CASE(10) EO_BG_CASE(11)
#undef EO_BG_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t batch_plan(const BatchJoin *joins, uint32_t njoins, uint32_t *pair_join, uint32_t npairs,
                      uint32_t *tile_start, BatchTile *tiles, uint32_t tiles_cap, void *scan_scratch, uint32_t *err,
                      hipStream_t stream) {
  if (njoins == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bpairs, dim3(njoins), dim3(256), 0, stream, joins, pair_join);
  hipLaunchKernelGGL(k_bplan, dim3((npairs + 1 + 255) / 256), dim3(256), 0, stream, joins, pair_join, npairs,
                     tile_start);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = scan_u32(tile_start, tile_start, (uint64_t)npairs + 1, nullptr, scan_scratch, err, stream);
  if (e != hipSuccess || tiles_cap == 0 || npairs == 0) return e;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(((uint64_t)tiles_cap + 255) / 256, 4096);
  hipLaunchKernelGGL(k_btiles, dim3(blocks), dim3(256), 0, stream, joins, pair_join, tile_start, npairs, tiles,
                     tiles_cap);
  return hipGetLastError();
}

hipError_t batch_emit_unmatched(const BatchJoin *joins, uint32_t njoins, const uint32_t *emit_start,
                                uint32_t blocks, hipStream_t stream) {
  if (njoins == 0 || blocks == 0) return hipSuccess;
  EO_TIMED("outer_rows", stream,
           hipLaunchKernelGGL(k_bemit, dim3(blocks), dim3(256), 0, stream, joins, njoins, emit_start));
  return hipGetLastError();
}

hipError_t batch_probe(const BatchJoin *joins, const uint32_t *pair_join, const uint32_t *tile_start, uint32_t npairs,
                       const BatchTile *tiles, uint32_t tiles_cap, uint32_t *queue_head, int cus,
                       hipStream_t stream) {
  if (npairs == 0) return hipSuccess;
  EO_TIMED("batch_probe", stream,
           hipLaunchKernelGGL(k_bprobe, dim3(2u * (uint32_t)std::max(cus, 1)), dim3(kFNT), 0, stream, joins,
                              pair_join, tile_start, npairs, tiles, tiles_cap, queue_head));
  return hipGetLastError();
}

}  // namespace eo
