// Stable radix partitioning of 8-byte tuples.
//
// The reference partitions with one thread per 256 consecutive tuples,
// 16-thread blocks and per-thread digit histograms (lcount_partitioning /
// lpartitioning, joins/hashJoin_gpu/main_street/partitioning.cu:15-131),
// 6 bits per pass.  Here one pass is:
//   1. k_part_hist: one workgroup per contiguous stripe, LDS histogram with
//      LDS atomics, counts written digit-major: counts[d * G + g];
//   2. exclusive scan of the F*G counts (decoupled look-back, scan.hip): the
//      result is, per digit and stripe, where that stripe's run of the digit
//      starts in the output; partition starts fall out as counts[d * G + 0];
//   3. k_part_scatter: each workgroup walks its stripe in tiles; every wave
//      ranks its 64 tuples per step by ballot-matched peers (stable), the tile
//      is placed digit-sorted into an LDS staging buffer and flushed so that
//      consecutive lanes write consecutive addresses of one digit's run.
// Stripes are dealt so that consecutive stripes share an XCD, which makes
// neighbouring runs of one digit meet in the same L2.
#include <algorithm>

#include "device.hpp"
#include "kernels.hpp"
#include "tile_kernels.hpp"

namespace eo {
namespace {

using tile::digit_of;
using tile::kFastNT;
using tile::kFastTile;

constexpr int kPartNT = 512;
constexpr int kMsdGroupsPerWg = 4;   // second-pass groups per persistent workgroup (exact partition)
constexpr int kGapGroupsPerWg = 16;  // the same for the sampled partition

template <int BITS>
struct PartCfg {
  static constexpr int F = 1 << BITS;
  static constexpr int NT = kPartNT;
  static constexpr int NW = NT / kWave;
  static constexpr int IPT = BITS <= 8 ? 16 : 8;   // tuples per thread per tile
  static constexpr int TILE = NT * IPT;
  static constexpr int WSPAN = TILE / NW;          // tuples per wave per tile
  static constexpr int DPT = (F + NT - 1) / NT;    // digits per thread
};


// Where a histogram / scatter workgroup's stripe lies and where its counts
// go.  Uniform (seg_start == nullptr): stripe g = [g*len, (g+1)*len) of the
// whole array, sub-stripe k of K; count of digit d at (d*G + g)*K + k.
// Segmented (second pass of an MSD partition): the array is already split into
// nseg segments [seg_start[i], seg_start[i+1]); segment i owns stripes
// [seg_base[i], seg_base[i+1]) of length len, and the counts of segment i are
// laid out [d][stripe] from seg_base[i] * F — one exclusive scan over all
// counts then orders the output segment-major, digit, stripe.
struct StripeGeo {
  uint64_t n;
  uint64_t len;
  uint32_t G, K;
  const uint32_t *seg_start;
  const uint32_t *seg_base;
  uint32_t nseg;
};

struct StripeRef {
  uint64_t beg, end;
  uint64_t cidx0;     // count index of digit 0
  uint64_t cstride;   // distance between consecutive digits' counts
  uint32_t seg;       // segment (segmented mode) or 0
};

// whole: the scatter's view — the full stripe, counts indexed at sub-stripe 0.
template <int F>
__device__ __forceinline__ bool stripe_of(const StripeGeo &geo, uint32_t g, uint32_t k, StripeRef *r,
                                          bool whole = false) {
  if (!geo.seg_start) {
    const uint64_t s0 = (uint64_t)g * geo.len;
    const uint64_t s1 = s0 + geo.len < geo.n ? s0 + geo.len : geo.n;
    if (geo.K == 1 || whole) {
      r->beg = s0;
      r->end = s1;
    } else {
      const uint64_t sub = (geo.len / geo.K + 1) & ~1ull;  // even: sub-stripes stay 16-byte aligned
      r->beg = s0 + k * sub;
      r->end = r->beg + sub;
      if (k == geo.K - 1 || r->end > s1) r->end = s1;
    }
    r->cidx0 = (uint64_t)g * geo.K + k;
    r->cstride = (uint64_t)geo.G * geo.K;
    r->seg = 0;
    return true;
  }
  if (g >= geo.seg_base[geo.nseg]) return false;
  uint32_t lo = 0, hi = geo.nseg;  // seg_base[lo] <= g < seg_base[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (geo.seg_base[mid] <= g) lo = mid; else hi = mid;
  }
  const uint32_t sidx = g - geo.seg_base[lo];
  const uint32_t k1 = geo.seg_base[lo + 1] - geo.seg_base[lo];
  r->beg = geo.seg_start[lo] + (uint64_t)sidx * geo.len;
  r->end = r->beg + geo.len < geo.seg_start[lo + 1] ? r->beg + geo.len : geo.seg_start[lo + 1];
  r->cidx0 = (uint64_t)geo.seg_base[lo] * F + sidx;
  r->cstride = k1;
  r->seg = lo;
  return true;
}

// Histogram of one pass.  The grid is G stripes x K sub-stripes; counts are
// laid out [digit][stripe][sub] so that one exclusive scan gives, for every
// (digit, stripe), the output offset of the stripe's run at index
// (d * G + g) * K — the scatter reads those.  V16: 16-byte loads (pairs of
// tuples), used when the input is 16-byte aligned.
template <int BITS, bool V16>
__global__ __launch_bounds__(kPartNT) void k_part_hist(const uint2 *__restrict__ in, StripeGeo geo, int shift,
                                                       uint32_t xm, uint32_t *__restrict__ counts) {
  using C = PartCfg<BITS>;
  __shared__ uint32_t hist[C::F];
  uint32_t g, k;
  if (geo.seg_start) {
    g = blockIdx.x;  // grid is an upper bound: spread the real stripes over all XCDs
    k = 0;
  } else {
    const uint32_t gk = xcd_local_index(blockIdx.x, geo.G * geo.K);
    g = gk / geo.K;
    k = gk % geo.K;
  }
  StripeRef sr;
  if (!stripe_of<C::F>(geo, g, k, &sr)) return;
  for (int d = threadIdx.x; d < C::F; d += C::NT) hist[d] = 0;
  __syncthreads();
  tile::hist_stripe<BITS, C::NT, V16>(hist, in, sr.beg, sr.end, shift, xm);
  __syncthreads();
  for (int d = threadIdx.x; d < C::F; d += C::NT) counts[sr.cidx0 + (uint64_t)d * sr.cstride] = hist[d];
}

template <int BITS>
__global__ __launch_bounds__(kPartNT) void k_part_scatter(const uint2 *__restrict__ in, uint64_t n, int shift,
                                                          uint32_t xm, uint32_t G, uint64_t stripe_len,
                                                          uint32_t K, const uint32_t *__restrict__ offsets,
                                                          uint2 *__restrict__ out, uint32_t *err) {
  using C = PartCfg<BITS>;
  constexpr int F = C::F;
  __shared__ uint32_t cursor[F];
  __shared__ uint32_t tbase[F + 1];
  __shared__ uint16_t wcnt[C::NW][F];
  __shared__ uint2 stage[C::TILE];
  __shared__ uint32_t scan_tmp[C::NW + 1];

  const uint32_t g = xcd_local_index(blockIdx.x, G);
  for (int d = threadIdx.x; d < F; d += C::NT) cursor[d] = offsets[((uint64_t)d * G + g) * K];
  const uint64_t beg = (uint64_t)g * stripe_len;
  const uint64_t end = beg + stripe_len < n ? beg + stripe_len : n;
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  const uint64_t lt = lanemask_lt();

  for (uint64_t t0 = beg; t0 < end; t0 += C::TILE) {
    const uint32_t tn = (uint32_t)((end - t0) < (uint64_t)C::TILE ? (end - t0) : (uint64_t)C::TILE);
    for (int i = threadIdx.x; i < C::NW * F / 2; i += C::NT) reinterpret_cast<uint32_t *>(&wcnt[0][0])[i] = 0;
    __syncthreads();

    uint2 v[C::IPT];
    uint32_t dg[C::IPT];
    uint32_t rk[C::IPT];
#pragma unroll
    for (int it = 0; it < C::IPT; ++it) {
      const uint32_t li = (uint32_t)(w * C::WSPAN + it * kWave + lane);
      const bool valid = li < tn;
      v[it] = valid ? in[t0 + li] : make_uint2(0u, 0u);
      dg[it] = digit_of(v[it].x, xm, shift, F - 1);
    }
#pragma unroll
    for (int it = 0; it < C::IPT; ++it) {
      const uint32_t li = (uint32_t)(w * C::WSPAN + it * kWave + lane);
      const bool valid = li < tn;
      const uint64_t peers = match_peers<BITS>(dg[it], __ballot(valid));
      const uint32_t before = (uint32_t)__popcll(peers & lt);
      uint32_t old = 0;
      if (valid) old = wcnt[w][dg[it]];
      rk[it] = old + before;
      if (valid && before == 0) wcnt[w][dg[it]] = (uint16_t)(old + (uint32_t)__popcll(peers));
    }
    __syncthreads();

    // Per digit: exclusive offsets over waves (in place) and the tile total.
    for (int d = threadIdx.x; d < F; d += C::NT) {
      uint32_t run = 0;
#pragma unroll
      for (int ww = 0; ww < C::NW; ++ww) {
        const uint32_t c = wcnt[ww][d];
        wcnt[ww][d] = (uint16_t)run;
        run += c;
      }
      tbase[d] = run;
    }
    __syncthreads();
    {
      uint32_t loc[C::DPT];
      uint32_t sum = 0;
#pragma unroll
      for (int j = 0; j < C::DPT; ++j) {
        const int d = threadIdx.x * C::DPT + j;
        loc[j] = d < F ? tbase[d] : 0u;
        sum += loc[j];
      }
      uint32_t tot;
      uint32_t ex = block_excl_scan<C::NT, uint32_t>(sum, scan_tmp, &tot);
#pragma unroll
      for (int j = 0; j < C::DPT; ++j) {
        const int d = threadIdx.x * C::DPT + j;
        if (d < F) tbase[d] = ex;
        ex += loc[j];
      }
      if (threadIdx.x == 0) tbase[F] = tn;
    }
    __syncthreads();

#pragma unroll
    for (int it = 0; it < C::IPT; ++it) {
      const uint32_t li = (uint32_t)(w * C::WSPAN + it * kWave + lane);
      if (li < tn) stage[tbase[dg[it]] + wcnt[w][dg[it]] + rk[it]] = v[it];
    }
    __syncthreads();

    for (uint32_t i = threadIdx.x; i < tn; i += C::NT) {
      const uint2 t = stage[i];
      const uint32_t d = digit_of(t.x, xm, shift, F - 1);
      const uint64_t pos = (uint64_t)cursor[d] + (i - tbase[d]);
      if (pos < n) out[pos] = t;
      else atomic_or_agent(err, kErrOverflow);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < F; d += C::NT) cursor[d] += tbase[d + 1] - tbase[d];
  }
}


// Unstable scatter for the join's internal partitioning: a workgroup of 1024
// threads stages 16K tuples per tile; ranks come from LDS atomics (the
// histogram and the within-tile ranks in one pass), the tile is placed
// digit-sorted in LDS and flushed in digit runs.  The next tile's loads are
// issued before the flush so that HBM reads overlap the writes.  Within a
// partition the tuple order depends on atomic order; the multiset per
// partition equals the stable scatter's.
template <int BITS, bool ORDERED = false>
__global__ __launch_bounds__(kFastNT) void k_part_scatter_fast(const uint2 *__restrict__ in, StripeGeo geo, int shift,
                                                               uint32_t xm, const uint32_t *__restrict__ offsets,
                                                               uint2 *__restrict__ out, uint32_t *err) {
  __shared__ tile::ScatterShared<BITS, ORDERED> sh;
  // segmented: the grid is an upper bound, so spread the real stripes over all XCDs
  const uint32_t g = geo.seg_start ? blockIdx.x : xcd_local_index(blockIdx.x, gridDim.x);
  StripeRef sr;
  if (!stripe_of<(1 << BITS)>(geo, g, 0, &sr, true)) return;
  for (int d = threadIdx.x; d < (1 << BITS); d += kFastNT) {
    sh.cursor[d] = offsets[sr.cidx0 + (uint64_t)d * sr.cstride];
    sh.cnt[d] = 0;
  }
  tile::scatter_stripe<BITS, ORDERED>(sh, in, sr.beg, sr.end, geo.n, shift, xm, out, err);
}

__global__ void k_part_starts(const uint32_t *offsets, uint32_t G, uint32_t K, uint32_t F, uint64_t n,
                              uint32_t *part_start) {
  const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d < F) part_start[d] = offsets[(uint64_t)d * G * K];
  if (d == F) part_start[F] = (uint32_t)n;
}

__global__ void k_part_bounds(const uint2 *__restrict__ sorted, uint64_t n, int shift, uint32_t xm, uint32_t F,
                              uint32_t *part_start) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x) {
    const int64_t dprev = i == 0 ? -1 : (int64_t)digit_of(sorted[i - 1].x, xm, shift, F - 1);
    const int64_t dcur = i == n ? (int64_t)F : (int64_t)digit_of(sorted[i].x, xm, shift, F - 1);
    for (int64_t d = dprev + 1; d <= dcur; ++d) part_start[d] = (uint32_t)i;
  }
}

template <int BITS>
hipError_t launch_pass(const PartPassPlan &p, const eo_tuple *in, eo_tuple *out, uint32_t *counts,
                       void *scan_scratch, uint32_t *err, hipStream_t stream) {
  const uint2 *vin = reinterpret_cast<const uint2 *>(in);
  uint2 *vout = reinterpret_cast<uint2 *>(out);
  const bool aligned = ((uintptr_t)in & 7) == 0;
  StripeGeo geo{p.n, p.stripe_len, p.stripes, p.hist_split, nullptr, nullptr, 0};
  const uint32_t hist_grid = p.stripes * p.hist_split;
  if (aligned)
    EO_TIMED("part_hist", stream, hipLaunchKernelGGL((k_part_hist<BITS, true>), dim3(hist_grid), dim3(kPartNT), 0, stream, vin, geo, p.shift,
                       p.xor_mask, counts));
  else
    EO_TIMED("part_hist", stream, hipLaunchKernelGGL((k_part_hist<BITS, false>), dim3(hist_grid), dim3(kPartNT), 0, stream, vin, geo, p.shift,
                       p.xor_mask, counts));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const uint64_t m = ((uint64_t)p.stripes * p.hist_split) << BITS;
  e = scan_u32(counts, counts, m, nullptr, scan_scratch, err, stream);
  if (e != hipSuccess) return e;
  if (p.fast) {
    if ((uintptr_t)in & 15) return hipErrorInvalidValue;  // stripe 0 starts at the array start
    if constexpr (BITS <= 7) {
      if (p.ordered) {
        EO_TIMED("part_scatter", stream,
                 hipLaunchKernelGGL((k_part_scatter_fast<BITS, true>), dim3(p.stripes), dim3(kFastNT), 0, stream, vin,
                                    geo, p.shift, p.xor_mask, counts, vout, err));
        return hipGetLastError();
      }
    } else {
      if (p.ordered) return hipErrorInvalidValue;  // per-wave counters fit up to 7 bits
    }
    EO_TIMED("part_scatter", stream, hipLaunchKernelGGL(k_part_scatter_fast<BITS>, dim3(p.stripes), dim3(kFastNT), 0, stream, vin, geo, p.shift,
                       p.xor_mask, counts, vout, err));
  } else {
    EO_TIMED("part_scatter_stable", stream, hipLaunchKernelGGL(k_part_scatter<BITS>, dim3(p.stripes), dim3(kPartNT), 0, stream, vin, p.n, p.shift,
                       p.xor_mask, p.stripes, p.stripe_len, p.hist_split, counts, vout, err));
  }
  return hipGetLastError();
}

// ---- segmented (second MSD) pass ------------------------------------------------

__global__ void k_seg_stripes(const uint32_t *seg_start, uint32_t nseg, uint64_t len, uint32_t *seg_base) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nseg) seg_base[i] = (uint32_t)(((uint64_t)(seg_start[i + 1] - seg_start[i]) + len - 1) / len);
  if (i == nseg) seg_base[nseg] = 0;
}

// part_start[seg * F + d] = start of digit d inside segment seg.
__global__ void k_seg_starts(const uint32_t *offsets, const uint32_t *seg_start, const uint32_t *seg_base,
                             uint32_t nseg, uint32_t F, uint64_t n, uint32_t *part_start) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t tot = (uint64_t)nseg * F;
  if (i < tot) {
    const uint32_t sg = (uint32_t)(i / F), d = (uint32_t)(i % F);
    const uint32_t k1 = seg_base[sg + 1] - seg_base[sg];
    part_start[i] = k1 ? offsets[(uint64_t)seg_base[sg] * F + (uint64_t)d * k1] : seg_start[sg];
  } else if (i == tot) {
    part_start[tot] = (uint32_t)n;
  }
}

template <int BITS>
hipError_t launch_seg_pass(const SegPassPlan &p, const eo_tuple *in, eo_tuple *out, const uint32_t *seg_start,
                           uint32_t *part_start, void *scratch, uint32_t *err, hipStream_t stream) {
  constexpr uint32_t F = 1u << BITS;
  char *sc = (char *)scratch;
  uint32_t *seg_base = (uint32_t *)sc;
  sc += ((p.nseg + 1) * sizeof(uint32_t) + 255) / 256 * 256;
  uint32_t *counts = (uint32_t *)sc;
  const uint64_t m = (uint64_t)p.max_stripes * F;
  sc += (m * sizeof(uint32_t) + 255) / 256 * 256;
  void *scan_scratch = sc;

  hipLaunchKernelGGL(k_seg_stripes, dim3((p.nseg + 1 + 255) / 256), dim3(256), 0, stream, seg_start, p.nseg,
                     p.stripe_len, seg_base);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = scan_u32(seg_base, seg_base, p.nseg + 1, nullptr, scan_scratch, err, stream);
  if (e == hipSuccess) e = fill_u32(counts, 0u, m, stream);  // counts of absent stripes stay 0
  if (e != hipSuccess) return e;
  const StripeGeo geo{p.n, p.stripe_len, 0, 1, seg_start, seg_base, p.nseg};
  const uint2 *vin = reinterpret_cast<const uint2 *>(in);
  const uint32_t grid = (p.max_stripes + kXcds - 1) / kXcds * kXcds;
  EO_TIMED("seg_hist", stream, hipLaunchKernelGGL((k_part_hist<BITS, true>), dim3(grid), dim3(kPartNT), 0, stream, vin, geo, p.shift, 0u,
                     counts));
  e = hipGetLastError();
  if (e == hipSuccess) e = scan_u32(counts, counts, m, nullptr, scan_scratch, err, stream);
  if (e != hipSuccess) return e;
  EO_TIMED("seg_scatter", stream, hipLaunchKernelGGL(k_part_scatter_fast<BITS>, dim3(grid), dim3(kFastNT), 0, stream, vin, geo, p.shift, 0u, counts,
                     reinterpret_cast<uint2 *>(out), err));
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const uint64_t tot = (uint64_t)p.nseg * F + 1;
  hipLaunchKernelGGL(k_seg_starts, dim3((uint32_t)((tot + 255) / 256)), dim3(256), 0, stream, counts, seg_start,
                     seg_base, p.nseg, F, p.n, part_start);
  return hipGetLastError();
}

// ---- two-pass MSD partition from one histogram ------------------------------------
// The histogram counts the combined digit bin = key & (2^(bh+bl) - 1) =
// dh * F2 + dl.  The first pass runs one workgroup per kFastTile-tuple tile
// (grid = the tile count, no loop over a stripe): a radix scatter writes into
// every digit's region at once, and with persistent stripe workgroups the
// writes in flight are spread over every stripe's share of every region,
// while tiles handed out in address order (consecutive tiles on one XCD)
// keep them near each other, each region filled from a few moving fronts.
// Its offsets come from per-tile counts of the high digit dh, which the
// histogram workgroup of each coarse stripe g (TPS consecutive tiles) writes
// as it walks its tiles, laid out [dh][tile] so that one exclusive scan gives
// tile t's run start at off1[dh * T + t].  The same workgroup accumulates the
// combined-digit histogram of the whole stripe and writes it as one row; a
// transpose (k_msd_reduce, instead of device-scope atomics, which bypass the
// per-XCD L2) turns the rows into counts2[bin * G + g].  The first pass writes
// stripe g's tuples of high digit dh (its tiles' runs) as one contiguous
// chunk (dh, g), chunks ordered by g inside segment dh, so after one exclusive
// scan of counts2 (order dh, dl, g) the tuples of digit pair (dh, dl) that
// come from chunks g0..g1-1 have one contiguous output range starting at
// off2[(dh * F2 + dl) * G + g0].  The second pass therefore needs no
// histogram of its own: its work items are groups of consecutive chunks of one
// segment, sized so that the groups fill the GPU in about one wave.
template <int TB, int BL>
__global__ __launch_bounds__(kPartNT) void k_msd_hist_tiles(const uint2 *__restrict__ in, uint64_t n, uint32_t G,
                                                            uint32_t TPS, uint32_t *__restrict__ counts1,
                                                            uint32_t *__restrict__ partial) {
  constexpr uint32_t FT = 1u << TB, F1 = FT >> BL;
  constexpr uint32_t NW = kPartNT / kWave;
  constexpr uint32_t kPairsPerTile = kFastTile / 2;
  constexpr uint32_t kGroups = kPairsPerTile / (4 * kPartNT);  // 4-load groups per thread per tile
  static_assert(kPairsPerTile % (4 * kPartNT) == 0, "a tile is whole 4-load groups");
  __shared__ uint32_t hist[FT];
  __shared__ uint32_t th_w[NW][F1];  // per-wave tile counts of the high digit
  const uint32_t g = xcd_local_index(blockIdx.x, G);
  uint32_t *th = th_w[threadIdx.x / kWave];
  for (uint32_t b = threadIdx.x; b < FT; b += kPartNT) hist[b] = 0;
  for (uint32_t b = threadIdx.x; b < NW * F1; b += kPartNT) (&th_w[0][0])[b] = 0;
  __syncthreads();
  const uint4 *in4 = reinterpret_cast<const uint4 *>(in);
  const uint64_t npairs = n / 2;  // complete pairs; an odd last tuple is counted once at the end
  auto count = [&](uint32_t key) {
    const uint32_t b = key & (FT - 1);
    atomicAdd(&hist[b], 1u);
    atomicAdd(&th[b >> BL], 1u);
  };
  // the two tuples of a load; every lane of the wave calls it.  A wave whose
  // valid tuples all share one high digit (sorted input: 128 consecutive
  // keys) adds its count of them with one atomic instead of 64 on one word.
  auto count_pair = [&](uint4 v, bool ok) {
    const uint32_t b0 = v.x & (FT - 1), b1 = v.z & (FT - 1);
    if (ok) {
      atomicAdd(&hist[b0], 1u);
      atomicAdd(&hist[b1], 1u);
    }
    const uint32_t h0 = b0 >> BL, h1 = b1 >> BL;
    const uint32_t hr = __builtin_amdgcn_readfirstlane(h0);
    if (__ballot(ok && (h0 != hr || h1 != hr)) == 0) {  // wave-uniform
      const uint32_t nv = 2u * (uint32_t)__popcll(__ballot(ok));
      if (lane_id() == 0 && nv) atomicAdd(&th[hr], nv);
    } else if (ok) {
      atomicAdd(&th[h0], 1u);
      atomicAdd(&th[h1], 1u);
    }
  };
  // (loads of the next tile issued before each tile's barrier measured no
  // faster: 0.270 against 0.264 ms per C2 launch)
  for (uint32_t k = 0; k < TPS; ++k) {
    const uint64_t t = (uint64_t)g * TPS + k;
    const uint64_t p0 = t * kPairsPerTile;
    for (uint32_t j = 0; j < kGroups; ++j) {
      uint4 q[4];
      bool ok[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint64_t p = p0 + (uint64_t)(j * 4 + u) * kPartNT + threadIdx.x;
        ok[u] = p < npairs;
        q[u] = ok[u] ? ld_stream(in4 + p) : make_uint4(0u, 0u, 0u, 0u);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) count_pair(q[u], ok[u]);
    }
    // the odd last tuple of the input belongs to the tile that holds it
    if ((n & 1) && threadIdx.x == 0 && (n - 1) / kFastTile == t) count(in[n - 1].x);
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < F1; d += kPartNT) {
      uint32_t c = 0;
#pragma unroll
      for (uint32_t w = 0; w < NW; ++w) {
        c += th_w[w][d];
        th_w[w][d] = 0;
      }
      counts1[((uint64_t)d * G + g) * TPS + k] = c;
    }
    __syncthreads();
  }
  uint32_t *row = partial + (uint64_t)g * FT;
  for (uint32_t b = threadIdx.x; b < FT; b += kPartNT) row[b] = hist[b];
}

// First pass, one workgroup per tile: tile t's run of digit d starts at
// offsets[d * T + t] (the exclusive scan of the per-tile counts, digit-major).
// RES: the runs are reserved from region cursors instead: grid T is a
// multiple of 8 and the tiles of XCD range x = t / (T/8) use the cursors
// res[x 2^BITS + d] and region ends offsets[x 2^BITS + d].
template <int BITS, bool ORDERED, bool RES = false>
__global__ __launch_bounds__(kFastNT) void k_tile_scatter(const uint2 *__restrict__ in, uint64_t n, int shift,
                                                          const uint32_t *__restrict__ offsets,
                                                          uint2 *__restrict__ out, uint32_t *err,
                                                          uint32_t *res = nullptr) {
  __shared__ tile::ScatterShared<BITS, ORDERED> sh;
  const uint32_t T = gridDim.x;
  const uint32_t t = xcd_local_index(blockIdx.x, T);  // consecutive tiles on one XCD
  const uint64_t beg = (uint64_t)t * kFastTile;
  if (beg >= n) return;
  const uint64_t end = beg + kFastTile < n ? beg + kFastTile : n;
  // RES: the tile's XCD range x owns regions [x 2^BITS, (x + 1) 2^BITS) of
  // the ends (offsets) and cursors (res)
  const uint32_t rx = RES ? t / (T / kXcds) : 0u;
  // RES: once the error is raised (the layout found the sampled regions too
  // large for their buffers, or a region overflowed) the caller redoes the
  // join exactly, so the tiles still to come move nothing.  Read once per
  // workgroup beside the cursors; scatter_stripe tests it after its first
  // barrier (the tile's loads already issued), so the decision is the
  // workgroup's and costs no extra round trip.
  __shared__ uint32_t s_err;
  if (RES && threadIdx.x == 0) s_err = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (int d = threadIdx.x; d < (1 << BITS); d += kFastNT) {
    sh.cursor[d] = RES ? offsets[((uint64_t)rx << BITS) + d] : offsets[(uint64_t)d * T + t];
    sh.cnt[d] = 0;
  }
  tile::scatter_stripe<BITS, ORDERED, RES>(sh, in, beg, end, n, shift, 0u, out, err,
                                           RES ? res + ((uint64_t)rx << BITS) : nullptr, RES ? &s_err : nullptr);
}

// counts2[b * G + g] = sum over k of partial[(g * K + k) * FT + b]: a tiled
// transpose (64 bins x 64 stripes per workgroup) so that reads and writes are
// both coalesced.
__global__ __launch_bounds__(256) void k_msd_reduce(const uint32_t *__restrict__ partial, uint32_t G, uint32_t K,
                                                    uint32_t FT, uint32_t *__restrict__ counts2) {
  __shared__ uint32_t t[64][65];
  const uint32_t b0 = blockIdx.x * 64, g0 = blockIdx.y * 64;
  const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
  for (uint32_t gi = ty; gi < 64; gi += 4) {
    const uint32_t g = g0 + gi, b = b0 + tx;
    uint32_t sum = 0;
    if (g < G && b < FT)
      for (uint32_t k = 0; k < K; ++k) sum += partial[(uint64_t)(g * K + k) * FT + b];
    t[gi][tx] = sum;
  }
  __syncthreads();
  for (uint32_t bi = ty; bi < 64; bi += 4) {
    const uint32_t b = b0 + bi, g = g0 + tx;
    if (b < FT && g < G) counts2[(uint64_t)b * G + g] = t[tx][bi];
  }
}

// grp[dh] = number of second-pass groups of segment dh (>= 1, <= G); grp[F1] = 0
// (the scan turns grp into group bases); grp[F1 + 1] is the work ticket of
// k_msd_scatter, cleared here because this kernel runs just before it.
__global__ void k_msd_groups(const uint32_t *__restrict__ off1, uint32_t G, uint32_t K, uint32_t F1, uint64_t n,
                             uint64_t L, uint32_t *__restrict__ grp) {
  const uint32_t dh = blockIdx.x * blockDim.x + threadIdx.x;
  if (dh < F1) {
    const uint64_t beg = off1[(uint64_t)dh * G * K];
    const uint64_t end = dh + 1 < F1 ? off1[(uint64_t)(dh + 1) * G * K] : n;
    uint64_t J = (end - beg + L / 2) / L;
    grp[dh] = (uint32_t)(J < 1 ? 1 : (J > G ? G : J));
  } else if (dh == F1) {
    grp[F1] = 0;
    grp[F1 + 1] = 0;
  }
}

// Second pass, persistent: each workgroup takes groups from the ticket until
// none is left, so that segments of very different sizes (skewed keys) still
// keep every CU busy to the end.  A group is a run of consecutive first-pass
// chunks of one segment dh; it scatters by the low digit into the final
// partitions, whose starts off2 the combined histogram already holds.
template <int BL, bool ORDERED = false>
__global__ __launch_bounds__(kFastNT) void k_msd_scatter(const uint2 *__restrict__ in, uint64_t n,
                                                         const uint32_t *__restrict__ off1, uint32_t G, uint32_t K,
                                                         uint32_t F1, const uint32_t *__restrict__ grp_base,
                                                         uint32_t *ticket, const uint32_t *__restrict__ off2,
                                                         uint2 *__restrict__ out, uint32_t *err) {
  constexpr uint32_t F2 = 1u << BL;
  __shared__ tile::ScatterShared<BL, ORDERED> sh;
  __shared__ uint32_t s_w;
  const uint32_t total = grp_base[F1];
  // (one group per workgroup, consecutive groups on one XCD, measured slower:
  // 0.738 -> 0.772 ms per C2 launch)
  for (;;) {
    if (threadIdx.x == 0) s_w = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t w = s_w;
    if (w >= total) break;  // uniform: every workgroup leaves once the tickets pass total
    const uint32_t dh = wave_find(grp_base, F1, w);  // grp_base[dh] <= w < grp_base[dh + 1]
    const uint32_t J = grp_base[dh + 1] - grp_base[dh], j = w - grp_base[dh];
    const uint32_t g0 = (uint32_t)((uint64_t)j * G / J), g1 = (uint32_t)((uint64_t)(j + 1) * G / J);
    const uint64_t c0 = (uint64_t)dh * G + g0, c1 = (uint64_t)dh * G + g1;
    const uint64_t beg = off1[c0 * K];
    const uint64_t end = c1 == (uint64_t)F1 * G ? n : off1[c1 * K];
    for (uint32_t d = threadIdx.x; d < F2; d += kFastNT) {
      sh.cursor[d] = off2[((uint64_t)dh * F2 + d) * G + g0];
      sh.cnt[d] = 0;
    }
    __syncthreads();
    tile::scatter_stripe<BL, ORDERED>(sh, in, beg, end, n, 0, 0u, out, err);
    __syncthreads();  // sh and s_w are reused by the next group
  }
}

__global__ void k_msd_starts(const uint32_t *__restrict__ off2, uint32_t G, uint32_t FT, uint64_t n,
                             uint32_t *__restrict__ part_start) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < FT) part_start[i] = off2[(uint64_t)i * G];
  if (i == FT) part_start[FT] = (uint32_t)n;
}

// Geometry of the two-pass partition: T tiles of kFastTile tuples (the
// first pass's workgroups), G coarse stripes of TPS tiles each (the
// histogram's workgroups, and the second pass's unit of offsets).
struct Msd2Geo {
  uint32_t G, TPS, T;
};

Msd2Geo msd2_geo(uint64_t n, int cus) {
  const uint64_t tiles = std::max<uint64_t>(1, (n + kFastTile - 1) / kFastTile);
  uint64_t g = std::min<uint64_t>(tiles, (uint64_t)std::max(cus, kXcds));  // one histogram workgroup per CU
  g = (g + kXcds - 1) / kXcds * kXcds;
  const uint64_t tps = (tiles + g - 1) / g;
  return Msd2Geo{(uint32_t)g, (uint32_t)tps, (uint32_t)(g * tps)};
}

struct Msd2Layout {
  size_t o_c1, o_c2, o_part, o_grp, o_scan, total;
  uint64_t m1, m2, mp;
};

Msd2Layout msd2_layout(const Msd2Geo &geo, int bh, int bl) {
  Msd2Layout L{};
  const uint64_t F1 = 1ull << bh;
  L.m1 = (uint64_t)geo.T * F1;
  L.m2 = (F1 << bl) * geo.G;
  L.mp = L.m2;  // one combined-digit row per coarse stripe
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  L.o_c1 = 0;
  L.o_c2 = L.o_c1 + al(L.m1 * 4);
  L.o_part = L.o_c2 + al(L.m2 * 4);
  L.o_grp = L.o_part + al(L.mp * 4);
  L.o_scan = L.o_grp + al((F1 + 2) * 4);
  L.total = L.o_scan + std::max({scan_scratch_bytes_u32(L.m1), scan_scratch_bytes_u32(L.m2),
                                 scan_scratch_bytes_u32(F1 + 1)});
  return L;
}

template <int BH, int BL, bool ORDERED>
hipError_t launch_msd2(uint64_t n, const Msd2Geo &geo, uint64_t group_len, uint32_t workers, const eo_tuple *in,
                       eo_tuple *tmp, eo_tuple *out, uint32_t *part_start, void *scratch, uint32_t *err,
                       hipStream_t stream) {
  constexpr uint32_t F1 = 1u << BH, FT = 1u << (BH + BL);
  const Msd2Layout L = msd2_layout(geo, BH, BL);
  char *sc = (char *)scratch;
  uint32_t *c1 = (uint32_t *)(sc + L.o_c1), *c2 = (uint32_t *)(sc + L.o_c2), *grp = (uint32_t *)(sc + L.o_grp);
  uint32_t *part = (uint32_t *)(sc + L.o_part);
  void *scan_scratch = sc + L.o_scan;
  const uint2 *vin = reinterpret_cast<const uint2 *>(in);
  hipError_t e;
  EO_TIMED("part_hist", stream,
           hipLaunchKernelGGL((k_msd_hist_tiles<BH + BL, BL>), dim3(geo.G), dim3(kPartNT), 0, stream, vin, n, geo.G,
                              geo.TPS, c1, part));
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_msd_reduce, dim3((FT + 63) / 64, (geo.G + 63) / 64), dim3(256), 0, stream, part, geo.G, 1u,
                     FT, c2);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = scan_u32(c1, c1, L.m1, nullptr, scan_scratch, err, stream)) != hipSuccess) return e;
  if ((e = scan_u32(c2, c2, L.m2, nullptr, scan_scratch, err, stream)) != hipSuccess) return e;
  EO_TIMED("part_scatter", stream,
           hipLaunchKernelGGL((k_tile_scatter<BH, ORDERED>), dim3(geo.T), dim3(kFastNT), 0, stream, vin, n, BL, c1,
                              reinterpret_cast<uint2 *>(tmp), err));
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // the second pass's chunks are (high digit, coarse stripe): its first tile's
  // run starts the chunk, so the stripe's tiles act as its sub-stripes
  hipLaunchKernelGGL(k_msd_groups, dim3((F1 + 1 + 255) / 256), dim3(256), 0, stream, c1, geo.G, geo.TPS, F1, n,
                     group_len, grp);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = scan_u32(grp, grp, F1 + 1, nullptr, scan_scratch, err, stream)) != hipSuccess) return e;
  const uint64_t max_groups = F1 + n / group_len + 1;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(max_groups, std::max<uint32_t>(workers, 1));
  EO_TIMED("seg_scatter", stream,
           hipLaunchKernelGGL((k_msd_scatter<BL, ORDERED>), dim3(grid), dim3(kFastNT), 0, stream,
                              reinterpret_cast<const uint2 *>(tmp), n, c1, geo.G, geo.TPS, F1, grp, grp + F1 + 1,
                              c2, reinterpret_cast<uint2 *>(out), err));
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_msd_starts, dim3((FT + 1 + 255) / 256), dim3(256), 0, stream, c2, geo.G, FT, n, part_start);
  return hipGetLastError();
}

// ---- two-pass MSD partition into sampled regions ----------------------------
// The exact two-pass partition above reads its whole input once for the
// histogram before the first scatter.  This variant reads a sample instead
// (tile::gap_line: one 128-byte line in 16, 1/16 of the bytes).  From it, it
// sizes every final partition (combined digit) and every first-pass region at
// the scaled estimate plus kGapZ standard deviations of the sampling error,
// and each scatter tile reserves its runs from per-region cursors with one
// atomic per digit.  The first pass keeps one region per (high digit, XCD):
// the tiles of one XCD (a contiguous range of tiles) fill their own region of
// each digit, so each cursor takes an eighth of the atomics and the runs that
// meet in a cache line come from one L2.  The margins widen with the sample's
// measured line clumping (sorted or grouped keys); a layout too large for the
// buffers raises kErrCapacity before any tuple moves, and the first-pass
// scatter then moves nothing.  A region whose estimate proves too small takes
// no tuple past its end and raises kErrCapacity too; either way the caller
// redoes the join with the exact partition.  Every region starts on a
// 128-byte line, and the final partitions are [part_start[p], part_end[p])
// with gaps between them.  The work runs in two phases (kGapPrepare:
// sample, reduce, layout; kGapScatter: both passes and the ends) so that a
// join can lay out both relations before scattering either.
using tile::gap_cap;
using tile::gap_kappa;
using tile::line_same_counts;
using tile::gap_line;
using tile::kGapClumpEvery;
using tile::kGapGroup;
using tile::kGapKappaMax;
using tile::kGapLinesPerTile;
using tile::kGapLineTuples;
using tile::kGapPad;
using tile::kGapSampleNT;
using tile::kGapSampleUnroll;
using tile::kGapZ;
constexpr uint32_t kGapSampleWgs = 256;  // sample workgroups (rows of counts)
constexpr uint32_t kGapLayoutNT = 1024;
// first-pass regions per high digit: one per XCD range of tiles
constexpr uint32_t kGapRanges = kXcds;

// The XCD range of first-pass tile t in a grid of Tg (a multiple of 8) tiles:
// k_tile_scatter's blockIdx -> tile mapping (xcd_local_index) gives range x
// the tiles [x Tg/8, (x+1) Tg/8).
__device__ __forceinline__ uint32_t gap_range(uint32_t t, uint32_t Tg) { return t / (Tg / kGapRanges); }

// Sampled counts, one row per workgroup (tiles b, b + grid, ...: every
// workgroup samples the whole relation's spread) of gap_cols columns: 2^TB
// combined-digit bins; the high-digit counts of each XCD range, [range][high
// digit]; per XCD range the sum of same-digit counts of the high digit (for
// each measured tuple, the tuples of its line with its high digit, itself
// included: the clumping that gap_kappa reads) and the number of lines
// measured; and the one sum of same-digit counts of the combined digit over
// all measured tuples.  The clumping is measured on one sampled line in
// kGapClumpEvery: wave w of a workgroup measures the lines it holds of the
// u = w-th of the kGapSampleUnroll tiles the workgroup loads at a time, so
// each wave measures one tile in kGapSampleUnroll.
__host__ __device__ constexpr uint32_t gap_cols(uint32_t FT, uint32_t F1) {
  return FT + kGapRanges * F1 + 2 * kGapRanges + 1;
}

template <int TB, int BH>
__global__ __launch_bounds__(kGapSampleNT) void k_gap_sample(const uint4 *__restrict__ in4, uint64_t npairs,
                                                             uint32_t T, uint32_t Tg, uint32_t *__restrict__ rows) {
  constexpr uint32_t FT = 1u << TB, F1 = 1u << BH, BL = TB - BH;
  constexpr uint32_t COLS = gap_cols(FT, F1), QX = FT + kGapRanges * F1;
  static_assert(FT <= 0x8000u, "digits packed two to a word, 0xffff free");
  static_assert(kGapSampleNT / kWave == kGapSampleUnroll && kGapClumpEvery == kGapSampleUnroll,
                "one wave per tile of an unrolled step measures the clumping");
  __shared__ uint32_t hist[COLS];
  for (uint32_t b = threadIdx.x; b < COLS; b += kGapSampleNT) hist[b] = 0;
  __syncthreads();
  const uint32_t j = threadIdx.x >> 3, q = threadIdx.x & 7;
  uint32_t qb = 0;  // same-digit counts of the combined digit
  for (uint32_t t0 = blockIdx.x; t0 < T; t0 += gridDim.x * kGapSampleUnroll) {
    uint4 v[kGapSampleUnroll];
    bool ok[kGapSampleUnroll];
#pragma unroll
    for (int u = 0; u < kGapSampleUnroll; ++u) {
      const uint32_t t = t0 + (uint32_t)u * gridDim.x;
      const uint64_t p = (uint64_t)t * (kFastTile / 2) + (uint64_t)gap_line(t, j) * (kGapLineTuples / 2) + q;
      ok[u] = t < T && p < npairs;
      v[u] = ok[u] ? ld_stream(in4 + p) : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int u = 0; u < kGapSampleUnroll; ++u) {
      const uint32_t b0 = v[u].x & (FT - 1), b1 = v[u].z & (FT - 1);
      const uint32_t h0 = b0 >> BL, h1 = b1 >> BL;  // the high digit: the combined one's top BH bits
      const uint32_t t = t0 + (uint32_t)u * gridDim.x;
      if (u == (int)(threadIdx.x / kWave)) {  // wave-uniform: this wave's tile of the step
        // each line's 8 lanes compare their digits (all lanes of the wave take part)
        uint32_t cb0, cb1, ch0, ch1;
        line_same_counts<BL>(ok[u] ? b0 | b1 << 16 : ~0u, cb0, cb1, ch0, ch1);
        const uint32_t qh = wave_sum(ok[u] ? ch0 + ch1 : 0u);  // the tile's range is the wave's
        const uint32_t lines = (uint32_t)__popcll(__ballot(ok[u] && q == 0));  // a line's first pair is in
        if (t < T && (threadIdx.x & (kWave - 1)) == 0) {
          atomicAdd(&hist[QX + gap_range(t, Tg)], qh);
          atomicAdd(&hist[QX + kGapRanges + gap_range(t, Tg)], lines);
        }
        if (ok[u]) qb += cb0 + cb1;
      }
      if (!ok[u]) continue;
      const uint32_t x = gap_range(t, Tg) * F1;
      atomicAdd(&hist[b0], 1u);
      atomicAdd(&hist[b1], 1u);
      atomicAdd(&hist[FT + x + h0], 1u);
      atomicAdd(&hist[FT + x + h1], 1u);
    }
  }
  qb = wave_sum(qb);
  if ((threadIdx.x & (kWave - 1)) == 0) atomicAdd(&hist[COLS - 1], qb);
  __syncthreads();
  uint32_t *row = rows + (uint64_t)blockIdx.x * COLS;
  for (uint32_t b = threadIdx.x; b < COLS; b += kGapSampleNT) row[b] = hist[b];
}

// hs[b] = sum of the G rows' counts of column b: 64 columns per workgroup,
// its four waves each summing a quarter of the rows (reads coalesced along
// the columns).
__global__ __launch_bounds__(256) void k_gap_reduce(const uint32_t *__restrict__ rows, uint32_t G, uint32_t cols,
                                                    uint32_t *__restrict__ hs) {
  __shared__ uint32_t part[4][64];
  const uint32_t b = blockIdx.x * 64 + (threadIdx.x & 63), qw = threadIdx.x >> 6;
  uint32_t sum = 0;
  if (b < cols) {
    const uint32_t g0 = qw * ((G + 3) / 4), g1 = min(G, g0 + (G + 3) / 4);
    uint32_t g = g0;
    for (; g + 8 <= g1; g += 8) {
      uint32_t c[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) c[u] = rows[(uint64_t)(g + u) * cols + b];
#pragma unroll
      for (int u = 0; u < 8; ++u) sum += c[u];
    }
    for (; g < g1; ++g) sum += rows[(uint64_t)g * cols + b];
  }
  part[qw][threadIdx.x & 63] = sum;
  __syncthreads();
  if (threadIdx.x < 64 && b < cols)
    hs[b] = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
}

// One workgroup: the layout of both passes from the sampled counts hs.
// Final partition b: [part_start[b], re2[b]), cursor cur2[b] = part_start[b].
// First-pass region of (high digit h, range x), index k = x F1 + h: [rs1[k],
// re1[k]), cursor cur1[k], laid out in (h, x) order so that the second pass
// reads each digit's regions in a row.  Each estimate's margin is scaled by
// its clumping (gap_kappa): per XCD range for the first pass (a contiguous
// part of the input, which may be sorted where the rest is not), one for the
// whole relation for the final partitions.  The host sized
// the buffers (budget1, budget2 tuples) for any sample of unclumped input; a
// total beyond them (clumped input, or an empty sample) raises kErrCapacity
// before any tuple moves, and every region is clipped to its buffer.
template <int BH, int BL>
__global__ __launch_bounds__(kGapLayoutNT) void k_gap_layout(const uint32_t *__restrict__ hs, uint64_t n,
                                                              uint32_t T, uint32_t Tg, uint64_t budget1,
                                                              uint64_t budget2, uint32_t *__restrict__ rs1,
                                                              uint32_t *__restrict__ re1, uint32_t *__restrict__ cur1,
                                                              uint32_t *__restrict__ part_start,
                                                              uint32_t *__restrict__ re2,
                                                              uint32_t *__restrict__ cur2, uint32_t *err) {
  constexpr uint32_t FT = 1u << (BH + BL), F1 = 1u << BH;
  constexpr uint32_t BPT = FT >= kGapLayoutNT ? FT / kGapLayoutNT : 1;  // consecutive bins per thread
  constexpr uint32_t NSEG = F1 * kGapRanges, QX = FT + NSEG, QM = QX + kGapRanges, QB = gap_cols(FT, F1) - 1;
  static_assert(NSEG <= kGapLayoutNT, "one first-pass region per thread");
  __shared__ uint64_t red[kGapLayoutNT / kWave + 1];
  const uint32_t b0 = threadIdx.x * BPT;
  uint32_t sv[BPT];
  uint64_t loc = 0;
  if constexpr (BPT % 4 == 0) {  // whole 16-byte vectors of consecutive bins
#pragma unroll
    for (uint32_t k = 0; k < BPT; k += 4) {
      const uint4 v = *reinterpret_cast<const uint4 *>(hs + b0 + k);
      sv[k] = v.x;
      sv[k + 1] = v.y;
      sv[k + 2] = v.z;
      sv[k + 3] = v.w;
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < BPT; ++k) sv[k] = b0 + k < FT ? hs[b0 + k] : 0u;
  }
  uint64_t sq = 0;  // sum of squared counts (< ns^2 <= 2^56)
#pragma unroll
  for (uint32_t k = 0; k < BPT; ++k) {
    loc += sv[k];
    sq += (uint64_t)sv[k] * sv[k];
  }
  const uint64_t ns = block_sum<kGapLayoutNT, uint64_t>(loc, red);
  const uint64_t ns2 = block_sum<kGapLayoutNT, uint64_t>(sq, red);
  const double f = ns ? (double)n / (double)ns : 0.0;
  if (ns == 0 && threadIdx.x == 0) atomic_or_agent(err, kErrCapacity);
  // groups of lines the sample draws from: 64 per tile, and in the last
  // (partial) tile those of its tail tuples (a group counted with no tuple
  // only makes gap_kappa's correction smaller)
  const uint64_t tail = n - (uint64_t)(T - 1) * kFastTile;
  const uint32_t glast = (uint32_t)((tail + kGapLineTuples * kGapGroup - 1) / (kGapLineTuples * kGapGroup));
  // final partitions: one kappa for all of them, gap_kappa's with q and s^2 /
  // g summed over the bins (every bin draws on all the groups)
  // the clumping sums were measured on some of the lines: scaled by lines
  // sampled over lines measured (the ranges' line counts of the one in
  // kGapClumpEvery)
  const double gall = (double)(T - 1) * kGapLinesPerTile + glast;
  uint32_t mall = 0;
#pragma unroll
  for (uint32_t r = 0; r < kGapRanges; ++r) mall += hs[QM + r];
  const float kb = mall ? gap_kappa((double)ns, (double)ns2, (double)hs[QB] * gall / mall, gall) : 1.0f;
  uint64_t c2[BPT];
  uint64_t sum2 = 0;
#pragma unroll
  for (uint32_t k = 0; k < BPT; ++k) {
    c2[k] = (ns && b0 + k < FT) ? gap_cap(sv[k], kb, f) : 0;
    sum2 += c2[k];
  }
  uint64_t tot2;
  uint64_t at = block_excl_scan<kGapLayoutNT, uint64_t>(sum2, red, &tot2);
  if (tot2 > budget2 && threadIdx.x == 0) atomic_or_agent(err, kErrCapacity);
  uint32_t va[BPT], ve[BPT];
#pragma unroll
  for (uint32_t k = 0; k < BPT; ++k) {
    va[k] = (uint32_t)min(at, budget2);
    ve[k] = (uint32_t)min(at + c2[k], budget2);
    at += c2[k];
  }
  if constexpr (BPT % 4 == 0) {
#pragma unroll
    for (uint32_t k = 0; k < BPT; k += 4) {
      const uint4 a4 = make_uint4(va[k], va[k + 1], va[k + 2], va[k + 3]);
      *reinterpret_cast<uint4 *>(part_start + b0 + k) = a4;
      *reinterpret_cast<uint4 *>(cur2 + b0 + k) = a4;
      *reinterpret_cast<uint4 *>(re2 + b0 + k) = make_uint4(ve[k], ve[k + 1], ve[k + 2], ve[k + 3]);
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < BPT; ++k) {
      if (b0 + k < FT) {
        part_start[b0 + k] = va[k];
        cur2[b0 + k] = va[k];
        re2[b0 + k] = ve[k];
      }
    }
  }
  if (threadIdx.x == 0) part_start[FT] = (uint32_t)min(tot2, budget2);
  // first-pass regions: thread i = h * 8 + x
  const uint32_t i = threadIdx.x, h = i / kGapRanges, x = i % kGapRanges;
  const uint32_t s1 = i < NSEG ? hs[FT + x * F1 + h] : 0u;
  // per range x, the sum of its regions' counts and of their squares: the
  // lanes of one x are 8 apart in a wave, so xor-shuffles by 8, 16 and 32
  // leave lane x with the wave's sums, added to LDS once per range per wave
  __shared__ unsigned long long xs[kGapRanges], xs2[kGapRanges];
  if (threadIdx.x < kGapRanges) xs[threadIdx.x] = xs2[threadIdx.x] = 0;
  __syncthreads();
  unsigned long long a = s1, a2 = (unsigned long long)s1 * s1;  // < 2^56 each, < 2^63 over 128 digits
#pragma unroll
  for (int m = kGapRanges; m < kWave; m <<= 1) {
    a += __shfl_xor(a, m);
    a2 += __shfl_xor(a2, m);
  }
  if ((threadIdx.x & (kWave - 1)) < kGapRanges && (threadIdx.x & ~(kWave - 1)) < NSEG) {
    atomicAdd(&xs[threadIdx.x & (kWave - 1)], a);
    atomicAdd(&xs2[threadIdx.x & (kWave - 1)], a2);
  }
  __syncthreads();
  uint64_t c1 = 0;
  if (i < NSEG && ns) {
    // range x holds tiles [x Tg/8, (x + 1) Tg/8) of the T
    const uint32_t per = Tg / kGapRanges, lo = x * per;
    const uint32_t tiles = T > lo ? min(per, T - lo) : 0u;
    const uint32_t g = tiles * kGapLinesPerTile - ((T - 1) / per == x ? kGapLinesPerTile - glast : 0u);
    const uint32_t m = hs[QM + x];
    const float kx = m ? gap_kappa((double)xs[x], (double)xs2[x], (double)hs[QX + x] * g / m, (double)g) : 1.0f;
    c1 = gap_cap(s1, kx, f);
  }
  uint64_t tot1;
  const uint64_t a1 = block_excl_scan<kGapLayoutNT, uint64_t>(c1, red, &tot1);
  if (tot1 > budget1 && threadIdx.x == 0) atomic_or_agent(err, kErrCapacity);
  if (i < NSEG) {
    const uint32_t k = x * F1 + h;
    rs1[k] = (uint32_t)min(a1, budget1);
    cur1[k] = (uint32_t)min(a1, budget1);
    re1[k] = (uint32_t)min(a1 + c1, budget1);
  }
}

// Second-pass queues: the segments of high digit h go to queue h % Q (Q =
// min(8, F1)), whose groups a workgroup on XCD q takes first, so that the runs
// of one final partition are written from one XCD (their partial lines merge
// in its L2); a workgroup whose queue is empty takes groups from the others.
// Segment i in queue order: i = ((h % Q) (F1 / Q) + h / Q) 8 + x, for
// first-pass region (h, x), k = x F1 + h.
__host__ __device__ __forceinline__ uint32_t gap_queues(uint32_t F1) { return F1 < kXcds ? F1 : kXcds; }
__device__ __forceinline__ uint32_t gap_seg_digit(uint32_t i, uint32_t F1) {
  const uint32_t Q = gap_queues(F1), r = i / kGapRanges;  // r = (h % Q) (F1 / Q) + h / Q
  return (r % (F1 / Q)) * Q + r / (F1 / Q);
}

// Second-pass work items: segment i (the first-pass region's filled part
// [rs1[k], min(cur1[k], re1[k]))) splits into about len / L groups; grp
// receives their exclusive prefix over the segments in queue order
// (grp[nseg] = total), and grp[nseg + 1 + q], the ticket of queue q, is
// cleared.  Also raises kErrCapacity for a region that overflowed.
__global__ __launch_bounds__(kGapLayoutNT) void k_gap_groups(const uint32_t *__restrict__ rs1,
                                                             const uint32_t *__restrict__ cur1,
                                                             const uint32_t *__restrict__ re1, uint32_t F1,
                                                             uint64_t L, uint32_t *__restrict__ grp, uint32_t *err) {
  __shared__ uint32_t red[kGapLayoutNT / kWave + 1];
  const uint32_t i = threadIdx.x, nseg = F1 * kGapRanges;
  uint32_t J = 0;
  if (i < nseg) {
    const uint32_t k = (i % kGapRanges) * F1 + gap_seg_digit(i, F1);
    if (cur1[k] > re1[k]) atomic_or_agent(err, kErrCapacity);
    const uint64_t len = (uint64_t)min(cur1[k], re1[k]) - rs1[k];
    const uint64_t j = (len + L / 2) / L;
    J = (uint32_t)(j < 1 ? 1 : (j > 65536 ? 65536 : j));
  }
  uint32_t tot;
  const uint32_t ex = block_excl_scan<kGapLayoutNT, uint32_t>(J, red, &tot);
  if (i < nseg) grp[i] = ex;
  if (i == 0) grp[nseg] = tot;
  if (i < kXcds) grp[nseg + 1 + i] = 0;
}

// Second pass into the final partitions, persistent like k_msd_scatter, with
// one work queue per XCD (above): group w of segment (h, x) is the j-th of its
// J equal parts (even bounds: 16-byte loads), scattered by the low digit with
// runs reserved from cur2[h F2 + d].
template <int BL>
__global__ __launch_bounds__(kFastNT) void k_gap_scatter(const uint2 *__restrict__ in,
                                                         const uint32_t *__restrict__ rs1,
                                                         const uint32_t *__restrict__ cur1,
                                                         const uint32_t *__restrict__ re1, uint32_t F1,
                                                         const uint32_t *__restrict__ grp, uint32_t *tickets,
                                                         const uint32_t *__restrict__ re2, uint32_t *cur2,
                                                         uint2 *__restrict__ out, uint32_t *err) {
  constexpr uint32_t F2 = 1u << BL;
  __shared__ tile::ScatterShared<BL, false> sh;
  __shared__ uint32_t s_w;
  const uint32_t nseg = F1 * kGapRanges, Q = gap_queues(F1), spq = nseg / Q;  // segments per queue
  const uint32_t home = (blockIdx.x % kXcds) % Q;  // the hardware deals workgroup b to XCD b % 8
  uint32_t qi = 0;                                 // queues visited: home, home + 1, ...
  for (;;) {
    const uint32_t q = (home + qi) % Q;
    const uint32_t qb = grp[q * spq], qe = grp[(q + 1) * spq];
    if (threadIdx.x == 0) {
      const uint32_t t = atomicAdd(tickets + q, 1u);
      s_w = t < qe - qb ? qb + t : ~0u;
    }
    __syncthreads();
    const uint32_t w = s_w;
    __syncthreads();  // s_w is written again by the next round
    if (w == ~0u) {   // uniform: this queue is drained, go on to the next one
      if (++qi == Q) break;
      continue;
    }
    const uint32_t i = wave_find(grp, nseg, w);  // grp[i] <= w < grp[i + 1]
    const uint32_t h = gap_seg_digit(i, F1), k = (i % kGapRanges) * F1 + h;
    const uint32_t J = grp[i + 1] - grp[i], j = w - grp[i];
    const uint64_t b0 = rs1[k];
    const uint64_t len = (uint64_t)min(cur1[k], re1[k]) - b0;
    const uint64_t beg = b0 + (((uint64_t)j * len / J) & ~1ull);
    const uint64_t end = j + 1 == J ? b0 + len : b0 + (((uint64_t)(j + 1) * len / J) & ~1ull);
    for (uint32_t d = threadIdx.x; d < F2; d += kFastNT) {
      sh.cursor[d] = re2[(uint64_t)h * F2 + d];
      sh.cnt[d] = 0;
    }
    __syncthreads();
    tile::scatter_stripe<BL, false, true>(sh, in, beg, end, 0, 0, 0u, out, err, cur2 + (uint64_t)h * F2);
    __syncthreads();  // sh is reused by the next group
  }
}

// part_end[b] = the filled end of final partition b.
__global__ void k_gap_ends(const uint32_t *__restrict__ cur2, const uint32_t *__restrict__ re2, uint32_t FT,
                           uint32_t *__restrict__ part_end) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < FT) part_end[b] = min(cur2[b], re2[b]);
}

// Tuples a relation of n needs in a buffer of P sampled regions (the first
// pass's 8 2^bh, the final 2^(bh+bl)) for any sample whose clumping ratios
// are at most K.  A region of sampled count s and ratio kappa gets f kappa m w^3
// with m = s / kappa + 1, and w <= 1 + z/(3 sqrt(m)) gives m w^3 <= m + z
// sqrt(m) + z^2/3 + z^3/27, so its capacity is at most f (s + kappa + z
// sqrt(kappa (s + kappa)) + kappa (z^2/3 + z^3/27)) + pad + line; over the P
// regions, with kappa <= K and sum(sqrt(s_b + K)) <= sqrt(P (ns + K P))
// (Cauchy-Schwarz), at most f (ns + K P (1 + z^2/3 + z^3/27) + z sqrt(K P (ns
// + K P))) + P (pad + line + 1), and f ns = n; f = n / ns falls as the sample
// grows, so the smallest sample (full tiles only) bounds it.  A layout whose
// ratios pass K may not fit: it raises kErrCapacity before any tuple moves.
uint64_t gap_bound(uint64_t n, uint64_t P, double K) {
  const uint64_t full = n / kFastTile;
  const double ns = (double)full * kGapLinesPerTile * kGapLineTuples;
  if (ns <= 0.0) return ~0ull;
  const double f = (double)n / ns, z = kGapZ, p = (double)P;
  const double extra = f * (K * p * (1.0 + z * z / 3.0 + z * z * z / 27.0) + z * std::sqrt(K * p * (ns + K * p)));
  return n + (uint64_t)std::ceil(extra) + P * (kGapPad + kGapLineTuples + 2);
}

struct GapLayout {
  size_t o_rows, o_hs, o_re2, o_cur2, o_rs1, o_re1, o_cur1, o_grp, total;
  uint32_t G, cols;
};

GapLayout gap_layout(uint64_t n, int bh, int bl) {
  GapLayout L{};
  const uint64_t FT = 1ull << (bh + bl), NSEG = (1ull << bh) * kGapRanges;
  const uint64_t T = (n + kFastTile - 1) / kFastTile;
  L.G = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(T, kGapSampleWgs));
  L.cols = gap_cols((uint32_t)FT, (uint32_t)(1u << bh));
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  L.o_rows = 0;
  L.o_hs = L.o_rows + al((size_t)L.G * L.cols * 4);
  L.o_re2 = L.o_hs + al(L.cols * 4);
  L.o_cur2 = L.o_re2 + al(FT * 4);
  L.o_rs1 = L.o_cur2 + al(FT * 4);
  L.o_re1 = L.o_rs1 + al(NSEG * 4);
  L.o_cur1 = L.o_re1 + al(NSEG * 4);
  L.o_grp = L.o_cur1 + al(NSEG * 4);
  L.total = L.o_grp + al((NSEG + 1 + kXcds) * 4);
  return L;
}

template <int BH, int BL>
hipError_t launch_gap(uint64_t n, uint64_t group_len, uint32_t workers, const eo_tuple *in, eo_tuple *tmp,
                      eo_tuple *out, uint32_t *part_start, uint32_t *part_end, void *scratch, uint32_t *err,
                      unsigned phases, hipStream_t stream) {
  constexpr uint32_t F1 = 1u << BH, FT = 1u << (BH + BL), NSEG = F1 * kGapRanges;
  const GapLayout L = gap_layout(n, BH, BL);
  char *sc = (char *)scratch;
  uint32_t *rows = (uint32_t *)(sc + L.o_rows), *hs = (uint32_t *)(sc + L.o_hs);
  uint32_t *re2 = (uint32_t *)(sc + L.o_re2), *cur2 = (uint32_t *)(sc + L.o_cur2);
  uint32_t *rs1 = (uint32_t *)(sc + L.o_rs1), *re1 = (uint32_t *)(sc + L.o_re1), *cur1 = (uint32_t *)(sc + L.o_cur1);
  uint32_t *grp = (uint32_t *)(sc + L.o_grp);
  const uint32_t T = (uint32_t)((n + kFastTile - 1) / kFastTile);
  // whole rounds over the XCDs (xcd_local_index is a bijection only on
  // multiples of 8); the blocks past the last tile return at once
  const uint32_t Tg = (T + kXcds - 1) / kXcds * kXcds;
  hipError_t e;
  if (phases & kGapPrepare) {
    EO_TIMED("part_sample", stream,
             hipLaunchKernelGGL((k_gap_sample<BH + BL, BH>), dim3(L.G), dim3(kGapSampleNT), 0, stream,
                                reinterpret_cast<const uint4 *>(in), n / 2, T, Tg, rows));
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_gap_reduce, dim3((L.cols + 63) / 64), dim3(256), 0, stream, rows, L.G, L.cols, hs);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL((k_gap_layout<BH, BL>), dim3(1), dim3(kGapLayoutNT), 0, stream, hs, n, T, Tg,
                       gap_bound(n, NSEG, kGapKappaMax), gap_bound(n, FT, kGapKappaMax), rs1, re1, cur1, part_start, re2,
                       cur2, err);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (!(phases & kGapScatter)) return hipSuccess;
  EO_TIMED("part_scatter", stream,
           hipLaunchKernelGGL((k_tile_scatter<BH, false, true>), dim3(Tg), dim3(kFastNT), 0, stream,
                              reinterpret_cast<const uint2 *>(in), n, BL, re1, reinterpret_cast<uint2 *>(tmp), err,
                              cur1));
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_gap_groups, dim3(1), dim3(kGapLayoutNT), 0, stream, rs1, cur1, re1, F1, group_len, grp, err);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const uint64_t max_groups = NSEG + n / group_len + 1;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(max_groups, std::max<uint32_t>(workers, 1));
  EO_TIMED("seg_scatter", stream,
           hipLaunchKernelGGL((k_gap_scatter<BL>), dim3(grid), dim3(kFastNT), 0, stream,
                              reinterpret_cast<const uint2 *>(tmp), rs1, cur1, re1, F1, grp, grp + NSEG + 1, re2,
                              cur2, reinterpret_cast<uint2 *>(out), err));
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_gap_ends, dim3((FT + 255) / 256), dim3(256), 0, stream, cur2, re2, FT, part_end);
  return hipGetLastError();
}

uint32_t tile_of_bits(int bits) {
  switch (bits) {
#define EO_TILE_CASE(B) \
  case B:               \
    return PartCfg<B>::TILE;
    EO_TILE_CASE(1) EO_TILE_CASE(2) EO_TILE_CASE(3) EO_TILE_CASE(4) EO_TILE_CASE(5) EO_TILE_CASE(6)
    EO_TILE_CASE(7) EO_TILE_CASE(8) EO_TILE_CASE(9) EO_TILE_CASE(10) EO_TILE_CASE(11)
#undef EO_TILE_CASE
    default:
      return 0;
  }
}

}  // namespace

PartPassPlan plan_partition_pass(uint64_t n, int bits, int shift, uint32_t xor_mask, int cus, bool fast) {
  PartPassPlan p{};
  p.n = n;
  p.bits = bits;
  p.shift = shift;
  p.xor_mask = xor_mask;
  p.fast = fast;
  cus = std::max(cus, kXcds);
  const uint64_t tile = fast ? (uint64_t)kFastTile : tile_of_bits(bits);
  const uint64_t tiles = std::max<uint64_t>(1, (n + tile - 1) / tile);
  // fast: kFastWgPerCu 64 KB-LDS workgroups per CU, one stripe each; stable: two per CU.
  uint64_t g = std::min<uint64_t>(tiles, (uint64_t)cus * (fast ? tile::kFastWgPerCu : 2));
  g = (g + kXcds - 1) / kXcds * kXcds;
  uint64_t len = (n + g - 1) / g;
  len = std::max<uint64_t>(tile, (len + 1) & ~1ull);
  if (!fast) len = (len + tile - 1) / tile * tile;
  p.stripes = (uint32_t)g;
  p.stripe_len = len;
  // the histogram reads with ~4 workgroups per CU
  p.hist_split = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(8, (uint64_t)cus * 4 / g));
  return p;
}

size_t partition_pass_scratch_bytes(const PartPassPlan &p) {
  const uint64_t m = ((uint64_t)p.stripes * p.hist_split) << p.bits;
  const size_t counts = (m * sizeof(uint32_t) + 255) / 256 * 256;
  return counts + scan_scratch_bytes_u32(m);
}

hipError_t partition_pass(const PartPassPlan &p, const eo_tuple *in, eo_tuple *out, uint32_t *part_start,
                          void *scratch, uint32_t *err, hipStream_t stream) {
  const uint64_t m = ((uint64_t)p.stripes * p.hist_split) << p.bits;
  uint32_t *counts = (uint32_t *)scratch;
  void *scan_scratch = (char *)scratch + (m * sizeof(uint32_t) + 255) / 256 * 256;
  hipError_t e;
  switch (p.bits) {
#define EO_PASS_CASE(B)                                                    \
  case B:                                                                  \
    e = launch_pass<B>(p, in, out, counts, scan_scratch, err, stream); \
    break;
    EO_PASS_CASE(1) EO_PASS_CASE(2) EO_PASS_CASE(3) EO_PASS_CASE(4) EO_PASS_CASE(5) EO_PASS_CASE(6)
    EO_PASS_CASE(7) EO_PASS_CASE(8) EO_PASS_CASE(9) EO_PASS_CASE(10) EO_PASS_CASE(11)
#undef EO_PASS_CASE
    default:
      return hipErrorInvalidValue;
  }
  if (e != hipSuccess || !part_start) return e;
  const uint32_t F = 1u << p.bits;
  hipLaunchKernelGGL(k_part_starts, dim3((F + 1 + 255) / 256), dim3(256), 0, stream, counts, p.stripes,
                     p.hist_split, F, p.n, part_start);
  return hipGetLastError();
}

hipError_t partition_bounds(const eo_tuple *sorted, uint64_t n, int bits, int shift, uint32_t xor_mask,
                            uint32_t *part_start, hipStream_t stream) {
  const uint32_t F = 1u << bits;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 1 + 255) / 256, 8192);
  hipLaunchKernelGGL(k_part_bounds, dim3(blocks), dim3(256), 0, stream, reinterpret_cast<const uint2 *>(sorted), n,
                     shift, xor_mask, F, part_start);
  return hipGetLastError();
}

SegPassPlan plan_seg_pass(uint64_t n, uint32_t nseg, int bits, int shift, int cus) {
  SegPassPlan p{};
  p.n = n;
  p.nseg = nseg;
  p.bits = bits;
  p.shift = shift;
  // stripes of ~n / (2 CUs) tuples, at least one fast tile
  uint64_t len = std::max<uint64_t>(kFastTile, n / std::max<uint64_t>(1, (uint64_t)std::max(cus, 8) * 2));
  len = (len + kFastTile - 1) / kFastTile * kFastTile;
  p.stripe_len = len;
  p.max_stripes = (uint32_t)(nseg + (n + len - 1) / len);
  return p;
}

size_t seg_pass_scratch_bytes(const SegPassPlan &p) {
  const uint64_t m = (uint64_t)p.max_stripes << p.bits;
  return ((p.nseg + 1) * sizeof(uint32_t) + 255) / 256 * 256 + (m * sizeof(uint32_t) + 255) / 256 * 256 +
         std::max(scan_scratch_bytes_u32(m), scan_scratch_bytes_u32(p.nseg + 1));
}

hipError_t partition_seg_pass(const SegPassPlan &p, const eo_tuple *in, eo_tuple *out, const uint32_t *seg_start,
                              uint32_t *part_start, void *scratch, uint32_t *err, hipStream_t stream) {
  if (((uintptr_t)in & 15) || ((uintptr_t)out & 7)) return hipErrorInvalidValue;
  switch (p.bits) {
#define EO_SEG_CASE(B) \
  case B:              \
    return launch_seg_pass<B>(p, in, out, seg_start, part_start, scratch, err, stream);
    EO_SEG_CASE(1) EO_SEG_CASE(2) EO_SEG_CASE(3) EO_SEG_CASE(4) EO_SEG_CASE(5) EO_SEG_CASE(6)
    EO_SEG_CASE(7) EO_SEG_CASE(8) EO_SEG_CASE(9) EO_SEG_CASE(10) EO_SEG_CASE(11)
#undef EO_SEG_CASE
    default:
      return hipErrorInvalidValue;
  }
}


bool msd2_supported(int bits_hi, int bits_lo) {
  return bits_hi >= 1 && bits_lo >= 1 && bits_hi <= 7 && bits_lo <= 7 && bits_hi + bits_lo <= kMsd2MaxBits;
}

uint64_t msd2_group_len(uint64_t n, int cus) {
  // kMsdGroupsPerWg groups per persistent second-pass workgroup (kFastWgPerCu
  // per CU): small enough that the dynamic hand-out evens out segments of
  // unequal size, large enough that a group spans several tiles
  const uint64_t len = n / ((uint64_t)std::max(cus, 8) * tile::kFastWgPerCu * kMsdGroupsPerWg);
  return std::max<uint64_t>(len, (uint64_t)kFastTile);
}

size_t msd2_scratch_bytes(uint64_t n, int bh, int bl, int cus) {
  return msd2_layout(msd2_geo(n, cus), bh, bl).total;
}

hipError_t partition_msd2(uint64_t n, int bh, int bl, const eo_tuple *in, eo_tuple *tmp, eo_tuple *out,
                          uint32_t *part_start, void *scratch, uint32_t *err, int cus, hipStream_t stream,
                          bool ordered) {
  if (!msd2_supported(bh, bl)) return hipErrorInvalidValue;
  if (((uintptr_t)in & 15) || ((uintptr_t)tmp & 15) || ((uintptr_t)out & 7)) return hipErrorInvalidValue;
  const Msd2Geo geo = msd2_geo(n, cus);
  const uint64_t L = msd2_group_len(n, cus);
  const uint32_t workers = (uint32_t)std::max(cus, 8) * tile::kFastWgPerCu;
  switch (bh * 16 + bl) {
#define EO_MSD_CASE(BH, BL) \
  case BH * 16 + BL:        \
    return ordered ? launch_msd2<BH, BL, true>(n, geo, L, workers, in, tmp, out, part_start, scratch, err, stream) \
                   : launch_msd2<BH, BL, false>(n, geo, L, workers, in, tmp, out, part_start, scratch, err, stream);
#define EO_MSD_ROW(BH) \
  EO_MSD_CASE(BH, 1) EO_MSD_CASE(BH, 2) EO_MSD_CASE(BH, 3) EO_MSD_CASE(BH, 4) EO_MSD_CASE(BH, 5) EO_MSD_CASE(BH, 6) \
  EO_MSD_CASE(BH, 7)
    EO_MSD_ROW(1) EO_MSD_ROW(2) EO_MSD_ROW(3) EO_MSD_ROW(4) EO_MSD_ROW(5) EO_MSD_ROW(6) EO_MSD_ROW(7)
#undef EO_MSD_ROW
#undef EO_MSD_CASE
    default:
      return hipErrorInvalidValue;
  }
}

// Second-pass group length of the sampled partition: about 16 groups per
// persistent workgroup, so that the last groups to finish are short
// (measured per C2 launch: 4 groups per workgroup 0.666 ms, 16: 0.649, 32:
// 0.664, 64: 0.697; at Zipf 0.9 4: 0.705, 16: 0.650).
uint64_t gap_group_len(uint64_t n, int cus) {
  const uint64_t len = n / ((uint64_t)std::max(cus, 8) * tile::kFastWgPerCu * kGapGroupsPerWg);
  return std::max<uint64_t>(len, (uint64_t)kFastTile);
}

bool gap_supported(uint64_t n, int bits_hi, int bits_lo) {
  if (!msd2_supported(bits_hi, bits_lo) || bits_hi + bits_lo < 2 || n < kGapMinRows) return false;
  // the 32-bit region cursors start below the buffer size and grow by at
  // most n, so they never wrap
  return gap_bound(n, 1ull << (bits_hi + bits_lo), kGapKappaMax) + n < (1ull << 32) &&
         gap_bound(n, (1ull << bits_hi) * kGapRanges, kGapKappaMax) + n < (1ull << 32);
}

uint64_t gap_rows(uint64_t n, uint64_t parts) { return gap_bound(n, parts, kGapKappaMax); }

uint64_t gap_tmp_rows(uint64_t n, int bits_hi, int /*bits_lo*/) {
  return gap_bound(n, (1ull << bits_hi) * kGapRanges, kGapKappaMax);
}

uint64_t gap_out_rows(uint64_t n, int bits_hi, int bits_lo) {
  return gap_bound(n, 1ull << (bits_hi + bits_lo), kGapKappaMax);
}

size_t gap_scratch_bytes(uint64_t n, int bits_hi, int bits_lo) { return gap_layout(n, bits_hi, bits_lo).total; }

hipError_t partition_gapped(uint64_t n, int bh, int bl, const eo_tuple *in, eo_tuple *tmp, eo_tuple *out,
                            uint32_t *part_start, uint32_t *part_end, void *scratch, uint32_t *err, int cus,
                            hipStream_t stream, unsigned phases) {
  if (!phases || (phases & ~(kGapPrepare | kGapScatter))) return hipErrorInvalidValue;
  if (!gap_supported(n, bh, bl)) return hipErrorInvalidValue;
  if (((uintptr_t)in & 15) || ((uintptr_t)tmp & 15) || ((uintptr_t)out & 15)) return hipErrorInvalidValue;
  const uint64_t L = gap_group_len(n, cus);
  const uint32_t workers = (uint32_t)std::max(cus, 8) * tile::kFastWgPerCu;
  switch (bh * 16 + bl) {
#define EO_GAP_CASE(BH, BL) \
  case BH * 16 + BL:        \
    return launch_gap<BH, BL>(n, L, workers, in, tmp, out, part_start, part_end, scratch, err, phases, stream);
#define EO_GAP_ROW(BH) \
  EO_GAP_CASE(BH, 1) EO_GAP_CASE(BH, 2) EO_GAP_CASE(BH, 3) EO_GAP_CASE(BH, 4) EO_GAP_CASE(BH, 5) EO_GAP_CASE(BH, 6) \
  EO_GAP_CASE(BH, 7)
    EO_GAP_ROW(1) EO_GAP_ROW(2) EO_GAP_ROW(3) EO_GAP_ROW(4) EO_GAP_ROW(5) EO_GAP_ROW(6) EO_GAP_ROW(7)
#undef EO_GAP_ROW
#undef EO_GAP_CASE
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace eo
