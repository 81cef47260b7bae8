// Per-partition join of a radix-partitioned R and S.
//
// The reference's probe (count / join, joins/hashJoin_gpu/main_street/
// join_gpu.cu:9-113) holds a <=256-tuple S chunk in shared memory and compares
// every R tuple of the partition against all of it — a partitioned nested
// loop, run twice (count, then write).  Here each work tile
//   (partition p, R-chunk c of <= CAP tuples, S-range j of <= s_tile tuples)
// builds a bucketised table of its R chunk in LDS (LDS counting sort by a
// multiplicative hash: histogram with LDS atomics, workgroup scan, placement by
// atomic decrement), then streams its S range from HBM and probes the table.
// Output materialisation is count -> device scan -> write: the count kernel
// stores one total per tile, a u64 scan turns them into tile output offsets,
// and the write kernel re-probes and writes (R payload, S payload) pairs at
// tile offset + workgroup-scanned per-thread offsets.
//
// Tiles of one partition are numbered (j, c) with c fastest and dealt so that
// consecutive tiles share an XCD: the R-chunks probing one S range run side by
// side on one XCD and meet its S tuples in that XCD's L2.
#include <algorithm>

#include "device.hpp"
#include "kernels.hpp"
#include "tile_kernels.hpp"

namespace eo {
namespace {

template <int NT_, int CAP_, int LOG2NB_>
struct JoinCfg {
  static constexpr int NT = NT_;
  static constexpr int CAP = CAP_;
  static constexpr int LOG2NB = LOG2NB_;
  static constexpr int NB = 1 << LOG2NB_;
  static constexpr int RPT = CAP / NT;   // R tuples per thread while building
  static constexpr int BPT = NB / NT;    // buckets per thread in the bucket scan
  static constexpr int U = 4;            // S tuples per thread per probe step
};

using Cfg4K = JoinCfg<512, 4096, 10>;
using Cfg8K = JoinCfg<512, 8192, 11>;
using Cfg16K = JoinCfg<1024, 16384, 12>;

struct TileRef {
  uint32_t r_beg, r_n, s_beg, s_n;
};

// Locates tile t: p = last partition with tile_start[p] <= t (a wave-wide
// search: all threads of the workgroup call it), then (j, c).
template <int CAP>
__device__ __forceinline__ bool resolve_tile(uint32_t t, const uint32_t *tile_start, uint32_t P,
                                             const uint32_t *r_start, const uint32_t *s_start, uint32_t s_tile,
                                             TileRef *ref) {
  if (t >= tile_start[P]) return false;
  const uint32_t p = wave_find(tile_start, P, t);
  const uint32_t local = t - tile_start[p];
  const uint32_t rb = r_start[p], rn = r_start[p + 1] - rb;
  const uint32_t sb = s_start[p], sn = s_start[p + 1] - sb;
  const uint32_t nc = (rn + CAP - 1) / CAP;
  const uint32_t c = local % nc, j = local / nc;
  ref->r_beg = rb + c * CAP;
  ref->r_n = min((uint32_t)CAP, rn - c * CAP);
  ref->s_beg = sb + j * s_tile;
  ref->s_n = min(s_tile, sn - j * s_tile);
  return true;
}

// Builds the LDS table of R[rb, rb+rn): ent[] grouped by bucket, bucket b in
// ent[off[b], off[b+1]).  Ends with a barrier.
template <typename C>
__device__ __forceinline__ void build_table(const uint2 *__restrict__ R, uint32_t rb, uint32_t rn, uint2 *ent,
                                            uint32_t *off, uint32_t *scan_tmp) {
  for (int h = threadIdx.x; h <= C::NB; h += C::NT) off[h] = 0;
  __syncthreads();
  uint2 r[C::RPT];
  uint32_t hh[C::RPT];
#pragma unroll
  for (int k = 0; k < C::RPT; ++k) {
    const uint32_t i = k * C::NT + threadIdx.x;
    if (i < rn) {
      r[k] = R[rb + i];
      hh[k] = bucket_hash((int32_t)r[k].x, C::LOG2NB);
      atomicAdd(&off[hh[k]], 1u);
    }
  }
  __syncthreads();
  {  // inclusive scan of bucket sizes: off[b] = end of bucket b
    uint32_t loc[C::BPT];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < C::BPT; ++k) {
      loc[k] = off[threadIdx.x * C::BPT + k];
      sum += loc[k];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan<C::NT, uint32_t>(sum, scan_tmp, &tot);
#pragma unroll
    for (int k = 0; k < C::BPT; ++k) {
      run += loc[k];
      off[threadIdx.x * C::BPT + k] = run;
    }
    if (threadIdx.x == 0) off[C::NB] = rn;
  }
  __syncthreads();
  // Place every tuple at a slot fixed by the input alone, so that the count
  // and write passes, and every run, see one bucket order: lanes of a wave
  // that share a bucket are ranked by lane (match over the bucket bits), one
  // leader per group takes the group's slots with one atomic, and the waves
  // take turns.  Packed per tuple: rank in group | leader lane << 8 | group size << 16.
  const int w = threadIdx.x / kWave;
  const uint64_t lt = lanemask_lt();
  uint32_t grp[C::RPT];
#pragma unroll
  for (int k = 0; k < C::RPT; ++k) {
    const bool valid = k * C::NT + threadIdx.x < rn;
    const uint64_t peers = match_peers<C::LOG2NB>(valid ? hh[k] : 0u, __ballot(valid));
    const uint32_t lead = valid ? (uint32_t)(__ffsll((unsigned long long)peers) - 1) : (uint32_t)lane_id();
    grp[k] = (uint32_t)__popcll(peers & lt) | lead << 8 | (uint32_t)__popcll(peers) << 16;
  }
  for (int turn = 0; turn < C::NT / kWave; ++turn) {
    if (w == turn) {
#pragma unroll
      for (int k = 0; k < C::RPT; ++k) {
        const bool valid = k * C::NT + threadIdx.x < rn;
        const uint32_t before = grp[k] & 0xffu;
        uint32_t top = 0;
        if (valid && before == 0) top = atomicSub(&off[hh[k]], grp[k] >> 16);
        top = __shfl(top, (int)((grp[k] >> 8) & 0xffu));
        if (valid) ent[top - 1 - before] = r[k];
      }
    }
    __syncthreads();
  }
}

template <typename C>
__global__ __launch_bounds__(C::NT) void k_join_count(const uint2 *__restrict__ R, const uint32_t *__restrict__ r_start,
                                                      const uint2 *__restrict__ S, const uint32_t *__restrict__ s_start,
                                                      const uint32_t *__restrict__ tile_start, uint32_t P,
                                                      uint32_t s_tile, uint64_t *__restrict__ tile_count,
                                                      uint32_t *matched) {
  __shared__ uint2 ent[C::CAP];
  __shared__ uint32_t off[C::NB + 1];
  __shared__ uint32_t scan_tmp[C::NT / kWave + 1];
  __shared__ uint64_t red[C::NT / kWave + 1];
  const uint32_t t = xcd_local_index(blockIdx.x, gridDim.x);
  TileRef ref;
  if (!resolve_tile<C::CAP>(t, tile_start, P, r_start, s_start, s_tile, &ref)) return;
  build_table<C>(R, ref.r_beg, ref.r_n, ent, off, scan_tmp);

  uint32_t cnt = 0;
  for (uint32_t i0 = 0; i0 < ref.s_n; i0 += C::NT * C::U) {
    int32_t key[C::U];
#pragma unroll
    for (int u = 0; u < C::U; ++u) {
      const uint32_t i = i0 + u * C::NT + threadIdx.x;
      key[u] = i < ref.s_n ? (int32_t)S[ref.s_beg + i].x : 0;
    }
#pragma unroll
    for (int u = 0; u < C::U; ++u) {
      const uint32_t i = i0 + u * C::NT + threadIdx.x;
      uint32_t m = 0;
      if (i < ref.s_n) {
        const uint32_t h = bucket_hash(key[u], C::LOG2NB);
        const uint32_t e1 = off[h + 1];
        for (uint32_t e = off[h]; e < e1; ++e) m += (int32_t)ent[e].x == key[u];
      }
      cnt += m;
      if (matched)  // left outer join: S positions with a match in this tile
        mark_bits(matched, (uint64_t)ref.s_beg + i0 + u * C::NT + (threadIdx.x / kWave) * kWave, __ballot(m != 0));
    }
  }
  const uint64_t tot = block_sum<C::NT, uint64_t>((uint64_t)cnt, red);
  if (threadIdx.x == 0) tile_count[t] = tot;
}

template <typename C>
__global__ __launch_bounds__(C::NT) void k_join_write(const uint2 *__restrict__ R, const uint32_t *__restrict__ r_start,
                                                      const uint2 *__restrict__ S, const uint32_t *__restrict__ s_start,
                                                      const uint32_t *__restrict__ tile_start, uint32_t P,
                                                      uint32_t s_tile, const uint64_t *__restrict__ tile_off,
                                                      uint2 *__restrict__ out, uint64_t out_cap, int wide) {
  __shared__ uint2 ent[C::CAP];
  __shared__ uint32_t off[C::NB + 1];
  __shared__ uint32_t scan_tmp[C::NT / kWave + 1];
  __shared__ uint64_t red[C::NT / kWave + 1];
  const uint32_t t = xcd_local_index(blockIdx.x, gridDim.x);
  TileRef ref;
  if (!resolve_tile<C::CAP>(t, tile_start, P, r_start, s_start, s_tile, &ref)) return;
  build_table<C>(R, ref.r_beg, ref.r_n, ent, off, scan_tmp);

  uint64_t base = tile_off[t];
  for (uint32_t i0 = 0; i0 < ref.s_n; i0 += C::NT * C::U) {
    uint2 s[C::U];
    uint32_t b0[C::U], b1[C::U];
    uint32_t m = 0;
#pragma unroll
    for (int u = 0; u < C::U; ++u) {
      const uint32_t i = i0 + u * C::NT + threadIdx.x;
      s[u] = i < ref.s_n ? S[ref.s_beg + i] : make_uint2(0u, 0u);
      b0[u] = b1[u] = 0;
    }
#pragma unroll
    for (int u = 0; u < C::U; ++u) {
      const uint32_t i = i0 + u * C::NT + threadIdx.x;
      if (i < ref.s_n) {
        const uint32_t h = bucket_hash((int32_t)s[u].x, C::LOG2NB);
        b0[u] = off[h];
        b1[u] = off[h + 1];
        for (uint32_t e = b0[u]; e < b1[u]; ++e) m += ent[e].x == s[u].x;
      }
    }
    uint64_t step_total;
    uint64_t pos = base + block_excl_scan<C::NT, uint64_t>((uint64_t)m, red, &step_total);
#pragma unroll
    for (int u = 0; u < C::U; ++u) {
      for (uint32_t e = b0[u]; e < b1[u]; ++e) {
        const uint2 r = ent[e];
        if (r.x == s[u].x) {
          if (pos < out_cap) emit_pair(out, pos, r.y, s[u].y, s[u].x, wide != 0);
          ++pos;
        }
      }
    }
    base += step_total;
  }
}


// Fused probe, one pass over S.  A tile is (partition p, R-chunk c, S span j):
// the R chunk (<= 16384 tuples) is built once into an LDS table and the whole S
// span (up to s_span tuples) streams past it in batches of NT*SPT tuples held
// in VGPRs, the next batch's loads in flight while the current one probes.
//
// LDS table (SoA, 144 KB): keys[16384] and pays[16384] grouped by bucket,
// u16 bucket offsets off[8192 + 1].  Built in three steps without a cursor
// array: (1) every R tuple takes its rank inside its bucket from an LDS
// atomic on a u32 counter array that lives in the (still empty) keys region,
// (2) a workgroup scan turns the counters into u16 offsets, (3) each tuple is
// stored at off[h] + rank.
//
// Per batch, every S tuple is probed once (match count and first payload kept
// in registers), output positions are numbered register-major (so a wave's
// store instruction writes one contiguous run), and one 64-bit atomic per
// batch reserves the output range.  Pairs at positions >= out_cap are counted
// but not written.
using namespace tile;

__device__ __forceinline__ bool resolve_span(uint32_t t, const uint32_t *tile_start, uint32_t P,
                                             const uint32_t *r_start, const uint32_t *s_start, uint32_t s_span,
                                             uint32_t cap, SpanRef *ref) {
  if (t >= tile_start[P]) return false;
  const uint32_t p = wave_find(tile_start, P, t);
  const uint32_t local = t - tile_start[p];
  const uint32_t rb = r_start[p], rn = r_start[p + 1] - rb;
  const uint32_t sb = s_start[p], sn = s_start[p + 1] - sb;
  const uint32_t nc = (rn + cap - 1) / cap;
  const uint32_t c = local % nc, j = local / nc;
  ref->r_beg = rb + c * cap;
  ref->r_n = min(cap, rn - c * cap);
  ref->s_beg = sb + j * s_span;
  ref->s_n = min(s_span, sn - j * s_span);
  return true;
}

// The half-size table variant is limited to 64 VGPRs so that two 16-wave
// workgroups fit on a CU (8 waves per SIMD).
template <typename C>
__global__ __launch_bounds__(kFNT) __attribute__((amdgpu_waves_per_eu(C::Cap <= 8704 ? 8 : 1)))
void k_join_fused(const uint2 *__restrict__ R, const uint32_t *__restrict__ r_start,
                                                     const uint2 *__restrict__ S, const uint32_t *__restrict__ s_start,
                                                     const uint32_t *__restrict__ tile_start, uint32_t P,
                                                     uint32_t s_span, uint2 *__restrict__ out, uint64_t out_cap,
                                                     uint64_t *out_cursor, int wide, uint32_t *matched) {
  __shared__ tile::ProbeShared<C> sh;
  // round-robin over the XCDs: the span tiles of one hot (skewed) partition
  // are consecutive and must not all land on one XCD
  const uint32_t t = blockIdx.x;
  SpanRef ref;
  if (!resolve_span(t, tile_start, P, r_start, s_start, s_span, C::Cap, &ref)) return;
  tile::probe_tile<C>(sh, R, S, ref, out, out_cap, out_cursor, wide != 0, matched);
}

__global__ void k_plan_spans(const uint32_t *r_start, const uint32_t *s_start, uint32_t P, uint32_t s_span,
                             uint32_t cap, uint32_t *tiles) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > P) return;
  if (p == P) {
    tiles[P] = 0;
    return;
  }
  const uint32_t rn = r_start[p + 1] - r_start[p];
  const uint32_t sn = s_start[p + 1] - s_start[p];
  tiles[p] = (rn && sn) ? ((rn + cap - 1) / cap) * ((sn + s_span - 1) / s_span) : 0u;
}

__global__ void k_plan_tiles(const uint32_t *r_start, const uint32_t *s_start, uint32_t P, uint32_t cap,
                             uint32_t s_tile, uint32_t *tiles) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > P) return;
  if (p == P) {
    tiles[P] = 0;
    return;
  }
  const uint32_t rn = r_start[p + 1] - r_start[p];
  const uint32_t sn = s_start[p + 1] - s_start[p];
  tiles[p] = (rn && sn) ? ((rn + cap - 1) / cap) * ((sn + s_tile - 1) / s_tile) : 0u;
}

template <typename C>
hipError_t launch_count(const JoinTilePlan &jp, uint32_t n_tiles, const eo_tuple *R, const uint32_t *r_start,
                        const eo_tuple *S, const uint32_t *s_start, const uint32_t *tile_start, uint64_t *tile_count,
                        uint32_t *matched, hipStream_t stream) {
  const uint32_t grid = (n_tiles + kXcds - 1) / kXcds * kXcds;
  EO_TIMED("probe_count", stream, hipLaunchKernelGGL(k_join_count<C>, dim3(grid), dim3(C::NT), 0, stream, reinterpret_cast<const uint2 *>(R), r_start,
                     reinterpret_cast<const uint2 *>(S), s_start, tile_start, jp.parts, jp.s_tile, tile_count,
                     matched));
  return hipGetLastError();
}

template <typename C>
hipError_t launch_write(const JoinTilePlan &jp, uint32_t n_tiles, const eo_tuple *R, const uint32_t *r_start,
                        const eo_tuple *S, const uint32_t *s_start, const uint32_t *tile_start,
                        const uint64_t *tile_off, eo_pair *out, uint64_t out_cap, hipStream_t stream) {
  const uint32_t grid = (n_tiles + kXcds - 1) / kXcds * kXcds;
  EO_TIMED("probe_write", stream, hipLaunchKernelGGL(k_join_write<C>, dim3(grid), dim3(C::NT), 0, stream, reinterpret_cast<const uint2 *>(R), r_start,
                     reinterpret_cast<const uint2 *>(S), s_start, tile_start, jp.parts, jp.s_tile, tile_off,
                     reinterpret_cast<uint2 *>(out), out_cap, (int)jp.wide));
  return hipGetLastError();
}


}  // namespace

uint32_t join_max_lds_cap() { return Cfg16K::CAP; }

size_t join_plan_scratch_bytes(uint32_t parts) { return scan_scratch_bytes_u32((uint64_t)parts + 1); }

hipError_t join_plan_count(const JoinTilePlan &jp, const uint32_t *r_start, const uint32_t *s_start,
                           uint32_t *tile_start, uint64_t *n_tiles_dev, void *scan_scratch, uint32_t *err,
                           hipStream_t stream) {
  const uint32_t n = jp.parts + 1;
  hipLaunchKernelGGL(k_plan_tiles, dim3((n + 255) / 256), dim3(256), 0, stream, r_start, s_start, jp.parts,
                     jp.lds_cap, jp.s_tile, tile_start);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return scan_u32(tile_start, tile_start, n, n_tiles_dev, scan_scratch, err, stream);
}

hipError_t join_plan_map(const JoinTilePlan &, const uint32_t *, uint32_t *, hipStream_t) { return hipSuccess; }

hipError_t join_count(const JoinTilePlan &jp, uint32_t n_tiles, const eo_tuple *R, const uint32_t *r_start,
                      const eo_tuple *S, const uint32_t *s_start, const uint32_t *tile_start, uint32_t *matched,
                      uint64_t *tile_count, hipStream_t stream) {
  if (n_tiles == 0) return hipSuccess;
  switch (jp.lds_cap) {
    case Cfg4K::CAP:
      return launch_count<Cfg4K>(jp, n_tiles, R, r_start, S, s_start, tile_start, tile_count, matched, stream);
    case Cfg8K::CAP:
      return launch_count<Cfg8K>(jp, n_tiles, R, r_start, S, s_start, tile_start, tile_count, matched, stream);
    case Cfg16K::CAP:
      return launch_count<Cfg16K>(jp, n_tiles, R, r_start, S, s_start, tile_start, tile_count, matched, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t join_write(const JoinTilePlan &jp, uint32_t n_tiles, const eo_tuple *R, const uint32_t *r_start,
                      const eo_tuple *S, const uint32_t *s_start, const uint32_t *tile_start, const uint32_t *,
                      const uint64_t *tile_off, eo_pair *out, uint64_t out_cap, hipStream_t stream) {
  if (n_tiles == 0) return hipSuccess;
  switch (jp.lds_cap) {
    case Cfg4K::CAP:
      return launch_write<Cfg4K>(jp, n_tiles, R, r_start, S, s_start, tile_start, tile_off, out, out_cap, stream);
    case Cfg8K::CAP:
      return launch_write<Cfg8K>(jp, n_tiles, R, r_start, S, s_start, tile_start, tile_off, out, out_cap, stream);
    case Cfg16K::CAP:
      return launch_write<Cfg16K>(jp, n_tiles, R, r_start, S, s_start, tile_start, tile_off, out, out_cap, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace eo

namespace eo {

uint32_t join_fused_cap(uint32_t lds_cap) {
  return lds_cap <= 8192 ? (uint32_t)tile::ProbeSmall::Cap : (uint32_t)tile::ProbeBig::Cap;
}

hipError_t join_fused_plan(uint32_t parts, uint32_t s_span, uint32_t cap, const uint32_t *r_start,
                           const uint32_t *s_start, uint32_t *tile_start, uint64_t *n_tiles_dev, void *scan_scratch,
                           uint32_t *err, hipStream_t stream) {
  const uint32_t n = parts + 1;
  hipLaunchKernelGGL(k_plan_spans, dim3((n + 255) / 256), dim3(256), 0, stream, r_start, s_start, parts, s_span, cap,
                     tile_start);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return scan_u32(tile_start, tile_start, n, n_tiles_dev, scan_scratch, err, stream);
}

hipError_t join_fused(uint32_t parts, uint32_t s_span, uint32_t cap, uint32_t n_tiles, const eo_tuple *R,
                      const uint32_t *r_start, const eo_tuple *S, const uint32_t *s_start, const uint32_t *tile_start,
                      eo_pair *out, uint64_t out_cap, uint64_t *out_cursor, bool wide, uint32_t *matched,
                      hipStream_t stream) {
  if (n_tiles == 0) return hipSuccess;
  const uint32_t grid = (n_tiles + kXcds - 1) / kXcds * kXcds;
  const uint2 *vR = reinterpret_cast<const uint2 *>(R), *vS = reinterpret_cast<const uint2 *>(S);
  uint2 *vout = reinterpret_cast<uint2 *>(out);
  if (cap == (uint32_t)tile::ProbeSmall::Cap)
    EO_TIMED("probe", stream,
             hipLaunchKernelGGL(k_join_fused<tile::ProbeSmall>, dim3(grid), dim3(kFNT), 0, stream, vR, r_start, vS,
                                s_start, tile_start, parts, s_span, vout, out_cap, out_cursor, (int)wide,
                                matched));
  else if (cap == (uint32_t)tile::ProbeBig::Cap)
    EO_TIMED("probe", stream,
             hipLaunchKernelGGL(k_join_fused<tile::ProbeBig>, dim3(grid), dim3(kFNT), 0, stream, vR, r_start, vS,
                                s_start, tile_start, parts, s_span, vout, out_cap, out_cursor, (int)wide,
                                matched));
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

namespace {

// Left outer join: one row (null R side) per S tuple whose bit in `matched`
// is clear.  Each workgroup (256 words = 8192 positions) reserves its rows
// from *cursor with one atomic, or, with block_off, writes them at
// base + block_off[workgroup] (S order: deterministic).  out == NULL: only
// counts.
__global__ __launch_bounds__(256) void k_emit_unmatched(const uint2 *__restrict__ S, uint64_t n,
                                                        const uint32_t *__restrict__ matched, uint64_t *cursor,
                                                        const uint64_t *__restrict__ block_off, uint64_t base,
                                                        uint2 *__restrict__ out, uint64_t out_cap, int wide) {
  __shared__ uint64_t red[256 / kWave + 1];
  __shared__ uint64_t base_sh;
  const uint64_t wd = (uint64_t)blockIdx.x * 256 + threadIdx.x;  // one 32-position word per thread
  uint32_t unm = 0;
  if (wd * 32 < n) {
    const uint64_t left = n - wd * 32;
    const uint32_t valid = left >= 32 ? ~0u : ((1u << left) - 1u);
    unm = ~matched[wd] & valid;
  }
  uint64_t tot;
  const uint64_t excl = block_excl_scan<256, uint64_t>((uint64_t)__popc(unm), red, &tot);
  if (threadIdx.x == 0) {
    if (block_off) base_sh = base + block_off[blockIdx.x];
    else base_sh = tot ? atomic_add_agent(cursor, tot) : 0ull;
  }
  __syncthreads();
  if (!out) return;
  uint64_t pos = base_sh + excl;
  while (unm) {
    const uint32_t b = (uint32_t)__ffs(unm) - 1u;
    unm &= unm - 1u;
    const uint2 t = S[wd * 32 + b];
    if (pos < out_cap) emit_null(out, pos, t.y, t.x, wide != 0);
    ++pos;
  }
}

// Unmatched S positions per 256-word block (the counts k_emit_unmatched's
// deterministic mode scans into block offsets).
__global__ __launch_bounds__(256) void k_count_unmatched(uint64_t n, const uint32_t *__restrict__ matched,
                                                         uint64_t *__restrict__ block_cnt) {
  __shared__ uint64_t red[256 / kWave + 1];
  const uint64_t wd = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  uint32_t unm = 0;
  if (wd * 32 < n) {
    const uint64_t left = n - wd * 32;
    const uint32_t valid = left >= 32 ? ~0u : ((1u << left) - 1u);
    unm = ~matched[wd] & valid;
  }
  const uint64_t tot = block_sum<256, uint64_t>((uint64_t)__popc(unm), red);
  if (threadIdx.x == 0) block_cnt[blockIdx.x] = tot;
}

}  // namespace

size_t matched_words(uint64_t n) { return (n + 31) / 32; }

uint64_t unmatched_blocks(uint64_t n) { return (matched_words(n) + 255) / 256; }

hipError_t count_unmatched(uint64_t n, const uint32_t *matched, uint64_t *block_cnt, hipStream_t stream) {
  const uint64_t blocks = unmatched_blocks(n);
  if (blocks == 0) return hipSuccess;
  if (blocks >= (1ull << 31)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_count_unmatched, dim3((uint32_t)blocks), dim3(256), 0, stream, n, matched, block_cnt);
  return hipGetLastError();
}

hipError_t emit_unmatched_at(const eo_tuple *S, uint64_t n, const uint32_t *matched, const uint64_t *block_off,
                             uint64_t base, eo_pair *out, uint64_t out_cap, bool wide, hipStream_t stream) {
  const uint64_t blocks = unmatched_blocks(n);
  if (blocks == 0) return hipSuccess;
  if (blocks >= (1ull << 31)) return hipErrorInvalidValue;
  EO_TIMED("outer_rows", stream,
           hipLaunchKernelGGL(k_emit_unmatched, dim3((uint32_t)blocks), dim3(256), 0, stream,
                              reinterpret_cast<const uint2 *>(S), n, matched, nullptr, block_off, base,
                              reinterpret_cast<uint2 *>(out), out_cap, (int)wide));
  return hipGetLastError();
}

hipError_t emit_unmatched(const eo_tuple *S, uint64_t n, const uint32_t *matched, uint64_t *cursor, eo_pair *out,
                          uint64_t out_cap, bool wide, hipStream_t stream) {
  const uint64_t words = matched_words(n);
  if (words == 0) return hipSuccess;
  const uint64_t blocks = (words + 255) / 256;
  if (blocks >= (1ull << 31)) return hipErrorInvalidValue;
  EO_TIMED("outer_rows", stream,
           hipLaunchKernelGGL(k_emit_unmatched, dim3((uint32_t)blocks), dim3(256), 0, stream,
                              reinterpret_cast<const uint2 *>(S), n, matched, cursor, nullptr, 0ull,
                              reinterpret_cast<uint2 *>(out), out_cap, (int)wide));
  return hipGetLastError();
}

}  // namespace eo
