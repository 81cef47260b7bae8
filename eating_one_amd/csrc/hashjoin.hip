// Per-partition join of a radix-partitioned R and S.
//
// The reference's probe (count / join, joins/hashJoin_gpu/main_street/
// join_gpu.cu:9-113) holds a <=256-tuple S chunk in shared memory and compares
// every R tuple of the partition against all of it — a partitioned nested
// loop, run twice (count, then write).  Here each work tile
//   (partition p, R-chunk c of <= the LDS table's tuples, S span j)
// builds a bucketised table of its R chunk in LDS (tile::probe_tile: LDS
// counting sort by a multiplicative hash) and streams its S span past it.
// Default order: one fused pass per tile that reserves its rows with one
// atomic per S batch.  Deterministic order (the reference's count -> presum ->
// join): a count pass stores one total per tile, a u64 scan turns them into
// tile offsets, and a write pass re-probes with a table order fixed by the
// inputs and writes each tile's rows from its offset on.
#include <algorithm>

#include "device.hpp"
#include "kernels.hpp"
#include "plan.hpp"
#include "tile_kernels.hpp"

namespace eo {
namespace {

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

// Partition p of R is [r_start[p], r_end[p]) (r_end = r_start + 1 for
// contiguous partitions), of S [s_start[p], s_end[p]).
__device__ __forceinline__ bool resolve_span(uint32_t t, const uint32_t *tile_start, uint32_t P,
                                             const uint32_t *r_start, const uint32_t *r_end,
                                             const uint32_t *s_start, const uint32_t *s_end, uint32_t s_span,
                                             uint32_t cap, SpanRef *ref) {
  if (t >= tile_start[P]) return false;
  const uint32_t p = wave_find(tile_start, P, t);
  const uint32_t local = t - tile_start[p];
  const uint32_t rb = r_start[p], rn = r_end[p] - rb;
  const uint32_t sb = s_start[p], sn = s_end[p] - sb;
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
//
// The grid is sized on the host without knowing the tile count (no readback
// between planning and probing): block b runs tiles b, b + grid, ... below the
// planned total *n_tiles, so a grid of one block per partition covers the
// usual one-tile partitions exactly and the extra span tiles of skewed
// partitions go to the blocks that started first.
template <typename C>
__global__ __launch_bounds__(kFNT) __attribute__((amdgpu_waves_per_eu(C::Cap <= 8704 ? 8 : 1)))
void k_join_fused(const uint2 *__restrict__ R, const uint32_t *__restrict__ r_start,
                  const uint32_t *__restrict__ r_end, const uint2 *__restrict__ S,
                  const uint32_t *__restrict__ s_start, const uint32_t *__restrict__ s_end,
                  const uint32_t *__restrict__ tile_start, uint32_t P,
                                                     uint32_t s_span, uint2 *__restrict__ out, uint64_t out_cap,
                                                     uint64_t *out_cursor, int wide, uint32_t *matched,
                                                     const uint64_t *__restrict__ n_tiles, uint32_t *err,
                                                     int swapped) {
  __shared__ tile::ProbeShared<C> sh;
  const uint64_t nt = *n_tiles;
  if (nt >= (1ull << 31)) {  // tile ids are 32-bit: report, probe nothing
    if (blockIdx.x == 0 && threadIdx.x == 0) atomic_or_agent(err, kErrOverflow);
    return;
  }
  // round-robin over the XCDs: the span tiles of one hot (skewed) partition
  // are consecutive and must not all land on one XCD
  for (uint32_t t = blockIdx.x; t < (uint32_t)nt; t += gridDim.x) {
    if (t != blockIdx.x) __syncthreads();  // the previous tile's LDS reads are done
    SpanRef ref;
    if (!resolve_span(t, tile_start, P, r_start, r_end, s_start, s_end, s_span, C::Cap, &ref)) return;
    tile::probe_tile<C>(sh, R, S, ref, out, out_cap, out_cursor, wide != 0, matched, 0, swapped != 0);
  }
}

// Deterministic order on the fused probe's tiles: a count pass (one count per
// tile, and the matched bits of a left outer join), a scan of the counts, and
// a write pass whose tiles write from their scanned offsets with a table
// order fixed by the inputs (kProbeOrdered).
template <typename C>
__global__ __launch_bounds__(kFNT) __attribute__((amdgpu_waves_per_eu(C::Cap <= 8704 ? 8 : 1)))
void k_span_count(const uint2 *__restrict__ R, const uint32_t *__restrict__ r_start, const uint2 *__restrict__ S,
                  const uint32_t *__restrict__ s_start, const uint32_t *__restrict__ tile_start, uint32_t P,
                  uint32_t s_span, uint64_t *__restrict__ tile_count, uint32_t *matched) {
  __shared__ tile::ProbeShared<C> sh;
  for (uint32_t t = blockIdx.x;; t += gridDim.x) {  // a grid below the tile count loops
    if (t != blockIdx.x) __syncthreads();
    SpanRef ref;
    if (!resolve_span(t, tile_start, P, r_start, r_start + 1, s_start, s_start + 1, s_span, C::Cap, &ref)) return;
    const uint64_t n = tile::probe_tile<C, tile::kProbeCount>(sh, R, S, ref, nullptr, 0, nullptr, false, matched);
    if (threadIdx.x == 0) tile_count[t] = n;
  }
}

template <typename C>
__global__ __launch_bounds__(kFNT) __attribute__((amdgpu_waves_per_eu(C::Cap <= 8704 ? 8 : 1)))
void k_span_write(const uint2 *__restrict__ R, const uint32_t *__restrict__ r_start, const uint2 *__restrict__ S,
                  const uint32_t *__restrict__ s_start, const uint32_t *__restrict__ tile_start, uint32_t P,
                  uint32_t s_span, const uint64_t *__restrict__ tile_off, uint2 *__restrict__ out, uint64_t out_cap,
                  int wide) {
  __shared__ tile::ProbeShared<C> sh;
  for (uint32_t t = blockIdx.x;; t += gridDim.x) {  // a grid below the tile count loops
    if (t != blockIdx.x) __syncthreads();
    SpanRef ref;
    if (!resolve_span(t, tile_start, P, r_start, r_start + 1, s_start, s_start + 1, s_span, C::Cap, &ref)) return;
    tile::probe_tile<C, tile::kProbeOrdered>(sh, R, S, ref, out, out_cap, nullptr, wide != 0, nullptr, tile_off[t]);
  }
}

__global__ void k_plan_spans(const uint32_t *r_start, const uint32_t *r_end, const uint32_t *s_start,
                             const uint32_t *s_end, uint32_t P, uint32_t s_span, uint32_t cap, uint32_t *tiles) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > P) return;
  if (p == P) {
    tiles[P] = 0;
    return;
  }
  const uint32_t rn = r_end[p] - r_start[p];
  const uint32_t sn = s_end[p] - s_start[p];
  tiles[p] = (rn && sn) ? ((rn + cap - 1) / cap) * ((sn + s_span - 1) / s_span) : 0u;
}

}  // namespace

size_t join_plan_scratch_bytes(uint32_t parts) { return scan_scratch_bytes_u32((uint64_t)parts + 1); }

}  // namespace eo

namespace eo {

static_assert(tile::ProbeBig::Cap == kFusedTableFull && tile::ProbeSmall::Cap == kFusedTableHalf,
              "the planner's table sizes are the probe's");

uint32_t join_fused_cap(uint32_t lds_cap) {
  return lds_cap <= 8192 ? (uint32_t)tile::ProbeSmall::Cap : (uint32_t)tile::ProbeBig::Cap;
}

hipError_t join_fused_plan(uint32_t parts, uint32_t s_span, uint32_t cap, const uint32_t *r_start,
                           const uint32_t *r_end, const uint32_t *s_start, const uint32_t *s_end,
                           uint32_t *tile_start, uint64_t *n_tiles_dev, void *scan_scratch, uint32_t *err,
                           hipStream_t stream) {
  const uint32_t n = parts + 1;
  hipLaunchKernelGGL(k_plan_spans, dim3((n + 255) / 256), dim3(256), 0, stream, r_start, r_end, s_start, s_end,
                     parts, s_span, cap, tile_start);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return scan_u32(tile_start, tile_start, n, n_tiles_dev, scan_scratch, err, stream);
}

// Largest grid of 1024-thread blocks one dispatch takes (2^32 work-items),
// in whole rounds over the 8 XCDs.
constexpr uint64_t kMaxFNTBlocks = (0xffffffffull / kFNT) / kXcds * kXcds;

// One block per tile, up to the dispatch limit (the kernels loop beyond it).
static uint32_t exact_grid(uint32_t n_tiles) {
  return (uint32_t)std::min<uint64_t>(((uint64_t)n_tiles + kXcds - 1) / kXcds * kXcds, kMaxFNTBlocks);
}

uint32_t join_fused_grid(uint32_t parts, uint64_t nR, uint64_t nS, uint64_t tiles_est, int cus) {
  // one block per partition that can hold a tile (the common one tile per
  // partition; a tile needs an R and an S tuple, so at most min(|R|, |S|)
  // partitions have one), at least two blocks per CU, never more than the
  // estimated tile count, and within the dispatch limit of 2^32 work-items
  uint64_t g = std::min<uint64_t>(parts, std::min(nR, nS));
  g = std::max<uint64_t>(g, 2ull * (uint64_t)cus);
  g = std::min<uint64_t>(g, std::max<uint64_t>(tiles_est, 1));
  g = (g + kXcds - 1) / kXcds * kXcds;
  return (uint32_t)std::min(g, kMaxFNTBlocks);
}

hipError_t join_fused(uint32_t parts, uint32_t s_span, uint32_t cap, uint32_t grid, const uint64_t *n_tiles_dev,
                      const eo_tuple *R, const uint32_t *r_start, const uint32_t *r_end, const eo_tuple *S,
                      const uint32_t *s_start, const uint32_t *s_end, const uint32_t *tile_start, eo_pair *out,
                      uint64_t out_cap, uint64_t *out_cursor, bool wide, uint32_t *matched, uint32_t *err,
                      hipStream_t stream, bool swapped) {
  if (grid == 0) return hipSuccess;
  const uint2 *vR = reinterpret_cast<const uint2 *>(R), *vS = reinterpret_cast<const uint2 *>(S);
  uint2 *vout = reinterpret_cast<uint2 *>(out);
  if (cap == (uint32_t)tile::ProbeSmall::Cap)
    EO_TIMED("probe", stream,
             hipLaunchKernelGGL(k_join_fused<tile::ProbeSmall>, dim3(grid), dim3(kFNT), 0, stream, vR, r_start, r_end,
                                vS, s_start, s_end, tile_start, parts, s_span, vout, out_cap, out_cursor, (int)wide,
                                matched, n_tiles_dev, err, (int)swapped));
  else if (cap == (uint32_t)tile::ProbeBig::Cap)
    EO_TIMED("probe", stream,
             hipLaunchKernelGGL(k_join_fused<tile::ProbeBig>, dim3(grid), dim3(kFNT), 0, stream, vR, r_start, r_end,
                                vS, s_start, s_end, tile_start, parts, s_span, vout, out_cap, out_cursor, (int)wide,
                                matched, n_tiles_dev, err, (int)swapped));
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t span_count(uint32_t parts, uint32_t s_span, uint32_t cap, uint32_t n_tiles, const eo_tuple *R,
                      const uint32_t *r_start, const eo_tuple *S, const uint32_t *s_start, const uint32_t *tile_start,
                      uint32_t *matched, uint64_t *tile_count, hipStream_t stream) {
  if (n_tiles == 0) return hipSuccess;
  const uint32_t grid = exact_grid(n_tiles);
  const uint2 *vR = reinterpret_cast<const uint2 *>(R), *vS = reinterpret_cast<const uint2 *>(S);
  if (cap == (uint32_t)tile::ProbeSmall::Cap)
    EO_TIMED("probe_count", stream,
             hipLaunchKernelGGL(k_span_count<tile::ProbeSmall>, dim3(grid), dim3(kFNT), 0, stream, vR, r_start, vS,
                                s_start, tile_start, parts, s_span, tile_count, matched));
  else if (cap == (uint32_t)tile::ProbeBig::Cap)
    EO_TIMED("probe_count", stream,
             hipLaunchKernelGGL(k_span_count<tile::ProbeBig>, dim3(grid), dim3(kFNT), 0, stream, vR, r_start, vS,
                                s_start, tile_start, parts, s_span, tile_count, matched));
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t span_write(uint32_t parts, uint32_t s_span, uint32_t cap, uint32_t n_tiles, const eo_tuple *R,
                      const uint32_t *r_start, const eo_tuple *S, const uint32_t *s_start, const uint32_t *tile_start,
                      const uint64_t *tile_off, eo_pair *out, uint64_t out_cap, bool wide, hipStream_t stream) {
  if (n_tiles == 0) return hipSuccess;
  const uint32_t grid = exact_grid(n_tiles);
  const uint2 *vR = reinterpret_cast<const uint2 *>(R), *vS = reinterpret_cast<const uint2 *>(S);
  uint2 *vout = reinterpret_cast<uint2 *>(out);
  if (cap == (uint32_t)tile::ProbeSmall::Cap)
    EO_TIMED("probe_write", stream,
             hipLaunchKernelGGL(k_span_write<tile::ProbeSmall>, dim3(grid), dim3(kFNT), 0, stream, vR, r_start, vS,
                                s_start, tile_start, parts, s_span, tile_off, vout, out_cap, (int)wide));
  else if (cap == (uint32_t)tile::ProbeBig::Cap)
    EO_TIMED("probe_write", stream,
             hipLaunchKernelGGL(k_span_write<tile::ProbeBig>, dim3(grid), dim3(kFNT), 0, stream, vR, r_start, vS,
                                s_start, tile_start, parts, s_span, tile_off, vout, out_cap, (int)wide));
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
