// Block nested-loop equi-join.
//
// The reference (joins/blockJoin_gpu/main_street/join_gpu.cu:28-118) holds
// 512 R tuples in shared memory per block and one S tuple per thread, writes
// one count per (thread, R slice) — 2^27 counts for a 2^18 x 2^18 tile — and
// repeats every compare in the write kernel.  Here a workgroup of 4 waves
// keeps an R tile of 4096 keys in LDS and 8 S keys per lane in VGPRs
// (a 2048-tuple S tile), and reads R keys from LDS four at a time with one
// wave-uniform (broadcast) load.  Each (R key, S register) compare is one
// VALU compare whose 64-lane ballot is OR-ed into a scalar accumulator; only
// a group of 8 R keys whose accumulator shows a hit is recounted exactly.
// Counts are kept per wave, so the write pass (same loop, same order) writes
// every pair at its wave's scanned offset in a deterministic order.
#include <algorithm>

#include "device.hpp"
#include "kernels.hpp"

namespace eo {
namespace {

constexpr int kNljNT = 256;
constexpr int kNljNW = kNljNT / kWave;
constexpr int kNljK = 8;                     // S keys per lane
constexpr int kNljSTile = kNljNT * kNljK;    // 2048
constexpr int kNljRTile = 4096;
constexpr int kNljGroup = 8;                 // R keys per OR-accumulated group

struct SRegs {
  int32_t key[kNljK];
  int32_t pay[kNljK];
  uint64_t valid[kNljK];  // ballot of lanes holding a real S tuple in register k
};

__device__ __forceinline__ void load_s(const uint2 *__restrict__ S, uint64_t nS, uint64_t s0, SRegs &sr,
                                       bool need_pay) {
  const int w = threadIdx.x / kWave, lane = lane_id();
#pragma unroll
  for (int k = 0; k < kNljK; ++k) {
    // wave w owns S[s0 + w*64*K, +64*K); register k holds element k*64 + lane of it
    const uint64_t i = s0 + (uint64_t)w * kWave * kNljK + (uint64_t)k * kWave + lane;
    const bool ok = i < nS;
    uint2 t = ok ? S[i] : make_uint2(0u, 0u);
    sr.key[k] = (int32_t)t.x;
    sr.pay[k] = need_pay ? (int32_t)t.y : 0;
    sr.valid[k] = __ballot(ok);
  }
}

// matched (count pass, optional): bitmap over S positions; this wave's S
// register k covers positions s_wave + k*64 + lane.
template <bool WRITE>
__device__ __forceinline__ void nlj_tile(const int32_t *rkey, const int32_t *rpay, uint32_t rn, const SRegs &sr,
                                         uint64_t *count, uint64_t base, uint2 *out, uint64_t out_cap, bool wide,
                                         uint32_t *matched = nullptr, uint64_t s_wave = 0) {
  const uint64_t lt = lanemask_lt();
  uint64_t running = 0;
  uint64_t hitk[kNljK];
#pragma unroll
  for (int k = 0; k < kNljK; ++k) hitk[k] = 0;
  uint32_t g = 0;
  for (; g + kNljGroup <= rn; g += kNljGroup) {
    int32_t rk[kNljGroup];
    const int4 *q = reinterpret_cast<const int4 *>(rkey + g);
    const int4 a = q[0], b = q[1];
    rk[0] = a.x; rk[1] = a.y; rk[2] = a.z; rk[3] = a.w;
    rk[4] = b.x; rk[5] = b.y; rk[6] = b.z; rk[7] = b.w;
    uint64_t acc = 0;
#pragma unroll
    for (int q2 = 0; q2 < kNljGroup; ++q2)
#pragma unroll
      for (int k = 0; k < kNljK; ++k) acc |= __ballot(sr.key[k] == rk[q2]);
    if (acc) {
#pragma unroll
      for (int q2 = 0; q2 < kNljGroup; ++q2) rk[q2] = opaque(rk[q2]);  // recompute, do not keep the ballots
#pragma unroll
      for (int q2 = 0; q2 < kNljGroup; ++q2)
#pragma unroll
        for (int k = 0; k < kNljK; ++k) {
          const uint64_t m = __ballot(sr.key[k] == rk[q2]) & sr.valid[k];
          hitk[k] |= m;
          if (m) {
            if (WRITE && sr.key[k] == rk[q2] && ((sr.valid[k] >> lane_id()) & 1)) {
              const uint64_t pos = base + running + __popcll(m & lt);
              if (pos < out_cap) emit_pair(out, pos, (uint32_t)rpay[g + q2], (uint32_t)sr.pay[k], (uint32_t)rk[q2], wide);
            }
            running += __popcll(m);
          }
        }
    }
  }
  for (; g < rn; ++g) {  // tail of the R tile
    const int32_t r = rkey[g];
#pragma unroll
    for (int k = 0; k < kNljK; ++k) {
      const uint64_t m = __ballot(sr.key[k] == r) & sr.valid[k];
      hitk[k] |= m;
      if (m) {
        if (WRITE && sr.key[k] == r && ((sr.valid[k] >> lane_id()) & 1)) {
          const uint64_t pos = base + running + __popcll(m & lt);
          if (pos < out_cap) emit_pair(out, pos, (uint32_t)rpay[g], (uint32_t)sr.pay[k], (uint32_t)r, wide);
        }
        running += __popcll(m);
      }
    }
  }
  if (!WRITE && lane_id() == 0) *count = running;
  if (!WRITE && matched) {
#pragma unroll
    for (int k = 0; k < kNljK; ++k) mark_bits(matched, s_wave + (uint64_t)k * kWave, hitk[k]);
  }
}

template <bool WRITE>
__global__ __launch_bounds__(kNljNT) void k_nlj(const uint2 *__restrict__ R, uint64_t nR, const uint2 *__restrict__ S,
                                                uint64_t nS, uint32_t s_tiles, uint64_t *wave_count,
                                                const uint64_t *__restrict__ wave_off, uint2 *__restrict__ out,
                                                uint64_t out_cap, int wide, uint32_t *matched) {
  __shared__ __attribute__((aligned(16))) int32_t rkey[kNljRTile];
  __shared__ int32_t rpay[WRITE ? kNljRTile : 1];
  const uint32_t b = blockIdx.x;
  const uint32_t st = b % s_tiles, rt = b / s_tiles;
  const uint64_t r0 = (uint64_t)rt * kNljRTile;
  const uint32_t rn = (uint32_t)std::min<uint64_t>(kNljRTile, nR - r0);
  for (uint32_t i = threadIdx.x; i < rn; i += kNljNT) {
    const uint2 t = R[r0 + i];
    rkey[i] = (int32_t)t.x;
    if (WRITE) rpay[i] = (int32_t)t.y;
  }
  SRegs sr;
  load_s(S, nS, (uint64_t)st * kNljSTile, sr, WRITE);
  __syncthreads();
  const uint64_t slot = (uint64_t)b * kNljNW + threadIdx.x / kWave;
  if (WRITE)
    nlj_tile<true>(rkey, rpay, rn, sr, nullptr, wave_off[slot], out, out_cap, wide != 0);
  else
    nlj_tile<false>(rkey, rpay, rn, sr, wave_count + slot, 0, nullptr, 0, false, matched,
                    (uint64_t)st * kNljSTile + (uint64_t)(threadIdx.x / kWave) * kWave * kNljK);
}

// Fused single pass (default order): every (R key group, S register) compare
// happens once.  A wave records the R groups whose OR-accumulated ballot shows
// a hit in an LDS list (at most 512 groups per R tile), counts their matches
// exactly, reserves its output range with one 64-bit atomic, and then
// re-scans only the recorded groups to write the pairs.  WRITE = false only
// counts (one atomic per wave).
template <bool WRITE>
__global__ __launch_bounds__(kNljNT) void k_nlj_fused(const uint2 *__restrict__ R, uint64_t nR,
                                                      const uint2 *__restrict__ S, uint64_t nS, uint32_t s_tiles,
                                                      uint64_t *cursor, uint2 *__restrict__ out, uint64_t out_cap,
                                                      int wide, uint32_t *matched) {
  // payloads are not staged: only the recorded hit groups need them (read from
  // global memory in the write phase), which keeps LDS at 20 KB per workgroup
  __shared__ __attribute__((aligned(16))) int32_t rkey[kNljRTile];
  __shared__ uint16_t hits[WRITE ? kNljNW : 1][WRITE ? kNljRTile / kNljGroup : 1];
  const uint32_t b = blockIdx.x;
  const uint32_t st = b % s_tiles, rt = b / s_tiles;
  const uint64_t r0 = (uint64_t)rt * kNljRTile;
  const uint32_t rn = (uint32_t)std::min<uint64_t>(kNljRTile, nR - r0);
  for (uint32_t i = threadIdx.x; i < rn; i += kNljNT) rkey[i] = (int32_t)R[r0 + i].x;
  SRegs sr;
  const uint64_t s0 = (uint64_t)st * kNljSTile;
  load_s(S, nS, s0, sr, false);
  __syncthreads();
  const int w = threadIdx.x / kWave;
  const uint64_t lt = lanemask_lt();
  const uint32_t ngroups = (rn + kNljGroup - 1) / kNljGroup;
  uint64_t total = 0;
  uint32_t nh = 0;
  uint64_t hitk[kNljK];  // lanes of each S register that matched (left outer join)
#pragma unroll
  for (int k = 0; k < kNljK; ++k) hitk[k] = 0;
  for (uint32_t gi = 0; gi < ngroups; ++gi) {
    const uint32_t g = gi * kNljGroup;
    const uint32_t gn = rn - g < (uint32_t)kNljGroup ? rn - g : (uint32_t)kNljGroup;  // wave-uniform
    int32_t rk[kNljGroup];
    if (gn == kNljGroup) {
      const int4 *q = reinterpret_cast<const int4 *>(rkey + g);
      const int4 a = q[0], c = q[1];
      rk[0] = a.x; rk[1] = a.y; rk[2] = a.z; rk[3] = a.w;
      rk[4] = c.x; rk[5] = c.y; rk[6] = c.z; rk[7] = c.w;
    } else {
#pragma unroll
      for (int q2 = 0; q2 < kNljGroup; ++q2) rk[q2] = rkey[g + (q2 < (int)gn ? q2 : 0)];
    }
    uint64_t acc = 0;
#pragma unroll
    for (int q2 = 0; q2 < kNljGroup; ++q2)
#pragma unroll
      for (int k = 0; k < kNljK; ++k) acc |= __ballot(sr.key[k] == rk[q2]);
    if (acc) {
      uint64_t cnt = 0;
#pragma unroll
      for (int q2 = 0; q2 < kNljGroup; ++q2) rk[q2] = opaque(rk[q2]);  // recompute, do not keep the ballots
#pragma unroll
      for (int q2 = 0; q2 < kNljGroup; ++q2) {
        if (q2 >= (int)gn) break;
#pragma unroll
        for (int k = 0; k < kNljK; ++k) {
          const uint64_t mm = __ballot(sr.key[k] == rk[q2]) & sr.valid[k];
          cnt += __popcll(mm);
          hitk[k] |= mm;
        }
      }
      if (cnt) {
        total += cnt;
        if (WRITE) {
          if (lane_id() == 0) hits[w][nh] = (uint16_t)gi;
          ++nh;
        }
      }
    }
  }
  if (matched) {
#pragma unroll
    for (int k = 0; k < kNljK; ++k)
      mark_bits(matched, s0 + (uint64_t)w * kWave * kNljK + (uint64_t)k * kWave, hitk[k]);
  }
  uint64_t base = 0;
  if (total) {
    if (lane_id() == 0) base = atomic_add_agent(cursor, total);
    base = read_lane(base, 0);
  }
  if (!WRITE || total == 0) return;
  uint64_t running = 0;
  for (uint32_t h = 0; h < nh; ++h) {
    const uint32_t g = (uint32_t)hits[w][h] * kNljGroup;
    const uint32_t gn = rn - g < (uint32_t)kNljGroup ? rn - g : (uint32_t)kNljGroup;
    for (uint32_t q2 = 0; q2 < gn; ++q2) {
      const int32_t r = rkey[g + q2];
#pragma unroll
      for (int k = 0; k < kNljK; ++k) {
        const bool eq = sr.key[k] == r && ((sr.valid[k] >> lane_id()) & 1);
        const uint64_t m = __ballot(eq);
        if (m) {
          if (eq) {
            const uint64_t pos = base + running + __popcll(m & lt);
            const uint64_t si = s0 + (uint64_t)w * kWave * kNljK + (uint64_t)k * kWave + lane_id();
            if (pos < out_cap) emit_pair(out, pos, R[r0 + g + q2].y, S[si].y, (uint32_t)r, wide != 0);
          }
          running += __popcll(m);
        }
      }
    }
  }
}

// ---- distance predicate (FP64) ---------------------------------------------------
// The multi-attribute predicate of joins/CPU_JOIN/3d_cpu/nestLoopJoin.cc:73-81:
// a pair matches when (r0-s0)^2 + (r1-s1)^2 + (r2-s2)^2 < d2, accumulated in
// that order with every subtraction, product and sum rounded on its own
// (__dsub_rn / __dmul_rn / __dadd_rn: no fused multiply-add), which makes each
// decision bit-identical to the reference's unfused x86-64 arithmetic.  Same
// fused single-pass structure as k_nlj_fused: R tile (1024 points, SoA
// doubles) in LDS, 4 S points per lane, OR-accumulated ballots per group of 8
// R points, per-wave hit-group list, one output reservation per wave.
constexpr int kDistNT = 256;
constexpr int kDistNW = kDistNT / kWave;
constexpr int kDistK = 4;
constexpr int kDistSTile = kDistNT * kDistK;  // 1024
constexpr int kDistRTile = 1024;
constexpr int kDistGroup = 8;

struct Point3 {
  int32_t id, pad;
  double v[3];
};

__device__ __forceinline__ bool dist_lt(double sx, double sy, double sz, double rx, double ry, double rz, double d2) {
  const double t0 = __dsub_rn(rx, sx), t1 = __dsub_rn(ry, sy), t2 = __dsub_rn(rz, sz);
  double acc = __dmul_rn(t0, t0);
  acc = __dadd_rn(acc, __dmul_rn(t1, t1));
  acc = __dadd_rn(acc, __dmul_rn(t2, t2));
  return acc < d2;
}

template <bool WRITE>
__global__ __launch_bounds__(kDistNT) void k_nlj_dist(const Point3 *__restrict__ R, uint64_t nR,
                                                      const Point3 *__restrict__ S, uint64_t nS, uint32_t s_tiles,
                                                      double d2, uint64_t *cursor, uint2 *__restrict__ out,
                                                      uint64_t out_cap) {
  __shared__ double rx[kDistRTile], ry[kDistRTile], rz[kDistRTile];
  __shared__ uint16_t hits[WRITE ? kDistNW : 1][WRITE ? kDistRTile / kDistGroup : 1];
  const uint32_t b = blockIdx.x;
  const uint32_t st = b % s_tiles, rt = b / s_tiles;
  const uint64_t r0 = (uint64_t)rt * kDistRTile;
  const uint32_t rn = (uint32_t)std::min<uint64_t>(kDistRTile, nR - r0);
  for (uint32_t i = threadIdx.x; i < rn; i += kDistNT) {
    const Point3 &p = R[r0 + i];
    rx[i] = p.v[0];
    ry[i] = p.v[1];
    rz[i] = p.v[2];
  }
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  const uint64_t s0 = (uint64_t)st * kDistSTile + (uint64_t)w * kWave * kDistK;
  double sx[kDistK], sy[kDistK], sz[kDistK];
  uint64_t valid[kDistK];
#pragma unroll
  for (int k = 0; k < kDistK; ++k) {
    const uint64_t i = s0 + (uint64_t)k * kWave + lane;
    const bool ok = i < nS;
    sx[k] = ok ? S[i].v[0] : 0.0;
    sy[k] = ok ? S[i].v[1] : 0.0;
    sz[k] = ok ? S[i].v[2] : 0.0;
    valid[k] = __ballot(ok);
  }
  __syncthreads();
  const uint64_t lt = lanemask_lt();
  const uint32_t ngroups = (rn + kDistGroup - 1) / kDistGroup;
  uint64_t total = 0;
  uint32_t nh = 0;
  for (uint32_t gi = 0; gi < ngroups; ++gi) {
    const uint32_t g = gi * kDistGroup;
    const uint32_t gn = rn - g < (uint32_t)kDistGroup ? rn - g : (uint32_t)kDistGroup;  // wave-uniform
    uint64_t acc = 0;
#pragma unroll
    for (int q = 0; q < kDistGroup; ++q) {
      const uint32_t j = g + (q < (int)gn ? q : 0);
      const double x = rx[j], y = ry[j], z = rz[j];
#pragma unroll
      for (int k = 0; k < kDistK; ++k) acc |= __ballot(dist_lt(sx[k], sy[k], sz[k], x, y, z, d2));
    }
    if (!acc) continue;
    uint64_t cnt = 0;
    for (uint32_t q = 0; q < gn; ++q) {
      const double x = rx[g + q], y = ry[g + q], z = rz[g + q];
#pragma unroll
      for (int k = 0; k < kDistK; ++k) cnt += __popcll(__ballot(dist_lt(sx[k], sy[k], sz[k], x, y, z, d2)) & valid[k]);
    }
    if (cnt) {
      total += cnt;
      if (WRITE) {
        if (lane == 0) hits[w][nh] = (uint16_t)gi;
        ++nh;
      }
    }
  }
  uint64_t base = 0;
  if (total) {
    if (lane == 0) base = atomic_add_agent(cursor, total);
    base = read_lane(base, 0);
  }
  if (!WRITE || total == 0) return;
  uint64_t running = 0;
  for (uint32_t h = 0; h < nh; ++h) {
    const uint32_t g = (uint32_t)hits[w][h] * kDistGroup;
    const uint32_t gn = rn - g < (uint32_t)kDistGroup ? rn - g : (uint32_t)kDistGroup;
    for (uint32_t q = 0; q < gn; ++q) {
      const double x = rx[g + q], y = ry[g + q], z = rz[g + q];
#pragma unroll
      for (int k = 0; k < kDistK; ++k) {
        const bool eq = dist_lt(sx[k], sy[k], sz[k], x, y, z, d2) && ((valid[k] >> lane) & 1);
        const uint64_t m = __ballot(eq);
        if (m) {
          if (eq) {
            const uint64_t pos = base + running + __popcll(m & lt);
            const uint64_t si = s0 + (uint64_t)k * kWave + lane;
            if (pos < out_cap) out[pos] = make_uint2((uint32_t)R[r0 + g + q].id, (uint32_t)S[si].id);
          }
          running += __popcll(m);
        }
      }
    }
  }
}

}  // namespace

NljPlan plan_nlj(uint64_t nR, uint64_t nS) {
  NljPlan p{};
  p.nR = nR;
  p.nS = nS;
  p.r_tile = kNljRTile;
  p.s_tile = kNljSTile;
  p.r_tiles = (uint32_t)((nR + kNljRTile - 1) / kNljRTile);
  p.s_tiles = (uint32_t)((nS + kNljSTile - 1) / kNljSTile);
  return p;
}

uint64_t nlj_count_slots(const NljPlan &p) { return (uint64_t)p.r_tiles * p.s_tiles * kNljNW; }

hipError_t nlj_count(const NljPlan &p, const eo_tuple *R, const eo_tuple *S, uint64_t *wave_count,
                     hipStream_t stream, uint32_t *matched) {
  const uint64_t grid = (uint64_t)p.r_tiles * p.s_tiles;
  if (grid == 0) return hipSuccess;
  if (grid >= (1ull << 31)) return hipErrorInvalidValue;
  EO_TIMED("nlj_count", stream, hipLaunchKernelGGL(k_nlj<false>, dim3((uint32_t)grid), dim3(kNljNT), 0, stream, reinterpret_cast<const uint2 *>(R),
                     p.nR, reinterpret_cast<const uint2 *>(S), p.nS, p.s_tiles, wave_count, nullptr, nullptr, 0, 0,
                     matched));
  return hipGetLastError();
}

hipError_t nlj_fused(const NljPlan &p, const eo_tuple *R, const eo_tuple *S, uint64_t *cursor, eo_pair *out,
                     uint64_t out_cap, bool wide, uint32_t *matched, hipStream_t stream) {
  const uint64_t grid = (uint64_t)p.r_tiles * p.s_tiles;
  if (grid == 0) return hipSuccess;
  if (grid >= (1ull << 31)) return hipErrorInvalidValue;
  if (out)
    EO_TIMED("nlj_fused", stream,
             hipLaunchKernelGGL(k_nlj_fused<true>, dim3((uint32_t)grid), dim3(kNljNT), 0, stream,
                                reinterpret_cast<const uint2 *>(R), p.nR, reinterpret_cast<const uint2 *>(S), p.nS,
                                p.s_tiles, cursor, reinterpret_cast<uint2 *>(out), out_cap, (int)wide, matched));
  else
    EO_TIMED("nlj_count", stream,
             hipLaunchKernelGGL(k_nlj_fused<false>, dim3((uint32_t)grid), dim3(kNljNT), 0, stream,
                                reinterpret_cast<const uint2 *>(R), p.nR, reinterpret_cast<const uint2 *>(S), p.nS,
                                p.s_tiles, cursor, nullptr, 0, 0, matched));
  return hipGetLastError();
}

hipError_t nlj_write(const NljPlan &p, const eo_tuple *R, const eo_tuple *S, const uint64_t *wave_off, eo_pair *out,
                     uint64_t out_cap, bool wide, hipStream_t stream) {
  const uint64_t grid = (uint64_t)p.r_tiles * p.s_tiles;
  if (grid == 0) return hipSuccess;
  if (grid >= (1ull << 31)) return hipErrorInvalidValue;
  EO_TIMED("nlj_write", stream, hipLaunchKernelGGL(k_nlj<true>, dim3((uint32_t)grid), dim3(kNljNT), 0, stream, reinterpret_cast<const uint2 *>(R),
                     p.nR, reinterpret_cast<const uint2 *>(S), p.nS, p.s_tiles, nullptr, wave_off,
                     reinterpret_cast<uint2 *>(out), out_cap, (int)wide, nullptr));
  return hipGetLastError();
}

hipError_t nlj_dist(const eo_point3 *R, uint64_t nR, const eo_point3 *S, uint64_t nS, double d2, uint64_t *cursor,
                    eo_pair *out, uint64_t out_cap, hipStream_t stream) {
  const uint64_t r_tiles = (nR + kDistRTile - 1) / kDistRTile, s_tiles = (nS + kDistSTile - 1) / kDistSTile;
  const uint64_t grid = r_tiles * s_tiles;
  if (grid == 0) return hipSuccess;
  if (grid >= (1ull << 31)) return hipErrorInvalidValue;
  const Point3 *vR = reinterpret_cast<const Point3 *>(R), *vS = reinterpret_cast<const Point3 *>(S);
  if (out)
    EO_TIMED("nlj_dist", stream,
             hipLaunchKernelGGL(k_nlj_dist<true>, dim3((uint32_t)grid), dim3(kDistNT), 0, stream, vR, nR, vS, nS,
                                (uint32_t)s_tiles, d2, cursor, reinterpret_cast<uint2 *>(out), out_cap));
  else
    EO_TIMED("nlj_dist_count", stream,
             hipLaunchKernelGGL(k_nlj_dist<false>, dim3((uint32_t)grid), dim3(kDistNT), 0, stream, vR, nR, vS, nS,
                                (uint32_t)s_tiles, d2, cursor, nullptr, 0));
  return hipGetLastError();
}

}  // namespace eo
