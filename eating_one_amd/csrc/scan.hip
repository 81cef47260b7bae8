// Device-wide exclusive prefix sum, single pass with decoupled look-back
// (Merrill & Garland, "Single-pass Parallel Prefix Scan with Decoupled
// Look-back", NVIDIA TR NVR-2016-002 — the published method; written from the
// paper for wave64).  Replaces the reference's multi-launch presum()
// (joins/hashJoin_gpu/main_street/scan_main.cpp:38-111), which allocated a
// fresh output and freed its input on every call; this one allocates nothing.
//
// Tile status word (uint64): bits 62-63 = flag (0 empty, 1 aggregate,
// 2 inclusive prefix), bits 0-61 = value.  Tiles take ids from an atomic
// ticket, so a tile only ever waits on tiles that already started: the wait
// always ends.  It is additionally bounded; on expiry the kernel sets kErrSpin
// and finishes so the host reports an error instead of hanging.
#include "device.hpp"
#include "kernels.hpp"

namespace eo {
namespace {

constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagPrefix = 2ull << 62;
constexpr uint64_t kValMask = (1ull << 62) - 1;
constexpr uint32_t kSpinLimit = 1u << 24;

template <typename T, int NT, int IPT>
__global__ __launch_bounds__(NT) void k_scan(const T *__restrict__ in, T *out, uint64_t n,
                                             uint64_t *status, uint32_t *ticket, uint64_t *total_out,
                                             uint32_t *err) {
  constexpr uint64_t TILE = (uint64_t)NT * IPT;
  __shared__ uint32_t s_tile;
  __shared__ T s_wave[NT / kWave + 1];
  __shared__ T s_prefix;
  if (threadIdx.x == 0) s_tile = atomic_add_agent(ticket, 1u);
  __syncthreads();
  const uint64_t tile = s_tile;
  const uint64_t first = tile * TILE + (uint64_t)threadIdx.x * IPT;

  T v[IPT];
  T sum = 0;
  if (first + IPT <= n) {
#pragma unroll
    for (int i = 0; i < IPT; ++i) v[i] = in[first + i];
  } else {
#pragma unroll
    for (int i = 0; i < IPT; ++i) v[i] = (first + i < n) ? in[first + i] : T(0);
  }
#pragma unroll
  for (int i = 0; i < IPT; ++i) sum += v[i];

  T agg;
  const T thread_prefix = block_excl_scan<NT, T>(sum, s_wave, &agg);

  if (threadIdx.x < kWave) {
    const int lane = lane_id();
    T prefix = 0;
    if (tile == 0) {
      if (lane == 0) st_release(&status[0], kFlagPrefix | (uint64_t)agg);
    } else {
      if (lane == 0) st_release(&status[tile], kFlagAgg | (uint64_t)agg);
      int64_t j = (int64_t)tile - 1;
      while (true) {
        const int64_t idx = j - lane;
        uint64_t s;
        uint32_t spins = 0;
        bool ready;
        do {
          s = (idx >= 0) ? ld_relaxed(&status[idx]) : kFlagPrefix;
          ready = (s >> 62) != 0;
          if (__all(ready)) break;
          __builtin_amdgcn_s_sleep(1);
        } while (++spins < kSpinLimit);
        if (spins >= kSpinLimit) {
          if (lane == 0) atomic_or_agent(err, kErrSpin);
          break;
        }
        const uint64_t pmask = __ballot((s >> 62) == 2);
        const T val = (T)(s & kValMask);
        if (pmask) {
          const int firstp = __ffsll((unsigned long long)pmask) - 1;
          prefix += wave_sum<T>(lane <= firstp ? val : T(0));
          break;
        }
        prefix += wave_sum<T>(val);
        j -= kWave;
      }
      if (lane == 0) st_release(&status[tile], kFlagPrefix | (uint64_t)(prefix + agg));
    }
    if (lane == 0) s_prefix = prefix;
  }
  __syncthreads();
  const T tile_prefix = s_prefix;
  T run = tile_prefix + thread_prefix;
  if (first + IPT <= n) {
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      out[first + i] = run;
      run += v[i];
    }
  } else {
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      if (first + i < n) out[first + i] = run;
      run += v[i];
    }
  }
  if (total_out && threadIdx.x == 0 && (tile + 1) * TILE >= n) *total_out = (uint64_t)(tile_prefix + agg);
}

template <typename T>
__global__ void k_fill(T *p, T v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

constexpr int kScanNT = 256;
constexpr int kScanNTBig = 1024;  // long inputs: 4x longer tiles, 4x shorter look-back chain
constexpr uint64_t kScanBigMin = 1ull << 20;
constexpr int kScanIPT32 = 16;
constexpr int kScanIPT64 = 8;

template <typename T, int IPT>
size_t scan_scratch(uint64_t n) {  // sized for the short tiles: covers both launch shapes
  const uint64_t tiles = (n + (uint64_t)kScanNT * IPT - 1) / ((uint64_t)kScanNT * IPT);
  return (tiles + 1) * sizeof(uint64_t);
}

template <typename T, int IPT>
hipError_t scan_impl(const T *in, T *out, uint64_t n, uint64_t *total_dev, void *scratch, uint32_t *err,
                     hipStream_t stream) {
  if (n == 0) {
    if (total_dev) return fill_u64(total_dev, 0, 1, stream);
    return hipSuccess;
  }
  const bool big = n >= kScanBigMin;
  const uint64_t tile = (uint64_t)(big ? kScanNTBig : kScanNT) * IPT;
  const uint64_t tiles = (n + tile - 1) / tile;
  if (tiles >= (1ull << 31)) return hipErrorInvalidValue;
  uint64_t *status = (uint64_t *)scratch;
  uint32_t *ticket = (uint32_t *)(status + tiles);
  hipError_t e = fill_u64(status, 0, tiles + 1, stream);
  if (e != hipSuccess) return e;
  if (big)
    hipLaunchKernelGGL((k_scan<T, kScanNTBig, IPT>), dim3((uint32_t)tiles), dim3(kScanNTBig), 0, stream, in, out, n,
                       status, ticket, total_dev, err);
  else
    hipLaunchKernelGGL((k_scan<T, kScanNT, IPT>), dim3((uint32_t)tiles), dim3(kScanNT), 0, stream, in, out, n,
                       status, ticket, total_dev, err);
  return hipGetLastError();
}

}  // namespace

size_t scan_scratch_bytes_u32(uint64_t n) { return scan_scratch<uint32_t, kScanIPT32>(n); }
size_t scan_scratch_bytes_u64(uint64_t n) { return scan_scratch<uint64_t, kScanIPT64>(n); }

hipError_t scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint64_t *total_dev, void *scratch,
                    uint32_t *err, hipStream_t stream) {
  return scan_impl<uint32_t, kScanIPT32>(in, out, n, total_dev, scratch, err, stream);
}
hipError_t scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *total_dev, void *scratch,
                    uint32_t *err, hipStream_t stream) {
  return scan_impl<uint64_t, kScanIPT64>(in, out, n, total_dev, scratch, err, stream);
}

hipError_t fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill<uint32_t>, dim3(blocks), dim3(256), 0, stream, p, v, n);
  return hipGetLastError();
}
hipError_t fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill<uint64_t>, dim3(blocks), dim3(256), 0, stream, p, v, n);
  return hipGetLastError();
}

}  // namespace eo
