// Streaming copy kernel: the measured HBM ceiling every join kernel is
// compared against, and the known-size reference that calibrates rocprof's
// FETCH_SIZE / WRITE_SIZE counters.
#include <algorithm>

#include "device.hpp"
#include "kernels.hpp"

namespace eo {
namespace {

// One 16-byte word per thread with streaming loads and stores, a grid of one
// 256-thread block per 4 KB: the fastest copy shape of the bandwidth probe's
// sweep (bench.py --config copy).  Loops only past the dispatch limit.
__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    st_stream(dst + i, ld_stream(src + i));
}

// ---- bandwidth probes: the measured HBM ceiling ----------------------------
// Grid-stride kernels over 16-byte words with U loads in flight per thread
// (issued before any store), nontemporal (streaming) or normal loads/stores.
template <int U, bool NTL, bool NTS>
__global__ void k_probe_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n; i += U * stride) {
    uint4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = NTL ? ld_stream(src + i + u * stride) : src[i + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (NTS) st_stream(dst + i + u * stride, v[u]);
      else dst[i + u * stride] = v[u];
    }
  }
  for (; i < n; i += stride) dst[i] = src[i];
}

// Read-only: the words are folded into one value per thread, stored only when
// it equals a value the data never produces (so the loads are not removed).
template <int U, bool NTL>
__global__ void k_probe_read(const uint4 *__restrict__ src, uint64_t n, uint32_t never, uint32_t *sink) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t acc = 0;
  for (; i + (U - 1) * stride < n; i += U * stride) {
    uint4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = NTL ? ld_stream(src + i + u * stride) : src[i + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
  }
  for (; i < n; i += stride) acc ^= src[i].x ^ src[i].y ^ src[i].z ^ src[i].w;
  if (acc == never) sink[threadIdx.x] = acc;
}

template <int U, bool NTS>
__global__ void k_probe_write(uint4 *__restrict__ dst, uint64_t n, uint32_t v0) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint4 v = make_uint4(v0, v0 + 1, v0 + 2, v0 + 3);
  for (; i + (U - 1) * stride < n; i += U * stride) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (NTS) st_stream(dst + i + u * stride, v);
      else dst[i + u * stride] = v;
    }
  }
  for (; i < n; i += stride) dst[i] = v;
}

// Run-scattered copy: sequential reads, writes in runs of 2^run_log2 words
// sent to pseudo-random run positions (run r -> r * odd mod the run count, a
// bijection when the count is a power of two): the write pattern of a radix
// scatter whose digit runs have that length.
template <int U, bool NTL, bool NTS>
__global__ void k_probe_scatter(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n, int run_log2) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t runs_mask = (n >> run_log2) - 1;  // n / run is a power of two (checked by the host)
  const uint64_t in_run = (1ull << run_log2) - 1;
  auto dest = [&](uint64_t i) { return ((((i >> run_log2) * 0x9E3779B97F4A7C15ull) & runs_mask) << run_log2) | (i & in_run); };
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n; i += U * stride) {
    uint4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = NTL ? ld_stream(src + i + u * stride) : src[i + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (NTS) st_stream(dst + dest(i + u * stride), v[u]);
      else dst[dest(i + u * stride)] = v[u];
    }
  }
  for (; i < n; i += stride) dst[dest(i)] = src[i];
}

template <int U, bool NTL, bool NTS>
hipError_t launch_probe(int mode, const uint4 *src, uint4 *dst, uint64_t n, uint32_t blocks, uint32_t nt,
                        uint32_t *sink, int run_log2, hipStream_t s) {
  if (mode == EO_PROBE_SCATTER)
    hipLaunchKernelGGL((k_probe_scatter<U, NTL, NTS>), dim3(blocks), dim3(nt), 0, s, src, dst, n, run_log2);
  else if (mode == EO_PROBE_COPY)
    hipLaunchKernelGGL((k_probe_copy<U, NTL, NTS>), dim3(blocks), dim3(nt), 0, s, src, dst, n);
  else if (mode == EO_PROBE_READ)
    hipLaunchKernelGGL((k_probe_read<U, NTL>), dim3(blocks), dim3(nt), 0, s, src, n, 0x9e3779b9u, sink);
  else
    hipLaunchKernelGGL((k_probe_write<U, NTS>), dim3(blocks), dim3(nt), 0, s, dst, n, 0x51u);
  return hipGetLastError();
}

template <int U>
hipError_t launch_probe_u(int mode, bool ntl, bool nts, const uint4 *src, uint4 *dst, uint64_t n, uint32_t blocks,
                          uint32_t nt, uint32_t *sink, int run_log2, hipStream_t s) {
  if (ntl && nts) return launch_probe<U, true, true>(mode, src, dst, n, blocks, nt, sink, run_log2, s);
  if (ntl) return launch_probe<U, true, false>(mode, src, dst, n, blocks, nt, sink, run_log2, s);
  if (nts) return launch_probe<U, false, true>(mode, src, dst, n, blocks, nt, sink, run_log2, s);
  return launch_probe<U, false, false>(mode, src, dst, n, blocks, nt, sink, run_log2, s);
}

// ---- integer VALU issue-rate probes -----------------------------------------
// EO_VALU_COMPARE: the block nested-loop join's inner loop without its hits:
// 8 keys per lane in VGPRs compared with a 4096-key tile in LDS, 8 tile keys
// per group read as two 16-byte LDS loads, each compare's lane mask (a wave
// ballot) OR-accumulated; the keys never match (odd against even), so every
// group is one OR of 64 ballots.  Counts compares (lanes x keys).
// EO_VALU_ADD: 8 independent 32-bit add chains per lane (integer VALU issue
// with no memory or scalar work).  Counts adds.
constexpr int kValuNT = 256;
constexpr int kValuTile = 4096;

__global__ __launch_bounds__(kValuNT) void k_valu_compare(uint32_t seed, uint32_t reps, uint64_t *sink) {
  __shared__ __attribute__((aligned(16))) int32_t tile[kValuTile];
  for (int i = threadIdx.x; i < kValuTile; i += kValuNT) tile[i] = (int32_t)((seed * 2654435761u + i * 2u) & ~1u);
  int32_t key[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) key[k] = (int32_t)(((blockIdx.x * kValuNT + threadIdx.x) * 8u + k) * 2u + 1u);
  __syncthreads();
  uint64_t acc = 0;
  for (uint32_t r = 0; r < reps; ++r) {
    for (int g = 0; g < kValuTile; g += 8) {
      const int4 *q = reinterpret_cast<const int4 *>(tile + g);
      const int4 a = q[0], c = q[1];
      const int32_t rk[8] = {a.x, a.y, a.z, a.w, c.x, c.y, c.z, c.w};
#pragma unroll
      for (int q2 = 0; q2 < 8; ++q2)
#pragma unroll
        for (int k = 0; k < 8; ++k) acc |= __ballot(key[k] == rk[q2]);
    }
  }
  if (acc && threadIdx.x == 0) sink[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kValuNT) void k_valu_add(uint32_t seed, uint32_t reps, uint64_t *sink) {
  uint32_t a[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) a[k] = seed + threadIdx.x * 8u + k;
  for (uint32_t r = 0; r < reps; ++r) {
#pragma unroll 16
    for (int i = 0; i < 512; ++i) {
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] += a[(k + 1) & 7] ^ (uint32_t)i;
    }
  }
  uint32_t x = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) x ^= a[k];
  if (x == 0x9e3779b9u) sink[blockIdx.x] = x;
}

__global__ void k_take_flags(uint32_t *err, uint64_t *dst) {
  if (threadIdx.x == 0) {
    *dst = *err;
    *err = 0u;
  }
}

}  // namespace

hipError_t take_error_flags(uint32_t *err, uint64_t *dst, hipStream_t stream) {
  hipLaunchKernelGGL(k_take_flags, dim3(1), dim3(64), 0, stream, err, dst);
  return hipGetLastError();
}

static thread_local LaunchTimer *g_timer = nullptr;

LaunchTimer *set_launch_timer(LaunchTimer *t) {
  LaunchTimer *prev = g_timer;
  g_timer = t;
  return prev;
}

void timer_begin(const char *name, hipStream_t s) {
  if (!g_timer) return;
  if (g_timer->n >= kMaxTimedLaunches) {
    g_timer->overflow = true;
    return;
  }
  g_timer->name[g_timer->n] = name;
  (void)hipEventRecord(g_timer->ev[2 * g_timer->n], s);
}

void timer_end(hipStream_t s) {
  if (!g_timer || g_timer->n >= kMaxTimedLaunches) return;
  (void)hipEventRecord(g_timer->ev[2 * g_timer->n + 1], s);
  ++g_timer->n;
}

hipError_t bandwidth_kernel(const eo_probe_opts &o, const void *src, void *dst, uint32_t *sink, int cus,
                            hipStream_t s) {
  const uint64_t n = o.bytes / 16;
  const uint32_t nt = o.threads ? (uint32_t)o.threads : 256u;
  const uint32_t blocks = o.blocks_per_cu ? (uint32_t)o.blocks_per_cu * (uint32_t)cus
                                          : (uint32_t)std::max<uint64_t>(1, (n + nt - 1) / nt);
  const bool ntl = (o.flags & EO_PROBE_NT_LOAD) != 0, nts = (o.flags & EO_PROBE_NT_STORE) != 0;
  const uint4 *a = (const uint4 *)src;
  uint4 *b = (uint4 *)dst;
  int run_log2 = 0;
  while ((16u << run_log2) < (uint32_t)o.run_bytes) ++run_log2;
  switch (o.unroll) {
    case 1: return launch_probe_u<1>(o.mode, ntl, nts, a, b, n, blocks, nt, sink, run_log2, s);
    case 2: return launch_probe_u<2>(o.mode, ntl, nts, a, b, n, blocks, nt, sink, run_log2, s);
    case 4: return launch_probe_u<4>(o.mode, ntl, nts, a, b, n, blocks, nt, sink, run_log2, s);
    case 8: return launch_probe_u<8>(o.mode, ntl, nts, a, b, n, blocks, nt, sink, run_log2, s);
  }
  return hipErrorInvalidValue;
}

hipError_t valu_kernel(int kind, uint32_t blocks, uint32_t reps, uint64_t *sink, hipStream_t s, double *ops) {
  if (kind == EO_VALU_COMPARE) {
    hipLaunchKernelGGL(k_valu_compare, dim3(blocks), dim3(kValuNT), 0, s, 12345u, reps, sink);
    *ops = (double)blocks * kValuNT * 8.0 * kValuTile * reps;
  } else {
    hipLaunchKernelGGL(k_valu_add, dim3(blocks), dim3(kValuNT), 0, s, 12345u, reps, sink);
    *ops = (double)blocks * kValuNT * 8.0 * 512.0 * reps;
  }
  return hipGetLastError();
}

hipError_t copy_bytes(const void *src, void *dst, uint64_t bytes, int cus, hipStream_t stream) {
  (void)cus;
  const uint64_t n = bytes / 16;
  const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 1ull << 24));
  hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, stream, (const uint4 *)src, (uint4 *)dst, n);
  return hipGetLastError();
}

}  // namespace eo
