// Streaming copy kernel: the measured HBM ceiling every join kernel is
// compared against, and the known-size reference that calibrates rocprof's
// FETCH_SIZE / WRITE_SIZE counters.
#include <algorithm>

#include "device.hpp"
#include "kernels.hpp"

namespace eo {
namespace {

__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const uint4 a = ld_stream(src + i), b = ld_stream(src + i + stride), c = ld_stream(src + i + 2 * stride),
                d = ld_stream(src + i + 3 * stride);
    st_stream(dst + i, a);
    st_stream(dst + i + stride, b);
    st_stream(dst + i + 2 * stride, c);
    st_stream(dst + i + 3 * stride, d);
  }
  for (; i < n; i += stride) dst[i] = src[i];
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

template <int U, bool NTL, bool NTS>
hipError_t launch_probe(int mode, const uint4 *src, uint4 *dst, uint64_t n, uint32_t blocks, uint32_t nt,
                        uint32_t *sink, hipStream_t s) {
  if (mode == EO_PROBE_COPY)
    hipLaunchKernelGGL((k_probe_copy<U, NTL, NTS>), dim3(blocks), dim3(nt), 0, s, src, dst, n);
  else if (mode == EO_PROBE_READ)
    hipLaunchKernelGGL((k_probe_read<U, NTL>), dim3(blocks), dim3(nt), 0, s, src, n, 0x9e3779b9u, sink);
  else
    hipLaunchKernelGGL((k_probe_write<U, NTS>), dim3(blocks), dim3(nt), 0, s, dst, n, 0x51u);
  return hipGetLastError();
}

template <int U>
hipError_t launch_probe_u(int mode, bool ntl, bool nts, const uint4 *src, uint4 *dst, uint64_t n, uint32_t blocks,
                          uint32_t nt, uint32_t *sink, hipStream_t s) {
  if (ntl && nts) return launch_probe<U, true, true>(mode, src, dst, n, blocks, nt, sink, s);
  if (ntl) return launch_probe<U, true, false>(mode, src, dst, n, blocks, nt, sink, s);
  if (nts) return launch_probe<U, false, true>(mode, src, dst, n, blocks, nt, sink, s);
  return launch_probe<U, false, false>(mode, src, dst, n, blocks, nt, sink, s);
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
  switch (o.unroll) {
    case 1: return launch_probe_u<1>(o.mode, ntl, nts, a, b, n, blocks, nt, sink, s);
    case 2: return launch_probe_u<2>(o.mode, ntl, nts, a, b, n, blocks, nt, sink, s);
    case 4: return launch_probe_u<4>(o.mode, ntl, nts, a, b, n, blocks, nt, sink, s);
    case 8: return launch_probe_u<8>(o.mode, ntl, nts, a, b, n, blocks, nt, sink, s);
  }
  return hipErrorInvalidValue;
}

hipError_t copy_bytes(const void *src, void *dst, uint64_t bytes, int cus, hipStream_t stream) {
  const uint64_t n = bytes / 16;
  const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, (uint64_t)cus * 32));
  hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, stream, (const uint4 *)src, (uint4 *)dst, n);
  return hipGetLastError();
}

}  // namespace eo
