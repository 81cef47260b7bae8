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

}  // namespace

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

hipError_t copy_bytes(const void *src, void *dst, uint64_t bytes, int cus, hipStream_t stream) {
  const uint64_t n = bytes / 16;
  const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, (uint64_t)cus * 32));
  hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, stream, (const uint4 *)src, (uint4 *)dst, n);
  return hipGetLastError();
}

}  // namespace eo
