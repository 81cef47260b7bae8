// C ABI of libeojoin: context management, device memory helpers, primitives
// and the hash / nested-loop joins.  Orchestration mirrors the reference's
// join() (joins/hashJoin_gpu/main_street/hj.cc:171-1269): partition S and R,
// plan the probe, count, scan, write — with the per-launch synchronisation,
// per-call allocation and host chunking loops of the reference replaced by
// device-side planning and one arena reservation per call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "ctx.hpp"
#include "device.hpp"
#include "eojoin.h"
#include "kernels.hpp"
#include "plan.hpp"

namespace eo {

eo_status hip_fail(hipError_t e, const char *what) {
  set_last_error(std::string(what) + ": " + hipGetErrorString(e));
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return EO_ENOMEM;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return EO_ENODEV;
  return EO_EHIP;
}

hipError_t Arena::reserve(size_t bytes) {
  if (bytes <= cap) return hipSuccess;
  release();
  size_t want = std::max<size_t>(bytes, (size_t)64 << 20);
  want = (want + (1u << 21) - 1) & ~(size_t)((1u << 21) - 1);
  hipError_t e = hipMalloc((void **)&base, want);
  if (e != hipSuccess) {
    base = nullptr;
    cap = 0;
    return e;
  }
  cap = want;
  return hipSuccess;
}

void Arena::release() {
  if (base) (void)hipFree(base);
  base = nullptr;
  cap = 0;
}

hipError_t PinnedBuf::reserve(size_t bytes) {
  if (bytes <= cap) return hipSuccess;
  release();
  const size_t want = (std::max<size_t>(bytes, (size_t)16 << 20) + (1u << 21) - 1) & ~(size_t)((1u << 21) - 1);
  hipError_t e = hipHostMalloc((void **)&base, want, hipHostMallocDefault);
  if (e != hipSuccess) {
    base = nullptr;
    cap = 0;
    return e;
  }
  cap = want;
  return hipSuccess;
}

void PinnedBuf::release() {
  if (base) (void)hipHostFree(base);
  base = nullptr;
  cap = 0;
}

bool is_pinned(const eo_ctx *c, const void *p, size_t bytes) {
  const uintptr_t a = (uintptr_t)p;
  auto it = c->pinned.upper_bound(a);
  if (it == c->pinned.begin()) return false;
  --it;
  return a >= it->first && a + bytes <= it->first + it->second;
}

eo_status read_u64(eo_ctx *ctx, const uint64_t *dev, uint64_t *host, int n, hipStream_t stream) {
  EO_TRY_HIP(hipMemcpyAsync(ctx->h_pinned, dev, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, stream), "readback");
  EO_TRY_HIP(hipStreamSynchronize(stream), "readback sync");
  std::memcpy(host, ctx->h_pinned, sizeof(uint64_t) * n);
  return EO_OK;
}

eo_status check_device_errors(eo_ctx *ctx, hipStream_t stream) {
  EO_TRY_HIP(hipMemcpyAsync(ctx->h_pinned + 63, ctx->d_err, sizeof(uint32_t), hipMemcpyDeviceToHost, stream),
             "error flag readback");
  EO_TRY_HIP(hipStreamSynchronize(stream), "sync");
  const uint32_t flags = (uint32_t)ctx->h_pinned[63];
  if (flags) {
    EO_TRY_HIP(hipMemsetAsync(ctx->d_err, 0, sizeof(uint32_t), stream), "error flag reset");
    set_last_error("device error flags 0x" + std::to_string(flags));
    return EO_EINTERNAL;
  }
  return EO_OK;
}

void fold_timer(const LaunchTimer &t, eo_join_stats *stats) {
  stats->n_kernels = 0;
  for (int i = 0; i < t.n; ++i) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.ev[2 * i], t.ev[2 * i + 1]) != hipSuccess) continue;
    const char *nm = t.name[i];
    uint32_t k = 0;
    while (k < stats->n_kernels && std::strncmp(stats->kernels[k].name, nm, sizeof(stats->kernels[k].name)))
      ++k;
    if (k == stats->n_kernels) {
      if (k >= EO_MAX_KERNEL_STATS) continue;
      std::memset(&stats->kernels[k], 0, sizeof(stats->kernels[k]));
      std::strncpy(stats->kernels[k].name, nm, sizeof(stats->kernels[k].name) - 1);
      ++stats->n_kernels;
    }
    stats->kernels[k].ms += ms;
    stats->kernels[k].launches += 1;
  }
}

void ScopedTimer::finish() {
  if (stats) fold_timer(ctx->timer, stats);
}

bool opts_ok(const eo_join_opts *o) {
  if (!o) return true;
  if (o->struct_size != 0 && o->struct_size < sizeof(eo_join_opts)) return false;
  return o->where == EO_HOST || o->where == EO_DEVICE;
}

eo_status partition_relation(eo_ctx *ctx, const eo_tuple *in, uint64_t n, const HashPlan &hp, eo_tuple *tmp,
                             eo_tuple *out, uint32_t *part_start, void *scratch, hipStream_t s, bool stable,
                             bool fixed) {
  const int cus = ctx->props.cus;
  if (hp.passes == 1) {
    // fixed order: the ordered fast scatter up to 7 bits, else the stable one
    const bool fast = aligned16(in) && !stable && (!fixed || hp.bits_pass1 <= 7);
    PartPassPlan pp = plan_partition_pass(n, hp.bits_pass1, 0, 0u, cus, fast);
    pp.ordered = fast && fixed;
    EO_TRY_HIP(partition_pass(pp, in, out, part_start, scratch, ctx->d_err, s), "partition pass");
    return EO_OK;
  }
  const int b_lo = hp.bits_pass1, b_hi = hp.bits_pass2;
  if (stable) {
    // fall through to the stable LSD passes below
  } else if (aligned16(in) && aligned16(tmp) && msd2_supported(b_hi, b_lo)) {
    EO_TRY_HIP(partition_msd2(n, b_hi, b_lo, in, tmp, out, part_start, scratch, ctx->d_err, cus, s, fixed),
               "partition (two passes, one histogram)");
    return EO_OK;
  }
  if (!stable && !fixed && aligned16(in) && aligned16(tmp)) {
    const uint32_t F1 = 1u << b_hi;
    uint32_t *start1 = (uint32_t *)scratch;
    char *rest = (char *)scratch + ((F1 + 1) * sizeof(uint32_t) + 255) / 256 * 256;
    PartPassPlan p1 = plan_partition_pass(n, b_hi, b_lo, 0u, cus, true);
    EO_TRY_HIP(partition_pass(p1, in, tmp, start1, rest, ctx->d_err, s), "partition pass 1");
    SegPassPlan p2 = plan_seg_pass(n, F1, b_lo, 0, cus);
    EO_TRY_HIP(partition_seg_pass(p2, tmp, out, start1, part_start, rest, ctx->d_err, s), "partition pass 2");
    return EO_OK;
  }
  PartPassPlan p1 = plan_partition_pass(n, b_lo, 0, 0u, cus);
  PartPassPlan p2 = plan_partition_pass(n, b_hi, b_lo, 0u, cus);
  EO_TRY_HIP(partition_pass(p1, in, tmp, nullptr, scratch, ctx->d_err, s), "partition pass 1");
  EO_TRY_HIP(partition_pass(p2, tmp, out, nullptr, scratch, ctx->d_err, s), "partition pass 2");
  EO_TRY_HIP(partition_bounds(out, n, hp.bits, 0, 0u, part_start, s), "partition bounds");
  return EO_OK;
}

size_t partition_scratch(uint64_t n, const HashPlan &hp, int cus) {
  size_t b = 0;
  if (hp.passes == 1) {
    b = std::max(partition_pass_scratch_bytes(plan_partition_pass(n, hp.bits_pass1, 0, 0u, cus, true)),
                 partition_pass_scratch_bytes(plan_partition_pass(n, hp.bits_pass1, 0, 0u, cus, false)));
  } else if (hp.passes == 2) {
    const int b_lo = hp.bits_pass1, b_hi = hp.bits_pass2;
    const size_t head = (((1u << b_hi) + 1) * sizeof(uint32_t) + 255) / 256 * 256;
    b = std::max(partition_pass_scratch_bytes(plan_partition_pass(n, b_lo, 0, 0u, cus)),
                 partition_pass_scratch_bytes(plan_partition_pass(n, b_hi, b_lo, 0u, cus)));
    b = std::max(b, head + partition_pass_scratch_bytes(plan_partition_pass(n, b_hi, b_lo, 0u, cus, true)));
    b = std::max(b, head + seg_pass_scratch_bytes(plan_seg_pass(n, 1u << b_hi, b_lo, 0, cus)));
    if (msd2_supported(b_hi, b_lo)) b = std::max(b, msd2_scratch_bytes(n, b_hi, b_lo, cus));
  }
  return b;
}

}  // namespace eo

using namespace eo;

extern "C" {

eo_status eo_ctx_create(eo_ctx **out, int device, const eo_ctx_opts *opts) {
  if (!out) return EO_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    set_last_error("no HIP device visible");
    return EO_ENODEV;
  }
  if (device < 0 || device >= ndev) return EO_EINVAL;
  DeviceGuard g(device);
  eo_ctx *c = new eo_ctx();
  c->device = device;
  hipDeviceProp_t pr;
  if (hipGetDeviceProperties(&pr, device) != hipSuccess) {
    delete c;
    return EO_ENODEV;
  }
  c->props.name = pr.name;
  c->props.arch = pr.gcnArchName;
  c->props.cus = pr.multiProcessorCount;
  c->props.clock_khz = pr.clockRate;
  c->props.lds_per_wg = pr.sharedMemPerBlock;
  c->props.hbm = pr.totalGlobalMem;
  c->props.l2 = pr.l2CacheSize;
  c->hbm_budget = (opts && opts->hbm_budget) ? opts->hbm_budget : 200ull * 1000 * 1000 * 1000;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void **)&c->d_err, 256);
  if (e == hipSuccess) e = hipMemset(c->d_err, 0, 256);
  if (e == hipSuccess) e = hipHostMalloc((void **)&c->h_pinned, 64 * sizeof(uint64_t), hipHostMallocDefault);
  for (int i = 0; i < 8 && e == hipSuccess; ++i) e = hipEventCreate(&c->ev[i]);
  for (int i = 0; i < 2 * kMaxTimedLaunches && e == hipSuccess; ++i) e = hipEventCreate(&c->timer.ev[i]);
  if (e == hipSuccess && opts && opts->arena_bytes) e = c->arena.reserve(opts->arena_bytes);
  if (e != hipSuccess) {
    eo_status st = hip_fail(e, "eo_ctx_create");
    eo_ctx_destroy(c);
    return st;
  }
  // the device address of the no-allocation entry point's scratch, taken
  // once per device with the device current
  eo_status st = noalloc_prepare(device);
  if (st != EO_OK) {
    eo_ctx_destroy(c);
    return st;
  }
  *out = c;
  return EO_OK;
}

eo_status eo_ctx_destroy(eo_ctx *c) {
  if (!c) return EO_EINVAL;
  DeviceGuard g(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (eo_batch *b : c->serve_pool) eo_batch_free(b);
  c->serve_pool.clear();
  c->arena.release();
  for (hipStream_t &x : c->xfer)
    if (x) {
      (void)hipStreamSynchronize(x);
      (void)hipStreamDestroy(x);
    }
  c->stage.release();
  for (auto &ev : c->ev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto &ev : c->timer.ev)
    if (ev) (void)hipEventDestroy(ev);
  if (c->d_err) (void)hipFree(c->d_err);
  if (c->h_pinned) (void)hipHostFree(c->h_pinned);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return EO_OK;
}

eo_status eo_ctx_sync(eo_ctx *c) {
  if (!c) return EO_EINVAL;
  DeviceGuard g(c->device);
  EO_TRY_HIP(hipStreamSynchronize(c->stream), "eo_ctx_sync");
  return EO_OK;
}

eo_status eo_device_query(eo_ctx *c, eo_device_info *info) {
  if (!c || !info) return EO_EINVAL;
  std::memset(info, 0, sizeof(*info));
  std::strncpy(info->name, c->props.name.c_str(), sizeof(info->name) - 1);
  std::strncpy(info->arch, c->props.arch.c_str(), sizeof(info->arch) - 1);
  info->cus = c->props.cus;
  info->xcds = kXcds;
  info->wave_size = kWave;
  info->clock_khz = c->props.clock_khz;
  info->lds_per_wg = c->props.lds_per_wg;
  info->hbm_bytes = c->props.hbm;
  info->l2_bytes = c->props.l2;
  return EO_OK;
}

eo_status eo_dev_alloc(eo_ctx *c, uint64_t bytes, void **dptr) {
  if (!c || !dptr) return EO_EINVAL;
  DeviceGuard g(c->device);
  *dptr = nullptr;
  EO_TRY_HIP(hipMalloc(dptr, bytes ? bytes : 1), "eo_dev_alloc");
  return EO_OK;
}

eo_status eo_host_alloc(eo_ctx *c, uint64_t bytes, void **hptr) {
  if (!c || !hptr) return EO_EINVAL;
  DeviceGuard g(c->device);
  *hptr = nullptr;
  EO_TRY_HIP(hipHostMalloc(hptr, bytes ? bytes : 1, hipHostMallocMapped | hipHostMallocPortable), "eo_host_alloc");
  c->pinned[(uintptr_t)*hptr] = bytes ? bytes : 1;
  return EO_OK;
}

eo_status eo_host_free(eo_ctx *c, void *hptr) {
  if (!c) return EO_EINVAL;
  if (!hptr) return EO_OK;
  DeviceGuard g(c->device);
  c->pinned.erase((uintptr_t)hptr);
  EO_TRY_HIP(hipHostFree(hptr), "eo_host_free");
  return EO_OK;
}

eo_status eo_dev_free(eo_ctx *c, void *dptr) {
  if (!c) return EO_EINVAL;
  DeviceGuard g(c->device);
  if (dptr) EO_TRY_HIP(hipFree(dptr), "eo_dev_free");
  return EO_OK;
}

static eo_status copy_sync(eo_ctx *c, void *dst, const void *src, uint64_t bytes, hipMemcpyKind kind) {
  if (!c || (bytes && (!dst || !src))) return EO_EINVAL;
  DeviceGuard g(c->device);
  if (!bytes) return EO_OK;
  EO_TRY_HIP(hipMemcpyAsync(dst, src, bytes, kind, c->stream), "copy");
  EO_TRY_HIP(hipStreamSynchronize(c->stream), "copy sync");
  return EO_OK;
}

eo_status eo_copy_h2d(eo_ctx *c, void *dst, const void *src, uint64_t bytes) {
  return copy_sync(c, dst, src, bytes, hipMemcpyHostToDevice);
}
eo_status eo_copy_d2h(eo_ctx *c, void *dst, const void *src, uint64_t bytes) {
  return copy_sync(c, dst, src, bytes, hipMemcpyDeviceToHost);
}
eo_status eo_copy_d2d(eo_ctx *c, void *dst, const void *src, uint64_t bytes) {
  return copy_sync(c, dst, src, bytes, hipMemcpyDeviceToDevice);
}

eo_status eo_bandwidth_probe(eo_ctx *c, uint64_t bytes, int iters, double *gbs) {
  if (!c || !gbs || bytes == 0 || bytes % 16 || iters < 1) return EO_EINVAL;
  DeviceGuard g(c->device);
  EO_TRY_HIP(hipStreamSynchronize(c->stream), "sync");
  EO_TRY_HIP(c->arena.reserve(2 * bytes), "arena");
  char *a = c->arena.base, *b = c->arena.base + bytes;
  EO_TRY_HIP(copy_bytes(a, b, bytes, c->props.cus, c->stream), "warm copy");
  EO_TRY_HIP(hipEventRecord(c->ev[0], c->stream), "event");
  for (int i = 0; i < iters; ++i) EO_TRY_HIP(copy_bytes(a, b, bytes, c->props.cus, c->stream), "copy");
  EO_TRY_HIP(hipEventRecord(c->ev[1], c->stream), "event");
  EO_TRY_HIP(hipEventSynchronize(c->ev[1]), "event sync");
  float ms = 0.f;
  EO_TRY_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]), "elapsed");
  *gbs = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;
  return EO_OK;
}

eo_status eo_bandwidth_probe_ex(eo_ctx *c, const eo_probe_opts *o, double *gbs) {
  if (!c || !o || !gbs || o->struct_size < sizeof(eo_probe_opts)) return EO_EINVAL;
  if (o->bytes == 0 || o->bytes % 16 || o->iters < 1 || o->mode < EO_PROBE_COPY || o->mode > EO_PROBE_SCATTER)
    return EO_EINVAL;
  if (o->mode == EO_PROBE_SCATTER) {
    const uint64_t rb = (uint64_t)o->run_bytes, words = o->bytes / 16;
    if (rb < 16 || rb > 65536 || (rb & (rb - 1)) || (words & (words - 1)) || words < rb / 16) return EO_EINVAL;
  }
  if (o->unroll != 1 && o->unroll != 2 && o->unroll != 4 && o->unroll != 8) return EO_EINVAL;
  if (o->threads != 0 && (o->threads < 64 || o->threads > 1024 || o->threads % 64)) return EO_EINVAL;
  if (o->blocks_per_cu < 0 || o->blocks_per_cu > 64 || (o->flags & ~3u)) return EO_EINVAL;
  if (o->blocks_per_cu == 0 && o->bytes / 16 / (o->threads ? o->threads : 256) >= (1ull << 31)) return EO_EINVAL;
  DeviceGuard g(c->device);
  EO_TRY_HIP(hipStreamSynchronize(c->stream), "sync");
  const uint64_t span = (o->bytes + 255) / 256 * 256;
  EO_TRY_HIP(c->arena.reserve(2 * span + 4096), "arena");
  char *a = c->arena.base, *b = c->arena.base + span;
  uint32_t *sink = (uint32_t *)(c->arena.base + 2 * span);
  EO_TRY_HIP(bandwidth_kernel(*o, a, b, sink, c->props.cus, c->stream), "warm launch");
  EO_TRY_HIP(hipEventRecord(c->ev[0], c->stream), "event");
  for (int i = 0; i < o->iters; ++i) EO_TRY_HIP(bandwidth_kernel(*o, a, b, sink, c->props.cus, c->stream), "launch");
  EO_TRY_HIP(hipEventRecord(c->ev[1], c->stream), "event");
  EO_TRY_HIP(hipEventSynchronize(c->ev[1]), "event sync");
  float ms = 0.f;
  EO_TRY_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]), "elapsed");
  const double moved =
      (o->mode == EO_PROBE_COPY || o->mode == EO_PROBE_SCATTER ? 2.0 : 1.0) * (double)o->bytes * o->iters;
  *gbs = moved / (ms * 1e-3) / 1e9;
  return EO_OK;
}

eo_status eo_valu_probe(eo_ctx *c, int kind, int blocks_per_cu, int reps, int iters, double *per_s) {
  if (!c || !per_s || (kind != EO_VALU_COMPARE && kind != EO_VALU_ADD) || blocks_per_cu < 1 || blocks_per_cu > 32 ||
      reps < 1 || iters < 1)
    return EO_EINVAL;
  DeviceGuard g(c->device);
  EO_TRY_HIP(hipStreamSynchronize(c->stream), "sync");
  const uint32_t blocks = (uint32_t)blocks_per_cu * (uint32_t)c->props.cus;
  EO_TRY_HIP(c->arena.reserve((size_t)blocks * sizeof(uint64_t) + 256), "arena");
  uint64_t *sink = (uint64_t *)c->arena.base;
  double ops = 0;
  EO_TRY_HIP(valu_kernel(kind, blocks, (uint32_t)reps, sink, c->stream, &ops), "warm launch");
  EO_TRY_HIP(hipEventRecord(c->ev[0], c->stream), "event");
  for (int i = 0; i < iters; ++i) EO_TRY_HIP(valu_kernel(kind, blocks, (uint32_t)reps, sink, c->stream, &ops), "launch");
  EO_TRY_HIP(hipEventRecord(c->ev[1], c->stream), "event");
  EO_TRY_HIP(hipEventSynchronize(c->ev[1]), "event sync");
  float ms = 0.f;
  EO_TRY_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]), "elapsed");
  *per_s = ops * iters / (ms * 1e-3);
  return EO_OK;
}

eo_status eo_gen_build(eo_ctx *c, eo_tuple *R, uint64_t n, uint64_t seed, const eo_join_opts *o) {
  if (!c || (n && !R) || n > kMaxRows || !opts_ok(o)) return EO_EINVAL;
  DeviceGuard g(c->device);
  hipStream_t s = pick_stream(c, o);
  EO_TRY_HIP(gen_build(R, n, seed, s), "gen_build");
  EO_TRY_HIP(hipStreamSynchronize(s), "sync");
  return EO_OK;
}

eo_status eo_gen_probe(eo_ctx *c, eo_tuple *S, uint64_t n, uint64_t nR, double selectivity, uint64_t seed,
                       const eo_join_opts *o) {
  if (!c || (n && !S) || n > kMaxRows || nR > kMaxRows || !opts_ok(o) || !(selectivity >= 0.0 && selectivity <= 1.0))
    return EO_EINVAL;
  DeviceGuard g(c->device);
  hipStream_t s = pick_stream(c, o);
  EO_TRY_HIP(gen_probe(S, n, nR, selectivity, seed, s), "gen_probe");
  EO_TRY_HIP(hipStreamSynchronize(s), "sync");
  return EO_OK;
}

// ---- primitives ---------------------------------------------------------------

}  // extern "C"

template <typename T>
static eo_status scan_entry(eo_ctx *c, const T *in, T *out, uint64_t n, uint64_t *total, const eo_join_opts *o) {
  if (!c || !opts_ok(o) || (n && (!in || !out))) return EO_EINVAL;
  DeviceGuard g(c->device);
  hipStream_t s = pick_stream(c, o);
  const bool host = where_of(o) == EO_HOST;
  const size_t sb = sizeof(T) == 4 ? scan_scratch_bytes_u32(n) : scan_scratch_bytes_u64(n);
  Layout L;
  const size_t o_data = L.add(host ? n * sizeof(T) : 0);
  const size_t o_tot = L.add(sizeof(uint64_t));
  const size_t o_scr = L.add(sb);
  EO_TRY_HIP(hipStreamSynchronize(s), "sync");
  EO_TRY_HIP(c->arena.reserve(L.total), "arena");
  T *dbuf = host ? (T *)(c->arena.base + o_data) : nullptr;
  uint64_t *dtot = (uint64_t *)(c->arena.base + o_tot);
  void *scr = c->arena.base + o_scr;
  const T *din = in;
  T *dout = out;
  if (host) {
    if (n) EO_TRY_HIP(hipMemcpyAsync(dbuf, in, n * sizeof(T), hipMemcpyHostToDevice, s), "h2d");
    din = dbuf;
    dout = dbuf;
  }
  hipError_t e = sizeof(T) == 4
                     ? scan_u32((const uint32_t *)din, (uint32_t *)dout, n, dtot, scr, c->d_err, s)
                     : scan_u64((const uint64_t *)din, (uint64_t *)dout, n, dtot, scr, c->d_err, s);
  EO_TRY_HIP(e, "scan");
  if (host && n) EO_TRY_HIP(hipMemcpyAsync(out, dbuf, n * sizeof(T), hipMemcpyDeviceToHost, s), "d2h");
  uint64_t tot = 0;
  eo_status st = read_u64(c, dtot, &tot, 1, s);
  if (st != EO_OK) return st;
  st = check_device_errors(c, s);
  if (st != EO_OK) return st;
  if (total) *total = tot;
  return EO_OK;
}

extern "C" eo_status eo_exclusive_scan_u32(eo_ctx *c, const uint32_t *in, uint32_t *out, uint64_t n, uint64_t *total,
                                const eo_join_opts *o) {
  return scan_entry<uint32_t>(c, in, out, n, total, o);
}
extern "C" eo_status eo_exclusive_scan_u64(eo_ctx *c, const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *total,
                                const eo_join_opts *o) {
  return scan_entry<uint64_t>(c, in, out, n, total, o);
}

__global__ void k_widen(const uint32_t *in, uint64_t *out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

extern "C" {

eo_status eo_radix_partition(eo_ctx *c, const eo_tuple *in, uint64_t n, int bits, int shift, eo_tuple *out,
                             uint64_t *part_off, const eo_join_opts *o) {
  if (!c || !opts_ok(o) || bits < 1 || bits > 16 || shift < 0 || shift + bits > 32 || n > kMaxRows)
    return EO_EINVAL;
  if (n && (!in || !out)) return EO_EINVAL;
  DeviceGuard g(c->device);
  hipStream_t s = pick_stream(c, o);
  const bool host = where_of(o) == EO_HOST;
  const int b1 = bits <= kMaxPassBits ? bits : bits / 2;
  const int b2 = bits - b1;
  const PartPassPlan p1 = plan_partition_pass(n, b1, shift, 0u, c->props.cus);
  const PartPassPlan p2 = plan_partition_pass(n, b2 ? b2 : 1, shift + b1, 0u, c->props.cus);
  const uint32_t F = 1u << bits;
  Layout L;
  const size_t o_in = L.add(host ? n * sizeof(eo_tuple) : 0);
  const size_t o_out = L.add(host ? n * sizeof(eo_tuple) : 0);
  const size_t o_tmp = L.add(b2 ? n * sizeof(eo_tuple) : 0);
  const size_t o_ps = L.add((F + 1) * sizeof(uint32_t));
  const size_t o_po = L.add((F + 1) * sizeof(uint64_t));
  const size_t o_scr =
      L.add(std::max(partition_pass_scratch_bytes(p1), b2 ? partition_pass_scratch_bytes(p2) : (size_t)0));
  EO_TRY_HIP(hipStreamSynchronize(s), "sync");
  EO_TRY_HIP(c->arena.reserve(L.total), "arena");
  char *A = c->arena.base;
  const eo_tuple *din = host ? (eo_tuple *)(A + o_in) : in;
  eo_tuple *dout = host ? (eo_tuple *)(A + o_out) : out;
  eo_tuple *dtmp = (eo_tuple *)(A + o_tmp);
  uint32_t *ps = (uint32_t *)(A + o_ps);
  uint64_t *po = (uint64_t *)(A + o_po);
  if (host && n) EO_TRY_HIP(hipMemcpyAsync((void *)din, in, n * sizeof(eo_tuple), hipMemcpyHostToDevice, s), "h2d");
  if (n == 0) {
    EO_TRY_HIP(fill_u32(ps, 0u, F + 1, s), "fill");
  } else if (!b2) {
    EO_TRY_HIP(partition_pass(p1, din, dout, ps, A + o_scr, c->d_err, s), "partition");
  } else {
    EO_TRY_HIP(partition_pass(p1, din, dtmp, nullptr, A + o_scr, c->d_err, s), "partition pass 1");
    EO_TRY_HIP(partition_pass(p2, dtmp, dout, nullptr, A + o_scr, c->d_err, s), "partition pass 2");
    EO_TRY_HIP(partition_bounds(dout, n, bits, shift, 0u, ps, s), "partition bounds");
  }
  hipLaunchKernelGGL(k_widen, dim3((F + 1 + 255) / 256), dim3(256), 0, s, ps, po, F + 1);
  EO_TRY_HIP(hipGetLastError(), "widen");
  if (host && n) EO_TRY_HIP(hipMemcpyAsync(out, dout, n * sizeof(eo_tuple), hipMemcpyDeviceToHost, s), "d2h");
  if (part_off)
    EO_TRY_HIP(hipMemcpyAsync(part_off, po, (F + 1) * sizeof(uint64_t),
                              host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s),
               "part_off copy");
  EO_TRY_HIP(hipStreamSynchronize(s), "sync");
  return check_device_errors(c, s);
}

// ---- hash join ----------------------------------------------------------------

}  // extern "C"

namespace eo {

eo_status hash_join_impl(eo_ctx *c, const eo_tuple *R, uint64_t nR, const eo_tuple *S, uint64_t nS, eo_pair *out,
                         uint64_t out_cap, uint64_t *n_out, const eo_join_opts *o, const FixedScratch *fx) {
  if (!c || !n_out || !opts_ok(o) || nR > kMaxRows || nS > kMaxRows) return EO_EINVAL;
  if ((nR && !R) || (nS && !S) || (out_cap && !out)) return EO_EINVAL;
  *n_out = 0;
  const eo_tuple *const R0 = R, *const S0 = S;  // the caller's sides, for a redo with exact partitions
  const uint64_t nR0 = nR, nS0 = nS;
  HashPlan hp;
  eo_join_opts eff{};
  if (o) eff = *o;
  apply_tuned(c, nR, nS, &eff);
  eo_status st = plan_hash_join(nR, nS, &eff, &hp, c->props.cus);
  if (st != EO_OK) return st;
  eo_join_stats *stats = o ? o->stats : nullptr;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (nS == 0 || (nR == 0 && !(o && (o->flags & EO_F_LEFT_OUTER)))) return EO_OK;
  const bool wide = o && (o->flags & EO_F_RESULT16);
  const size_t row = wide ? 16 : sizeof(eo_pair);
  const bool outer = o && (o->flags & EO_F_LEFT_OUTER);

  const bool host = where_of(o) == EO_HOST;
  if (host && !stats && !(o && o->stream) && !(o && (o->flags & EO_F_DETERMINISTIC)) && nS >= kStreamMinS) {
    // Large host-resident S: stream it in chunks past R (copies in both
    // directions overlapped with the join) instead of copying both relations
    // whole, joining, and copying the rows back.
    eo_stream_opts so{};
    so.struct_size = sizeof(so);
    so.op = EO_STREAM_HASH;
    so.flags = o ? (o->flags & (EO_F_RESULT16 | EO_F_LEFT_OUTER)) : 0;
    so.join = &eff;
    return eo_stream_join(c, R, nR, S, nS, out, out_cap, n_out, &so);
  }
  // An inner join in the fused order builds its LDS tables on the smaller
  // relation and streams the larger one past them (the probe writes each row
  // with the payloads exchanged back): the partition bits follow the smaller
  // side, and a side that fits one table is not partitioned at all.
  const bool swapped = !outer && !(o && (o->flags & EO_F_DETERMINISTIC)) && nS < nR;
  if (swapped) {
    std::swap(R, S);
    std::swap(nR, nS);
    eff = o ? *o : eo_join_opts{};
    apply_tuned(c, nR, nS, &eff);
    if ((st = plan_hash_join(nR, nS, &eff, &hp, c->props.cus)) != EO_OK) return st;
  }
  DeviceGuard g(c->device);
  hipStream_t s = pick_stream(c, o);
  const uint32_t P = 1u << hp.bits;
  // Fused inner joins of two passes size the partitions of each large enough
  // relation from a sample instead of a histogram pass (gapped partitions,
  // [start, end) each); the fixed-scratch entry point keeps the exact ones.
  const bool deterministic = o && (o->flags & EO_F_DETERMINISTIC);
  const bool gap_try = !fx && !outer && !deterministic && hp.passes == 2 && !(o && (o->flags & EO_F_EXACT_PARTITIONS));
  const bool gapR = gap_try && gap_supported(nR, hp.bits_pass2, hp.bits_pass1);
  const bool gapS = gap_try && gap_supported(nS, hp.bits_pass2, hp.bits_pass1);
  const uint64_t rp_rows = gapR ? gap_out_rows(nR, hp.bits_pass2, hp.bits_pass1) : nR;
  const uint64_t sp_rows = gapS ? gap_out_rows(nS, hp.bits_pass2, hp.bits_pass1) : nS;
  const uint64_t tmp_rows = std::max(gapR ? gap_tmp_rows(nR, hp.bits_pass2, hp.bits_pass1) : nR,
                                     gapS ? gap_tmp_rows(nS, hp.bits_pass2, hp.bits_pass1) : nS);

  Layout L;
  const size_t o_rin = L.add(host ? nR * sizeof(eo_tuple) : 0);
  const size_t o_sin = L.add(host ? nS * sizeof(eo_tuple) : 0);
  const size_t o_rp = L.add(hp.passes ? rp_rows * sizeof(eo_tuple) : 0);
  const size_t o_sp = L.add(hp.passes ? sp_rows * sizeof(eo_tuple) : 0);
  const size_t o_tmp = L.add(hp.passes == 2 ? tmp_rows * sizeof(eo_tuple) : 0);
  const size_t o_rs = L.add((P + 1) * sizeof(uint32_t));
  const size_t o_ss = L.add((P + 1) * sizeof(uint32_t));
  const size_t o_re = L.add(gapR ? P * sizeof(uint32_t) : 0);
  const size_t o_se = L.add(gapS ? P * sizeof(uint32_t) : 0);
  const size_t o_ts = L.add((P + 1) * sizeof(uint32_t));
  const size_t o_small = L.add(8 * sizeof(uint64_t));
  const size_t o_bits = L.add(outer ? matched_words(nS) * sizeof(uint32_t) : 0);
  const uint64_t ublocks = outer ? unmatched_blocks(nS) : 0;  // deterministic outer rows: block offsets
  const size_t o_ub = L.add(ublocks * sizeof(uint64_t));
  const size_t o_ubs = L.add(ublocks ? scan_scratch_bytes_u64(ublocks) : 0);
  const size_t o_pscr = L.add(std::max(partition_scratch(nR, hp, c->props.cus), partition_scratch(nS, hp, c->props.cus)));
  // sampled relations keep their layouts from their prepare step to their
  // scatters, so each has scratch of its own
  const size_t o_gr = L.add(gapR ? gap_scratch_bytes(nR, hp.bits_pass2, hp.bits_pass1) : 0);
  const size_t o_gs = L.add(gapS ? gap_scratch_bytes(nS, hp.bits_pass2, hp.bits_pass1) : 0);
  const size_t o_jscr = L.add(join_plan_scratch_bytes(P));
  const uint64_t tiles_est = (uint64_t)P + nR / hp.lds_cap + nS / hp.s_span + 1024;
  const size_t o_tc = L.add(tiles_est * sizeof(uint64_t) + scan_scratch_bytes_u64(tiles_est));
  char *A = nullptr;
  if (fx) {  // the caller's fixed scratch: no allocation, no synchronisation
    if (L.total > fx->cap) {
      set_last_error("hash join needs " + std::to_string(L.total) + " bytes of scratch, the fixed buffer holds " +
                     std::to_string(fx->cap));
      return EO_EINVAL;
    }
    A = fx->base;
  } else {
    // growing the arena frees the old one: nothing may still use it; when it
    // is large enough, this call's work follows earlier work on the stream
    if (L.total > c->arena.cap) {
      EO_TRY_HIP(hipStreamSynchronize(s), "sync");
      EO_TRY_HIP(c->arena.reserve(L.total), "arena");
    }
    A = c->arena.base;
  }
  const eo_tuple *dR = host ? (eo_tuple *)(A + o_rin) : R;
  const eo_tuple *dS = host ? (eo_tuple *)(A + o_sin) : S;
  uint32_t *r_start = (uint32_t *)(A + o_rs);
  uint32_t *s_start = (uint32_t *)(A + o_ss);
  const uint32_t *r_end = r_start + 1, *s_end = s_start + 1;  // contiguous partitions
  uint32_t *tile_start = (uint32_t *)(A + o_ts);
  uint64_t *small = (uint64_t *)(A + o_small);
  if (host) {
    EO_TRY_HIP(hipMemcpyAsync((void *)dR, R, nR * sizeof(eo_tuple), hipMemcpyHostToDevice, s), "h2d R");
    EO_TRY_HIP(hipMemcpyAsync((void *)dS, S, nS * sizeof(eo_tuple), hipMemcpyHostToDevice, s), "h2d S");
  }

  PhaseClock clk(c, s, stats != nullptr);
  ScopedTimer tmr(c, stats);
  clk.mark();  // 0
  const eo_tuple *Rp = dR, *Sp = dS;
  bool used_gap = false;
  if (hp.passes) {
    eo_tuple *rp = (eo_tuple *)(A + o_rp), *sp = (eo_tuple *)(A + o_sp), *tmp = (eo_tuple *)(A + o_tmp);
    // deterministic order needs partitions whose positions the inputs fix
    // (the fast passes rank by atomics): the ordered two-pass partition, or
    // stable passes, which a deterministic left outer join takes so that its
    // null rows come in S order within each partition
    const bool stable = deterministic && outer, fixed = deterministic && !outer;
    const bool gr = gapR && aligned16(dR) && aligned16(tmp) && aligned16(rp);
    const bool gs = gapS && aligned16(dS) && aligned16(tmp) && aligned16(sp);
    // sampled relations: both samples and layouts first, so that a layout
    // that finds its relation clumped stops both relations' scatters
    auto gap = [&](const eo_tuple *in, uint64_t n, eo_tuple *dst, uint32_t *start, uint32_t *end, size_t o_scr,
                   unsigned phases) -> hipError_t {
      return partition_gapped(n, hp.bits_pass2, hp.bits_pass1, in, tmp, dst, start, end, A + o_scr, c->d_err,
                              c->props.cus, s, phases);
    };
    uint32_t *r_end_g = (uint32_t *)(A + o_re), *s_end_g = (uint32_t *)(A + o_se);
    if (gr) EO_TRY_HIP(gap(dR, nR, rp, r_start, r_end_g, o_gr, kGapPrepare), "partition sample (R)");
    if (gs) EO_TRY_HIP(gap(dS, nS, sp, s_start, s_end_g, o_gs, kGapPrepare), "partition sample (S)");
    if (gr) {
      EO_TRY_HIP(gap(dR, nR, rp, r_start, r_end_g, o_gr, kGapScatter), "partition (sampled regions, R)");
      r_end = r_end_g;
    } else {
      st = partition_relation(c, dR, nR, hp, tmp, rp, r_start, A + o_pscr, s, stable, fixed);
      if (st != EO_OK) return st;
    }
    if (gs) {
      EO_TRY_HIP(gap(dS, nS, sp, s_start, s_end_g, o_gs, kGapScatter), "partition (sampled regions, S)");
      s_end = s_end_g;
    } else {
      st = partition_relation(c, dS, nS, hp, tmp, sp, s_start, A + o_pscr, s, stable, fixed);
      if (st != EO_OK) return st;
    }
    used_gap = gr || gs;
    Rp = rp;
    Sp = sp;
  } else {
    EO_TRY_HIP(fill_u32(r_start, 0u, 1, s), "fill");
    EO_TRY_HIP(fill_u32(r_start + 1, (uint32_t)nR, 1, s), "fill");
    EO_TRY_HIP(fill_u32(s_start, 0u, 1, s), "fill");
    EO_TRY_HIP(fill_u32(s_start + 1, (uint32_t)nS, 1, s), "fill");
  }
  clk.mark();  // 1

  // both orders use the fused probe's tiles: (partition, R chunk, S span)
  const uint32_t cap = join_fused_cap(hp.lds_cap);
  EO_TRY_HIP(join_fused_plan(P, hp.s_span, cap, r_start, r_end, s_start, s_end, tile_start, small, A + o_jscr,
                             c->d_err, s),
             "tile plan");
  if (stats) {
    stats->radix_bits = (uint32_t)hp.bits;
    stats->passes = (uint32_t)hp.passes;
    stats->lds_tuples = hp.lds_cap;
    stats->sampled = used_gap ? 1u : 0u;
  }
  eo_pair *dout = out;
  bool staged = false;
  uint64_t total = 0;

  if (!deterministic) {
    // One fused probe kernel; with out == NULL (or a too small buffer) it counts
    // every match and writes only the pairs that fit.  The probe's grid does
    // not depend on the tile count, so the join reads back once: the tile
    // count, the match count and the device error flags together.
    uint64_t cap_dev = out ? out_cap : 0;
    if (host && out && out_cap) {
      EO_TRY_HIP(hipMallocAsync((void **)&dout, out_cap * row, s), "output staging");
      staged = true;
    }
    EO_TRY_HIP(fill_u64(small + 1, 0, 1, s), "cursor reset");
    uint32_t *matched = outer ? (uint32_t *)(A + o_bits) : nullptr;
    if (outer) EO_TRY_HIP(fill_u32(matched, 0u, matched_words(nS), s), "match bitmap reset");
    clk.mark();  // 2
    hipError_t e = join_fused(P, hp.s_span, cap, join_fused_grid(P, nR, nS, tiles_est, c->props.cus), small, Rp,
                              r_start, r_end, Sp, s_start, s_end, tile_start, dout, cap_dev, small + 1, wide, matched,
                              c->d_err, s, swapped);
    if (e == hipSuccess && outer) e = emit_unmatched(Sp, nS, matched, small + 1, dout, cap_dev, wide, s);
    clk.mark();  // 3
    uint64_t words[3] = {0, 0, 0};  // tiles, matches, error flags
    if (e == hipSuccess) e = take_error_flags(c->d_err, small + 2, s);
    if (e == hipSuccess) {
      st = read_u64(c, small, words, 3, s);
      total = words[1];
      if (st == EO_OK && used_gap && words[2] == kErrCapacity) {
        // a sampled region was too small: the rows are incomplete, so the
        // join runs again from the caller's inputs with exact partitions
        if (staged) (void)hipFreeAsync(dout, s);
        eo_join_opts ex = o ? *o : eo_join_opts{};
        if (!o) ex.where = EO_HOST;
        ex.struct_size = sizeof(ex);
        ex.flags |= EO_F_EXACT_PARTITIONS;
        st = hash_join_impl(c, R0, nR0, S0, nS0, out, out_cap, n_out, &ex, fx);
        if (stats) stats->sampled = 2u;
        return st;
      }
      if (st == EO_OK && !words[2] && staged && total) {  // on overflow: the out_cap rows that were written
        e = hipMemcpyAsync(out, dout, std::min(total, out_cap) * row, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
      }
    }
    if (staged) (void)hipFreeAsync(dout, s);
    if (e != hipSuccess) return hip_fail(e, "fused probe");
    if (st != EO_OK) return st;
    if (words[2]) {
      set_last_error("device error flags 0x" + std::to_string(words[2]) +
                     (words[0] >= (1ull << 31) ? " (2^31 or more join tiles)" : ""));
      return EO_EINTERNAL;
    }
    *n_out = total;
    if (stats) {
      stats->tiles = (uint32_t)words[0];
      stats->n_out = total;
      stats->ms_partition = clk.ms(0, 1);
      stats->ms_count = 0.0;
      stats->ms_write = clk.ms(2, 3);
      stats->ms_total = clk.ms(0, 3);
      tmr.finish();
    }
    return (out && total > out_cap) ? EO_ECAPACITY : EO_OK;
  }

  // Deterministic order: the count and write grids are exact, so the tile
  // count is read first.
  uint64_t n_tiles = 0;
  st = read_u64(c, small, &n_tiles, 1, s);
  if (st != EO_OK) return st;
  if (n_tiles >= (1ull << 31)) {
    set_last_error("too many join tiles");
    return EO_EINTERNAL;
  }
  if (stats) stats->tiles = (uint32_t)n_tiles;

  // Deterministic order: count -> scan -> write, nothing written on EO_ECAPACITY.
  char *B = A + o_tc;
  bool side = false;
  size_t o_cnt = 0, o_cscr = tiles_est * sizeof(uint64_t);
  if (n_tiles > tiles_est) {
    Layout L2;
    o_cnt = L2.add(n_tiles * sizeof(uint64_t));
    o_cscr = L2.add(scan_scratch_bytes_u64(n_tiles));
    EO_TRY_HIP(hipMallocAsync((void **)&B, L2.total, s), "tile count buffer");
    side = true;
  }
  uint64_t *tile_cnt = (uint64_t *)(B + o_cnt);
  auto free_side = [&]() {
    if (side) (void)hipFreeAsync(B, s);
  };
  // Left outer: the count pass marks matched S tuples; the null rows follow
  // the inner rows, in S order, at scanned per-block offsets.
  uint32_t *bits = outer ? (uint32_t *)(A + o_bits) : nullptr;
  uint64_t *ub = (uint64_t *)(A + o_ub);
  if (outer) EO_TRY_HIP(fill_u32(bits, 0u, matched_words(nS), s), "match bitmap reset");
  EO_TRY_HIP(span_count(P, hp.s_span, cap, (uint32_t)n_tiles, Rp, r_start, Sp, s_start, tile_start, bits, tile_cnt, s),
             "count");
  EO_TRY_HIP(scan_u64(tile_cnt, tile_cnt, n_tiles, small + 1, B + o_cscr, c->d_err, s), "count scan");
  if (outer) {
    EO_TRY_HIP(count_unmatched(nS, bits, ub, s), "unmatched count");
    EO_TRY_HIP(scan_u64(ub, ub, ublocks, small + 2, A + o_ubs, c->d_err, s), "unmatched scan");
  } else {
    EO_TRY_HIP(fill_u64(small + 2, 0, 1, s), "fill");
  }
  clk.mark();  // 2
  uint64_t totals[2] = {0, 0};
  st = read_u64(c, small + 1, totals, 2, s);
  if (st == EO_OK) st = check_device_errors(c, s);
  if (st != EO_OK) {
    free_side();
    return st;
  }
  const uint64_t inner = totals[0];
  total = totals[0] + totals[1];
  *n_out = total;
  if (stats) stats->n_out = total;
  if (!out) {
    free_side();
    return EO_OK;
  }
  if (total > out_cap) {
    free_side();
    return EO_ECAPACITY;
  }
  if (host && total) {
    hipError_t e = hipMallocAsync((void **)&dout, total * row, s);
    if (e != hipSuccess) {
      free_side();
      return hip_fail(e, "output staging");
    }
  }
  if (total) {
    hipError_t e = inner ? span_write(P, hp.s_span, cap, (uint32_t)n_tiles, Rp, r_start, Sp, s_start, tile_start,
                                      tile_cnt, dout, total, wide, s)
                         : hipSuccess;
    if (e == hipSuccess && outer) e = emit_unmatched_at(Sp, nS, bits, ub, inner, dout, total, wide, s);
    clk.mark();  // 3
    if (e == hipSuccess && host)
      e = hipMemcpyAsync(out, dout, total * row, hipMemcpyDeviceToHost, s);
    if (host) (void)hipFreeAsync(dout, s);
    if (e != hipSuccess) {
      free_side();
      return hip_fail(e, "write");
    }
  } else {
    clk.mark();
  }
  free_side();
  EO_TRY_HIP(hipStreamSynchronize(s), "sync");
  st = check_device_errors(c, s);
  if (stats) {
    stats->ms_partition = clk.ms(0, 1);
    stats->ms_count = clk.ms(1, 2);
    stats->ms_write = clk.ms(2, 3);
    stats->ms_total = clk.ms(0, 3);
    tmr.finish();
  }
  return st;
}

}  // namespace eo

extern "C" {

eo_status eo_hash_join(eo_ctx *c, const eo_tuple *R, uint64_t nR, const eo_tuple *S, uint64_t nS, eo_pair *out,
                       uint64_t out_cap, uint64_t *n_out, const eo_join_opts *o) {
  return hash_join_impl(c, R, nR, S, nS, out, out_cap, n_out, o, nullptr);
}

// ---- nested-loop join -----------------------------------------------------------

eo_status eo_nlj_join(eo_ctx *c, const eo_tuple *R, uint64_t nR, const eo_tuple *S, uint64_t nS, eo_pair *out,
                      uint64_t out_cap, uint64_t *n_out, const eo_join_opts *o) {
  if (!c || !n_out || !opts_ok(o) || nR > kMaxRows || nS > kMaxRows) return EO_EINVAL;
  if ((nR && !R) || (nS && !S) || (out_cap && !out)) return EO_EINVAL;
  *n_out = 0;
  eo_join_stats *stats = o ? o->stats : nullptr;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  const bool outer = o && (o->flags & EO_F_LEFT_OUTER);
  if (nS == 0 || (nR == 0 && !outer)) return EO_OK;
  const bool wide = o && (o->flags & EO_F_RESULT16);
  const size_t row = wide ? 16 : sizeof(eo_pair);
  DeviceGuard g(c->device);
  hipStream_t s = pick_stream(c, o);
  const bool host = where_of(o) == EO_HOST;
  const NljPlan np = plan_nlj(nR, nS);
  const uint64_t slots = nlj_count_slots(np);
  Layout L;
  const size_t o_rin = L.add(host ? nR * sizeof(eo_tuple) : 0);
  const size_t o_sin = L.add(host ? nS * sizeof(eo_tuple) : 0);
  const size_t o_cnt = L.add(slots * sizeof(uint64_t));
  const size_t o_small = L.add(8 * sizeof(uint64_t));
  const size_t o_bits = L.add(outer ? matched_words(nS) * sizeof(uint32_t) : 0);
  const uint64_t ublocks = outer ? unmatched_blocks(nS) : 0;  // deterministic outer rows: block offsets
  const size_t o_ub = L.add(ublocks * sizeof(uint64_t));
  const size_t o_ubs = L.add(ublocks ? scan_scratch_bytes_u64(ublocks) : 0);
  const size_t o_scr = L.add(scan_scratch_bytes_u64(slots));
  EO_TRY_HIP(hipStreamSynchronize(s), "sync");
  EO_TRY_HIP(c->arena.reserve(L.total), "arena");
  char *A = c->arena.base;
  const eo_tuple *dR = host ? (eo_tuple *)(A + o_rin) : R;
  const eo_tuple *dS = host ? (eo_tuple *)(A + o_sin) : S;
  uint64_t *cnt = (uint64_t *)(A + o_cnt);
  uint64_t *small = (uint64_t *)(A + o_small);
  if (host) {
    EO_TRY_HIP(hipMemcpyAsync((void *)dR, R, nR * sizeof(eo_tuple), hipMemcpyHostToDevice, s), "h2d R");
    EO_TRY_HIP(hipMemcpyAsync((void *)dS, S, nS * sizeof(eo_tuple), hipMemcpyHostToDevice, s), "h2d S");
  }
  PhaseClock clk(c, s, stats != nullptr);
  ScopedTimer tmr(c, stats);
  clk.mark();
  if (!(o && (o->flags & EO_F_DETERMINISTIC))) {
    // one fused pass; with out == NULL (or a too small buffer) it counts every
    // match and writes only the pairs that fit
    eo_pair *dout = out;
    const uint64_t cap_dev = out ? out_cap : 0;
    bool staged = false;
    if (host && out && out_cap) {
      EO_TRY_HIP(hipMallocAsync((void **)&dout, out_cap * row, s), "output staging");
      staged = true;
    }
    EO_TRY_HIP(fill_u64(small, 0, 1, s), "cursor reset");
    uint32_t *matched = outer ? (uint32_t *)(A + o_bits) : nullptr;
    if (outer) EO_TRY_HIP(fill_u32(matched, 0u, matched_words(nS), s), "match bitmap reset");
    hipError_t e = nlj_fused(np, dR, dS, small, out ? dout : nullptr, cap_dev, wide, matched, s);
    if (e == hipSuccess && outer) e = emit_unmatched(dS, nS, matched, small, out ? dout : nullptr, cap_dev, wide, s);
    clk.mark();
    uint64_t total = 0;
    eo_status st = EO_OK;
    if (e == hipSuccess) {
      st = read_u64(c, small, &total, 1, s);
      if (st == EO_OK && staged && total)  // on overflow: the out_cap rows that were written
        e = hipMemcpyAsync(out, dout, std::min(total, out_cap) * row, hipMemcpyDeviceToHost, s);
    }
    if (staged) (void)hipFreeAsync(dout, s);
    if (e != hipSuccess) return hip_fail(e, "nlj");
    if (st != EO_OK) return st;
    EO_TRY_HIP(hipStreamSynchronize(s), "sync");
    st = check_device_errors(c, s);
    *n_out = total;
    if (stats) {
      stats->n_out = total;
      stats->ms_count = out ? 0.0 : clk.ms(0, 1);
      stats->ms_write = out ? clk.ms(0, 1) : 0.0;
      stats->ms_total = clk.ms(0, 1);
      tmr.finish();
    }
    if (st != EO_OK) return st;
    return (out && total > out_cap) ? EO_ECAPACITY : EO_OK;
  }
  // Left outer: the count pass marks matched S tuples; the null rows follow
  // the inner rows, in S order, at scanned per-block offsets.
  uint32_t *bits = outer ? (uint32_t *)(A + o_bits) : nullptr;
  uint64_t *ub = (uint64_t *)(A + o_ub);
  if (outer) EO_TRY_HIP(fill_u32(bits, 0u, matched_words(nS), s), "match bitmap reset");
  EO_TRY_HIP(nlj_count(np, dR, dS, cnt, s, bits), "nlj count");
  EO_TRY_HIP(scan_u64(cnt, cnt, slots, small, A + o_scr, c->d_err, s), "nlj scan");
  if (outer) {
    EO_TRY_HIP(count_unmatched(nS, bits, ub, s), "unmatched count");
    EO_TRY_HIP(scan_u64(ub, ub, ublocks, small + 1, A + o_ubs, c->d_err, s), "unmatched scan");
  } else {
    EO_TRY_HIP(fill_u64(small + 1, 0, 1, s), "fill");
  }
  clk.mark();
  uint64_t totals[2] = {0, 0};
  eo_status st = read_u64(c, small, totals, 2, s);
  if (st == EO_OK) st = check_device_errors(c, s);
  if (st != EO_OK) return st;
  const uint64_t inner = totals[0], total = totals[0] + totals[1];
  *n_out = total;
  if (stats) stats->n_out = total;
  if (!out) return EO_OK;
  if (total > out_cap) return EO_ECAPACITY;
  if (total) {
    eo_pair *dout = out;
    if (host) EO_TRY_HIP(hipMallocAsync((void **)&dout, total * row, s), "output staging");
    hipError_t e = inner ? nlj_write(np, dR, dS, cnt, dout, total, wide, s) : hipSuccess;
    if (e == hipSuccess && outer) e = emit_unmatched_at(dS, nS, bits, ub, inner, dout, total, wide, s);
    clk.mark();
    if (e == hipSuccess && host) e = hipMemcpyAsync(out, dout, total * row, hipMemcpyDeviceToHost, s);
    if (host) (void)hipFreeAsync(dout, s);
    if (e != hipSuccess) return hip_fail(e, "nlj write");
  } else {
    clk.mark();
  }
  EO_TRY_HIP(hipStreamSynchronize(s), "sync");
  if (stats) {
    stats->ms_count = clk.ms(0, 1);
    stats->ms_write = clk.ms(1, 2);
    stats->ms_total = clk.ms(0, 2);
    tmr.finish();
  }
  return check_device_errors(c, s);
}

eo_status eo_nlj_dist_join(eo_ctx *c, const eo_point3 *R, uint64_t nR, const eo_point3 *S, uint64_t nS,
                           double distance, eo_pair *out, uint64_t out_cap, uint64_t *n_out, const eo_join_opts *o) {
  if (!c || !n_out || !opts_ok(o) || nR > kMaxRows || nS > kMaxRows) return EO_EINVAL;
  if ((nR && !R) || (nS && !S) || (out_cap && !out) || !(distance >= 0.0)) return EO_EINVAL;
  if (o && (o->flags & (EO_F_DETERMINISTIC | EO_F_RESULT16 | EO_F_LEFT_OUTER))) {
    set_last_error("eo_nlj_dist_join: only the default output mode");
    return EO_EINVAL;
  }
  *n_out = 0;
  eo_join_stats *stats = o ? o->stats : nullptr;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (nR == 0 || nS == 0) return EO_OK;
  DeviceGuard g(c->device);
  hipStream_t s = pick_stream(c, o);
  const bool host = where_of(o) == EO_HOST;
  Layout L;
  const size_t o_rin = L.add(host ? nR * sizeof(eo_point3) : 0);
  const size_t o_sin = L.add(host ? nS * sizeof(eo_point3) : 0);
  const size_t o_small = L.add(8 * sizeof(uint64_t));
  EO_TRY_HIP(hipStreamSynchronize(s), "sync");
  EO_TRY_HIP(c->arena.reserve(L.total), "arena");
  char *A = c->arena.base;
  const eo_point3 *dR = host ? (eo_point3 *)(A + o_rin) : R;
  const eo_point3 *dS = host ? (eo_point3 *)(A + o_sin) : S;
  uint64_t *small = (uint64_t *)(A + o_small);
  if (host) {
    EO_TRY_HIP(hipMemcpyAsync((void *)dR, R, nR * sizeof(eo_point3), hipMemcpyHostToDevice, s), "h2d R");
    EO_TRY_HIP(hipMemcpyAsync((void *)dS, S, nS * sizeof(eo_point3), hipMemcpyHostToDevice, s), "h2d S");
  }
  PhaseClock clk(c, s, stats != nullptr);
  ScopedTimer tmr(c, stats);
  eo_pair *dout = out;
  bool staged = false;
  if (host && out && out_cap) {
    EO_TRY_HIP(hipMallocAsync((void **)&dout, out_cap * sizeof(eo_pair), s), "output staging");
    staged = true;
  }
  EO_TRY_HIP(fill_u64(small, 0, 1, s), "cursor reset");
  clk.mark();
  hipError_t e = nlj_dist(dR, nR, dS, nS, distance * distance, small, out ? dout : nullptr, out ? out_cap : 0, s);
  clk.mark();
  uint64_t total = 0;
  eo_status st = EO_OK;
  if (e == hipSuccess) {
    st = read_u64(c, small, &total, 1, s);
    if (st == EO_OK && staged && total)  // on overflow: the out_cap rows that were written
      e = hipMemcpyAsync(out, dout, std::min(total, out_cap) * sizeof(eo_pair), hipMemcpyDeviceToHost, s);
  }
  if (staged) (void)hipFreeAsync(dout, s);
  if (e != hipSuccess) return hip_fail(e, "distance join");
  if (st != EO_OK) return st;
  EO_TRY_HIP(hipStreamSynchronize(s), "sync");
  st = check_device_errors(c, s);
  *n_out = total;
  if (stats) {
    stats->n_out = total;
    stats->ms_total = clk.ms(0, 1);
    stats->ms_write = out ? clk.ms(0, 1) : 0.0;
    stats->ms_count = out ? 0.0 : clk.ms(0, 1);
    tmr.finish();
  }
  if (st != EO_OK) return st;
  return (out && total > out_cap) ? EO_ECAPACITY : EO_OK;
}

}  // extern "C"
