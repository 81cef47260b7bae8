// Library context: device, stream, scratch arena, error flag, readback buffer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "eojoin.h"
#include "kernels.hpp"

namespace eo {

void set_last_error(const std::string &msg);

// Bump allocator over one device buffer that only grows between calls.
struct Arena {
  char *base = nullptr;
  size_t cap = 0;
  hipError_t reserve(size_t bytes);  // may free and reallocate (caller has synchronised)
  void release();
};

// Computes offsets of sub-buffers inside one arena reservation.
struct Layout {
  size_t total = 0;
  size_t add(size_t bytes) {
    const size_t off = total;
    total += (bytes + 255) / 256 * 256;
    return off;
  }
};

// Autotuner table: size class (ceil log2 |R|, ceil log2 |S|) -> plan.
using TuneKey = std::pair<int, int>;
struct TunedPlan {
  int radix_bits, passes, lds_tuples, s_tile;
  double ms;
};

// Pinned host buffer that only grows between calls (staging of the
// streaming join when the caller's buffers are pageable).
struct PinnedBuf {
  char *base = nullptr;
  size_t cap = 0;
  hipError_t reserve(size_t bytes);  // may free and reallocate (nothing in flight)
  void release();
};

struct DeviceProps {
  std::string name, arch;
  int cus = 0;
  int clock_khz = 0;
  size_t lds_per_wg = 0, hbm = 0, l2 = 0;
};

}  // namespace eo

struct eo_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  eo::DeviceProps props;
  eo::Arena arena;
  uint32_t *d_err = nullptr;     // device-side error flags
  uint64_t *h_pinned = nullptr;  // 64 pinned words for small readbacks
  uint64_t hbm_budget = 0;
  hipEvent_t ev[8] = {};
  eo::LaunchTimer timer;  // per-launch events, installed while a call collects stats
  std::vector<eo_batch *> serve_pool;  // batch objects reused by eo_serve_run
  std::map<eo::TuneKey, eo::TunedPlan> tuned;  // autotuner results by size class
  std::map<uintptr_t, size_t> pinned;           // eo_host_alloc blocks: start -> bytes
  hipStream_t xfer[2] = {};                     // copy streams of eo_stream_join (H2D, D2H)
  eo::PinnedBuf stage;                          // its pinned staging
};

namespace eo {

// Scoped device selection.
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

eo_status hip_fail(hipError_t e, const char *what);
eo_status check_device_errors(eo_ctx *ctx, hipStream_t stream);
eo_status read_u64(eo_ctx *ctx, const uint64_t *dev, uint64_t *host, int n, hipStream_t stream);

constexpr uint64_t kMaxRows = (1ull << 31) - 1;
// Host-mode hash joins with at least this many S tuples run as streaming joins.
constexpr uint64_t kStreamMinS = 1ull << 22;

struct HashPlan;
bool opts_ok(const eo_join_opts *o);
inline bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }
// Partition `n` tuples into `out` and write the partition starts
// (2^bits + 1 u32) without any host synchronisation.  With a 16-byte aligned
// input: one fast pass, or two fast passes MSD-first (high digit over the
// whole array, then the low digit inside every first-pass partition).
// Otherwise, or with stable = true (order within a partition = input order),
// stable LSD passes (tmp holds the first pass).  fixed = true: positions
// fixed by the input (the same on every run): the two fast passes with
// ordered ranks where they apply, else the stable passes.
eo_status partition_relation(eo_ctx *ctx, const eo_tuple *in, uint64_t n, const HashPlan &hp, eo_tuple *tmp,
                             eo_tuple *out, uint32_t *part_start, void *scratch, hipStream_t s,
                             bool stable = false, bool fixed = false);
size_t partition_scratch(uint64_t n, const HashPlan &hp, int cus);
TuneKey tune_key(uint64_t nR, uint64_t nS);
// Fills radix_bits / passes / s_tile of *o from the tuning table when the
// caller left all plan fields at 0 and the size class has an entry.
void apply_tuned(const eo_ctx *c, uint64_t nR, uint64_t nS, eo_join_opts *o);
// Folds a launch timer's events into stats->kernels[] (stream synchronised).
void fold_timer(const LaunchTimer &t, eo_join_stats *stats);

// Device memory the caller provides in place of the context arena.
struct FixedScratch {
  char *base;
  size_t cap;
};
// eo_hash_join; with fx, its scratch comes from [fx->base, fx->base + cap)
// (a join that needs more is refused with EO_EINVAL) and nothing is
// allocated or synchronised before the kernels are enqueued.
eo_status hash_join_impl(eo_ctx *c, const eo_tuple *R, uint64_t nR, const eo_tuple *S, uint64_t nS, eo_pair *out,
                         uint64_t out_cap, uint64_t *n_out, const eo_join_opts *o, const FixedScratch *fx);
// Takes the device address of the no-allocation scratch on `device` (the
// current device) once; called by eo_ctx_create.
eo_status noalloc_prepare(int device);

// True when [p, p + bytes) lies inside one eo_host_alloc block of ctx.
bool is_pinned(const eo_ctx *ctx, const void *p, size_t bytes);

inline hipStream_t pick_stream(eo_ctx *ctx, const eo_join_opts *o) {
  return (o && o->stream) ? (hipStream_t)o->stream : ctx->stream;
}
inline int where_of(const eo_join_opts *o) { return o ? o->where : EO_HOST; }

// Installs the ctx's launch timer for the duration of a call when stats are
// requested; finish() folds the launch times into stats->kernels[].
struct ScopedTimer {
  eo_ctx *ctx;
  eo_join_stats *stats;
  LaunchTimer *prev = nullptr;
  ScopedTimer(eo_ctx *c, eo_join_stats *st) : ctx(c), stats(st) {
    if (stats) {
      ctx->timer.n = 0;
      ctx->timer.overflow = false;
      prev = set_launch_timer(&ctx->timer);
    }
  }
  ~ScopedTimer() {
    if (stats) set_launch_timer(prev);
  }
  void finish();  // call after the stream is synchronised
};

// Phase timing with events on the call's stream (only when stats requested).
struct PhaseClock {
  eo_ctx *ctx;
  hipStream_t stream;
  bool on;
  int n = 0;
  PhaseClock(eo_ctx *c, hipStream_t s, bool enabled) : ctx(c), stream(s), on(enabled) {}
  void mark() {
    if (on && n < 8) (void)hipEventRecord(ctx->ev[n++], stream);
  }
  double ms(int a, int b) const {
    float v = 0.f;
    if (!on || a >= n || b >= n) return 0.0;
    if (hipEventElapsedTime(&v, ctx->ev[a], ctx->ev[b]) != hipSuccess) return 0.0;
    return (double)v;
  }
};

}  // namespace eo

#define EO_TRY_HIP(call, what)                      \
  do {                                              \
    hipError_t eo_e_ = (call);                      \
    if (eo_e_ != hipSuccess) return eo::hip_fail(eo_e_, what); \
  } while (0)
