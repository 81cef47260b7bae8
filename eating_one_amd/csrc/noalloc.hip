// Hash join with no allocation: the join's scratch (partitioned copies of R
// and S, partition starts, tile lists, scan state, counters) lives in one
// fixed device array instead of the context's grow-only arena, so the call
// makes no hipMalloc / hipFree and can run where allocation is not allowed
// (under HIP graph capture, in serving loops that must not allocate).
//
// The array is a file-scope __device__ variable: every GPU the library runs
// on holds its own copy, at its own address, until the process ends.  Its
// address on a device is taken once, with that device current, when a
// context is created on it (noalloc_prepare, called by eo_ctx_create); the
// entry point passes that address to the kernels, never the variable's host
// name.  All calls on one GPU share the array and it keeps the last call's
// contents: every buffer in it is initialised by the join's own kernels, and
// calls that overlap on one GPU (from several host threads or contexts on
// the same device) race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "ctx.hpp"
#include "eojoin.h"

namespace {

constexpr size_t kNoAllocScratchBytes = (size_t)EO_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

std::mutex g_mu;
char *g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device

}  // namespace

namespace eo {

eo_status noalloc_prepare(int device) {
  if (device < 0 || device >= kMaxDevices) return EO_EINVAL;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return EO_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "no-allocation scratch");
  g_scratch[device] = (char *)p;
  return EO_OK;
}

}  // namespace eo

using namespace eo;

extern "C" eo_status eo_hash_join_noalloc(eo_ctx *c, const eo_tuple *R, uint64_t nR, const eo_tuple *S, uint64_t nS,
                                          eo_pair *out, uint64_t out_cap, uint64_t *n_out, const eo_join_opts *o) {
  if (!c || !o || !opts_ok(o) || o->where != EO_DEVICE) {
    set_last_error("eo_hash_join_noalloc: device pointers only (opts->where = EO_DEVICE)");
    return EO_EINVAL;
  }
  if (o->flags & EO_F_DETERMINISTIC) {
    set_last_error("eo_hash_join_noalloc: the deterministic order is not available without allocation");
    return EO_EINVAL;
  }
  char *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (c->device >= 0 && c->device < kMaxDevices) base = g_scratch[c->device];
  }
  if (!base) {
    set_last_error("eo_hash_join_noalloc: scratch not prepared on this device");
    return EO_EINTERNAL;
  }
  FixedScratch fx{base, kNoAllocScratchBytes};
  return hash_join_impl(c, R, nR, S, nS, out, out_cap, n_out, o, &fx);
}
