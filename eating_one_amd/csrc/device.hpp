// Device-side building blocks shared by every kernel of the library:
// wave64 scans on DPP row shifts/broadcasts, ballot-based peer matching,
// workgroup scans through LDS, and agent-scope (cross-XCD coherent) atomics.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace eo {

constexpr int kWave = 64;
constexpr int kXcds = 8;  // MI355X: 8 accelerator complex dies, WGs dealt round-robin

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Streaming (nontemporal) vector stores and loads for data written or read
// once: outputs of the scatters and the probe, the copy kernel.
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void st_stream(uint2 *p, uint2 v) {
  u32x2_t x;
  x.x = v.x;
  x.y = v.y;
  __builtin_nontemporal_store(x, reinterpret_cast<u32x2_t *>(p));
}
__device__ __forceinline__ void st_stream(uint4 *p, uint4 v) {
  u32x4_t x;
  x.x = v.x;
  x.y = v.y;
  x.z = v.z;
  x.w = v.w;
  __builtin_nontemporal_store(x, reinterpret_cast<u32x4_t *>(p));
}
// Result row pos: an 8-byte (R payload, S payload) pair, or with wide the
// reference's 16-byte RESULT {rkey, rval, lkey, lval} = (R payload, key,
// S payload, key).  stream: nontemporal store.
template <bool STREAM = false>
__device__ __forceinline__ void emit_pair(uint2 *out, uint64_t pos, uint32_t rp, uint32_t sp, uint32_t key,
                                          bool wide) {
  if (wide) {
    uint4 *o4 = reinterpret_cast<uint4 *>(out) + pos;
    const uint4 v = make_uint4(rp, key, sp, key);
    if (STREAM) st_stream(o4, v); else *o4 = v;
  } else {
    if (STREAM) st_stream(out + pos, make_uint2(rp, sp)); else out[pos] = make_uint2(rp, sp);
  }
}

// Row of an unmatched probe tuple in a left outer join: R side null.
constexpr uint32_t kNullPayload = 0x80000000u;  // INT32_MIN
template <bool STREAM = false>
__device__ __forceinline__ void emit_null(uint2 *out, uint64_t pos, uint32_t sp, uint32_t key, bool wide) {
  if (wide) {
    uint4 *o4 = reinterpret_cast<uint4 *>(out) + pos;
    const uint4 v = make_uint4(kNullPayload, kNullPayload, sp, key);
    if (STREAM) st_stream(o4, v); else *o4 = v;
  } else {
    if (STREAM) st_stream(out + pos, make_uint2(kNullPayload, sp)); else out[pos] = make_uint2(kNullPayload, sp);
  }
}

__device__ __forceinline__ uint4 ld_stream(const uint4 *p) {
  const u32x4_t x = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t *>(p));
  return make_uint4(x.x, x.y, x.z, x.w);
}

// ---- DPP wave scans ---------------------------------------------------------
// GFX9 DPP controls: row_shr:n = 0x110+n (within 16-lane rows),
// row_bcast:15 = 0x142 (lane 15 of row r -> all lanes of row r+1),
// row_bcast:31 = 0x143 (lane 31 -> rows 2 and 3).  Lanes whose source is
// outside the row, or whose row is masked off, receive `old` (= 0 here).
template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROWMASK, 0xf, false);
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  x += dpp_u32<0x111, 0xf>(x);
  x += dpp_u32<0x112, 0xf>(x);
  x += dpp_u32<0x114, 0xf>(x);
  x += dpp_u32<0x118, 0xf>(x);
  x += dpp_u32<0x142, 0xa>(x);
  x += dpp_u32<0x143, 0xc>(x);
  return x;
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t x) {
  uint32_t lo = dpp_u32<CTRL, ROWMASK>((uint32_t)x);
  uint32_t hi = dpp_u32<CTRL, ROWMASK>((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t x) {
  x += dpp_u64<0x111, 0xf>(x);
  x += dpp_u64<0x112, 0xf>(x);
  x += dpp_u64<0x114, 0xf>(x);
  x += dpp_u64<0x118, 0xf>(x);
  x += dpp_u64<0x142, 0xa>(x);
  x += dpp_u64<0x143, 0xc>(x);
  return x;
}

__device__ __forceinline__ uint32_t read_lane(uint32_t x, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, lane);
}
__device__ __forceinline__ uint64_t read_lane(uint64_t x, int lane) {
  uint32_t lo = read_lane((uint32_t)x, lane);
  uint32_t hi = read_lane((uint32_t)(x >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}

// FP64 wave scan (cross-lane shuffles; used off the hot path, e.g. model fits).
__device__ __forceinline__ double wave_incl_scan(double x) {
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const double y = __shfl_up(x, o, kWave);
    if (lane_id() >= o) x += y;
  }
  return x;
}
__device__ __forceinline__ double read_lane(double x, int lane) {
  return __longlong_as_double((long long)read_lane((uint64_t)__double_as_longlong(x), lane));
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
  return read_lane(wave_incl_scan(x), kWave - 1);
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Lanes whose `digit` (BITS low bits) equals this lane's, among the lanes in
// `active`.  BITS ballots; replaces a match_any instruction that CDNA lacks.
template <int BITS>
__device__ __forceinline__ uint64_t match_peers(uint32_t digit, uint64_t active) {
  uint64_t peers = active;
#pragma unroll
  for (int b = 0; b < BITS; ++b) {
    const bool bit = (digit >> b) & 1u;
    const uint64_t m = __ballot(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

// ---- workgroup scans --------------------------------------------------------
// Exclusive scan of one value per thread across a workgroup of NT threads.
// `lds` needs NT/64 + 1 entries.  Returns the exclusive prefix; *total gets
// the workgroup sum.  Ends with a barrier so `lds` may be reused.
template <int NT, typename T>
__device__ __forceinline__ T block_excl_scan(T x, T *lds, T *total) {
  constexpr int NW = NT / kWave;
  const int w = threadIdx.x / kWave;
  const int l = lane_id();
  T inc = wave_incl_scan(x);
  if (l == kWave - 1) lds[w] = inc;
  __syncthreads();
  if (w == 0) {
    T v = (l < NW) ? lds[l] : T(0);
    T vi = wave_incl_scan(v);
    if (l < NW) lds[l] = vi - v;
    if (l == NW - 1) lds[NW] = vi;
  }
  __syncthreads();
  T res = lds[w] + inc - x;
  *total = lds[NW];
  __syncthreads();
  return res;
}

template <int NT, typename T>
__device__ __forceinline__ T block_sum(T x, T *lds) {
  T total;
  (void)block_excl_scan<NT, T>(x, lds, &total);
  return total;
}

// Largest q in [0, n) with ts[q] <= t, for a non-decreasing ts with ts[0] <= t:
// a 64-ary search by one wave (one load per lane per level, 4 levels for 2^22).
// Every lane of the wave must take part (the ballot assumes a full wave).
__device__ __forceinline__ uint32_t wave_find(const uint32_t *__restrict__ ts, uint32_t n, uint32_t t) {
  uint32_t lo = 0;
  while (n > 1) {
    const uint32_t step = (n + kWave - 1) / kWave;
    const uint32_t idx = lo + (uint32_t)lane_id() * step;
    const bool le = idx < lo + n && ts[idx] <= t;
    const uint64_t m = __ballot(le);  // a prefix of the lanes; lane 0 always set
    const uint32_t k = 63u - (uint32_t)__clzll(m);
    const uint32_t nlo = lo + k * step;
    n = min(step, lo + n - nlo);
    lo = nlo;
  }
  return lo;
}

// Returns v unchanged through an empty asm statement (no instruction): the
// optimiser cannot merge expressions computed from the result with earlier
// ones.  Used where a rare branch recomputes compare ballots: merged, all of
// them would have to stay live in scalar registers (spilling) on the hot path.
__device__ __forceinline__ int32_t opaque(int32_t v) {
  asm volatile("" : "+v"(v));
  return v;
}

// ---- coherent atomics -------------------------------------------------------
// Agent scope makes loads/stores coherent across the XCDs' separate L2s.
__device__ __forceinline__ uint64_t ld_acquire(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_relaxed(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
// Returns p, stating that it points to global memory (not LDS, not scratch).
// Pointers read from memory (descriptor tables) are otherwise generic, and the
// compiler accesses them with flat instructions, which the LDS wait counter
// also tracks: every LDS wait would then also wait for those loads and stores.
// Rebuilt from its address as a global-address-space pointer (a plain cast
// pair would fold away), the compiler emits global instructions, as it does
// for kernel arguments.
template <typename T>
__device__ __forceinline__ T *as_global(T *p) {
  return (T *)(__attribute__((address_space(1))) T *)(uintptr_t)p;
}

// Takes the next ticket of a work queue (agent scope): returns *p and adds 1
// (the increment wraps only at 2^32).  An atomic add issued by one lane is
// rewritten by the compiler into a wave-combined form that reads its result
// right away, which makes the issuing wave wait for the round trip at once;
// the increment is left as is, so the wait falls at the first use of the
// ticket and a ticket taken ahead overlaps the work in between.
__device__ __forceinline__ uint32_t ticket_next(uint32_t *p) {
  return __builtin_amdgcn_atomic_inc32(p, 0xffffffffu, __ATOMIC_RELAXED, "agent");
}

__device__ __forceinline__ uint32_t atomic_add_agent(uint32_t *p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t atomic_add_agent(uint64_t *p, uint64_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_or_agent(uint32_t *p, uint32_t v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sets bit (base + lane) of the bitmap for every lane set in mask (the wave's
// 64 positions are consecutive): at most three word atomics per wave.
__device__ __forceinline__ void mark_bits(uint32_t *bits, uint64_t base, uint64_t mask) {
  if (!mask) return;
  const uint32_t s = (uint32_t)(base & 31);
  const int lane = lane_id();
  uint32_t v = 0;
  if (lane == 0) v = (uint32_t)(mask << s);
  else if (lane == 1) v = s ? (uint32_t)(mask >> (32 - s)) : (uint32_t)(mask >> 32);
  else if (lane == 2) v = s ? (uint32_t)(mask >> (64 - s)) : 0u;
  if (v) atomic_or_agent(bits + (base >> 5) + lane, v);
}

// Workgroup -> work-item mapping that keeps consecutive work items on one XCD:
// the hardware deals workgroup b to XCD b % 8, so item (b % 8) * (G/8) + b / 8
// puts items [x*G/8, (x+1)*G/8) on XCD x.  Bijective on [0, G) for G % 8 == 0.
__device__ __forceinline__ uint32_t xcd_local_index(uint32_t b, uint32_t grid) {
  const uint32_t per = grid / kXcds;
  return (b % kXcds) * per + b / kXcds;
}

// Multiplicative hash for LDS/global bucket tables (Knuth's golden ratio).
__device__ __forceinline__ uint32_t bucket_hash(int32_t key, int log2nb) {
  return log2nb == 0 ? 0u : (((uint32_t)key * 2654435769u) >> (32 - log2nb));
}

// Device error flags (bit set -> host reports EO_EHIP / EO_EINTERNAL).
enum : uint32_t {
  kErrSpin = 1u,       // a bounded wait gave up
  kErrOverflow = 2u,   // a device-side capacity was exceeded
  kErrCapacity = 4u,   // a partition region sized from a sample was too small (the join is redone exactly)
};

}  // namespace eo
