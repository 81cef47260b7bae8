// Seeded synthetic relations generated on the device.
//
// The reference builds its inputs on the host with rand() seeded by the
// clock (joins/hashJoin_gpu/main_street/hj.cc:62-160: unique R values by a
// partial Fisher-Yates shuffle over a `used` array, S = planted matches plus
// misses).  For multi-gigabyte request streams this generator runs on the
// device from counters instead, so that it is reproducible and costs one
// HBM write per tuple:
//   build side: key = pi(row), pi a seeded permutation of [0, n) (a bijection
//     of [0, 2^k) made of odd multiplications and xor-shifts mod 2^k, with
//     cycle walking down to [0, n)), payload = row;
//   probe side: with probability `selectivity` a key uniform in [0, nR)
//     (exactly one match against such a build side), otherwise a key uniform
//     in [nR, 2^31) (no match); payload = row.
#include "device.hpp"
#include "kernels.hpp"

namespace eo {
namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct PermKeys {
  uint64_t mask;
  int half;
  uint64_t mul[3], add[3];
};

__device__ __forceinline__ uint64_t perm_step(const PermKeys &k, uint64_t x) {
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    x = (x * k.mul[r] + k.add[r]) & k.mask;  // odd multiplier: bijective mod 2^k
    x ^= x >> k.half;                        // xor-shift: bijective on k bits
  }
  return x;
}

__global__ void k_gen_build(uint2 *__restrict__ out, uint64_t n, PermKeys k) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t x = perm_step(k, i);
    // cycle walking: the orbit of i under the permutation of [0, 2^k) returns
    // to [0, n) (it contains i); 2^k < 4n bounds the expected steps by 4
    while (x >= n) x = perm_step(k, x);
    out[i] = make_uint2((uint32_t)x, (uint32_t)i);
  }
}

__global__ void k_gen_probe(uint2 *__restrict__ out, uint64_t n, uint64_t nR, uint64_t thresh, uint64_t seed) {
  const uint64_t miss_range = (1ull << 31) - 1 - nR;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t u = splitmix64(seed ^ (i * 0xD1B54A32D192ED03ull));
    const uint64_t v = splitmix64(u);
    const bool hit = nR && (u >> 11) < thresh;
    const uint64_t r32 = v >> 32;
    const uint64_t key = hit ? (r32 * nR) >> 32 : nR + ((r32 * miss_range) >> 32);
    out[i] = make_uint2((uint32_t)key, (uint32_t)i);
  }
}

PermKeys perm_keys(uint64_t n, uint64_t seed) {
  int kb = 2;
  while ((1ull << kb) < n) kb += 2;  // even bit count: 2^k < 4n
  PermKeys k;
  k.mask = (1ull << kb) - 1;
  k.half = kb / 2 > 0 ? kb / 2 : 1;
  uint64_t s = seed;
  for (int r = 0; r < 3; ++r) {
    // host splitmix64 (same constants as the device one)
    auto next = [&s]() {
      s += 0x9E3779B97F4A7C15ull;
      uint64_t z = s;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      return z ^ (z >> 31);
    };
    k.mul[r] = (next() | 1ull) & k.mask;
    if (k.mul[r] == 0) k.mul[r] = 1;
    k.add[r] = next() & k.mask;
  }
  return k;
}

uint32_t gen_grid(uint64_t n) {
  const uint64_t g = (n + 255) / 256;
  return (uint32_t)(g < 8192 ? (g ? g : 1) : 8192);
}

}  // namespace

hipError_t gen_build(eo_tuple *out, uint64_t n, uint64_t seed, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gen_build, dim3(gen_grid(n)), dim3(256), 0, s, reinterpret_cast<uint2 *>(out), n,
                     perm_keys(n, seed));
  return hipGetLastError();
}

hipError_t gen_probe(eo_tuple *out, uint64_t n, uint64_t nR, double selectivity, uint64_t seed, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const double sel = selectivity < 0 ? 0 : (selectivity > 1 ? 1 : selectivity);
  const uint64_t thresh = sel >= 1.0 ? (1ull << 53) : (uint64_t)(sel * (double)(1ull << 53));
  hipLaunchKernelGGL(k_gen_probe, dim3(gen_grid(n)), dim3(256), 0, s, reinterpret_cast<uint2 *>(out), n, nR, thresh,
                     seed);
  return hipGetLastError();
}

}  // namespace eo
