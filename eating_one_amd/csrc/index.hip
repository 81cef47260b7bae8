// Resident indexes on R and their probe kernels.
//
// Hash index (reference: CPU counting sort into NB_BUCKET buckets,
// joins/innerJoin_gpu/main_street/inner-idx-bnlj.cc:149-180, probed by one
// thread per S tuple in count.cu:15-48 / join_gpu.cu:8-86).  Here the build is
// on the device: bucket b = multiplicative hash of the key (NB = power of two
// >= |R|), histogram with global atomics, exclusive scan, placement by atomic
// cursors; the index is SoA keys[] / pays[] grouped by bucket plus off[NB+1].
//
// Sorted index (reference: CPU quicksort of {val, adr},
// joins/innerJoin_gpu/sort_index/main_street/inner-idx-bnlj.cc:175-216, probed
// by a device binary search and expansion over equal keys, join_gpu.cu:9-113).
// Here: stable LSD radix sort of the R tuples by key on the device
// (partition.hip, four 8-bit passes on the sign-flipped key), SoA keys[] /
// pays[] (the payload itself instead of the reference's row address adr; the
// joined pairs are the same).
//
// Learned sorted index: a 2-stage model over the sorted keys.
//   root: piecewise-linear CDF through K1 = 256 knots at equal-count
//         positions (knot i = (keys[i*n/K1], i*n/K1)); a key's segment is found
//         by binary search over the knots (in LDS), the position interpolated
//         in FP32 and mapped to leaf = pred * L / n, L = 4096;
//   leaf: FP32 linear model pos = a + b * (key - k0) fit by least squares
//         (FP64 sums) on the keys the root routes to it; err_lo / err_hi = the
//         largest under- / over-prediction of the first position of any key
//         routed to the leaf, measured with this same inference code.
// A probe predicts, then binary-searches only [pred - err_hi, pred + err_lo]
// (clamped): every key present in R lies in its window, so results equal the
// plain binary search exactly; absent keys find no equal key in any window.
//
// Probes are count -> scan -> write: per tile of 2048 S tuples a count, a u64
// scan of those, then the write kernel re-probes and writes at tile offset +
// workgroup-scanned thread offsets.
#include <algorithm>

#include "device.hpp"
#include "kernels.hpp"

namespace eo {
namespace {

constexpr int kIdxNT = 256;
constexpr int kIdxU = 8;
constexpr int kIdxTile = kIdxNT * kIdxU;

// ---- hash index build ---------------------------------------------------------

__global__ void k_hidx_count(const uint2 *__restrict__ R, uint64_t n, int log2nb, uint32_t *cnt) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[bucket_hash((int32_t)R[i].x, log2nb)], 1u);
}

__global__ void k_hidx_place(const uint2 *__restrict__ R, uint64_t n, int log2nb, uint32_t *cursor, uint32_t *keys,
                             uint32_t *pays) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint2 t = R[i];
    const uint32_t pos = atomicAdd(&cursor[bucket_hash((int32_t)t.x, log2nb)], 1u);
    keys[pos] = t.x;
    pays[pos] = t.y;
  }
}

// ---- sorted index helpers --------------------------------------------------------

__global__ void k_split_soa(const uint2 *__restrict__ T, uint64_t n, uint32_t *keys, uint32_t *pays) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint2 t = T[i];
    keys[i] = t.x;
    pays[i] = t.y;
  }
}

__device__ __forceinline__ uint64_t lower_bound_i32(const uint32_t *__restrict__ keys, uint64_t lo, uint64_t hi,
                                                    int32_t k) {
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if ((int32_t)keys[mid] < k) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// ---- learned model ----------------------------------------------------------------

constexpr int kKnots = 256;     // root segments
constexpr int kLeaves = 4096;   // stage-2 linear models

struct LeafModel {
  float a, b;        // pos ~= a + b * (key - k0)
  int32_t k0;        // first key routed to the leaf
  uint32_t pad;
  int32_t err_lo;    // max (true - pred), >= 0
  int32_t err_hi;    // max (pred - true), >= 0
  uint32_t first, last;  // position range of the leaf's keys [first, last)
};

// Root: position estimate from the knot table (knot_key[K+1], knot_pos[K+1]).
__device__ __forceinline__ float root_pred(const int32_t *knot_key, const float *knot_pos, int32_t k) {
  int lo = 0, hi = kKnots;  // knot_key[lo] <= k < knot_key[hi] (clamped at the ends)
  if (k <= knot_key[0]) return knot_pos[0];
  if (k >= knot_key[kKnots]) return knot_pos[kKnots];
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (knot_key[mid] <= k) lo = mid; else hi = mid;
  }
  const float dk = (float)((int64_t)knot_key[hi] - (int64_t)knot_key[lo]);
  const float t = dk > 0.f ? (float)((int64_t)k - (int64_t)knot_key[lo]) / dk : 0.f;
  return knot_pos[lo] + t * (knot_pos[hi] - knot_pos[lo]);
}

__device__ __forceinline__ uint32_t leaf_of(float root, uint64_t n) {
  const float l = root * (float)kLeaves / (float)n;
  const int li = (int)l;
  return (uint32_t)min(max(li, 0), kLeaves - 1);
}

__device__ __forceinline__ int64_t leaf_pred(const LeafModel &m, int32_t k) {
  const float x = (float)((int64_t)k - (int64_t)m.k0);
  return (int64_t)floorf(m.a + m.b * x);
}

__global__ void k_knots(const uint32_t *__restrict__ keys, uint64_t n, int32_t *knot_key, float *knot_pos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > kKnots) return;
  uint64_t p = (uint64_t)((double)i * (double)(n - 1) / kKnots);
  knot_key[i] = (int32_t)keys[p];
  knot_pos[i] = (float)p;
}

// leaf id of every position (monotone in position because the root is)
__global__ void k_leaf_ids(const uint32_t *__restrict__ keys, uint64_t n, const int32_t *knot_key, const float *knot_pos,
                           uint16_t *leaf) {
  __shared__ int32_t kk[kKnots + 1];
  __shared__ float kp[kKnots + 1];
  for (int i = threadIdx.x; i <= kKnots; i += blockDim.x) {
    kk[i] = knot_key[i];
    kp[i] = knot_pos[i];
  }
  __syncthreads();
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    leaf[i] = (uint16_t)leaf_of(root_pred(kk, kp, (int32_t)keys[i]), n);
}

// One workgroup per leaf: locate its position range, least-squares fit in FP64,
// then the error bounds with the FP32 inference code.
__global__ __launch_bounds__(256) void k_fit_leaves(const uint32_t *__restrict__ keys, uint64_t n,
                                                    const uint16_t *__restrict__ leaf, LeafModel *models) {
  __shared__ double red[256 / kWave + 1];
  __shared__ uint64_t range[2];
  __shared__ float ab[2];
  const uint32_t l = blockIdx.x;
  if (threadIdx.x == 0) {
    uint64_t lo = 0, hi = n;  // first position with leaf >= l
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (leaf[mid] < l) lo = mid + 1; else hi = mid;
    }
    const uint64_t first = lo;
    hi = n;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (leaf[mid] <= l) lo = mid + 1; else hi = mid;
    }
    range[0] = first;
    range[1] = lo;
  }
  __syncthreads();
  const uint64_t first = range[0], last = range[1];
  LeafModel m{};
  m.first = (uint32_t)first;
  m.last = (uint32_t)last;
  if (first == last) {
    if (threadIdx.x == 0) {
      m.k0 = 0;
      m.a = (float)first;
      m.b = 0.f;
      m.err_lo = m.err_hi = 0;
      models[l] = m;
    }
    return;
  }
  const int32_t k0 = (int32_t)keys[first];
  double sx = 0, sy = 0, sxx = 0, sxy = 0;
  for (uint64_t i = first + threadIdx.x; i < last; i += blockDim.x) {
    const double x = (double)((int64_t)(int32_t)keys[i] - (int64_t)k0);
    const double y = (double)i;
    sx += x;
    sy += y;
    sxx += x * x;
    sxy += x * y;
  }
  sx = block_sum<256, double>(sx, red);
  sy = block_sum<256, double>(sy, red);
  sxx = block_sum<256, double>(sxx, red);
  sxy = block_sum<256, double>(sxy, red);
  if (threadIdx.x == 0) {
    const double cnt = (double)(last - first);
    const double den = cnt * sxx - sx * sx;
    double b = den > 0 ? (cnt * sxy - sx * sy) / den : 0.0;
    double a = (sy - b * sx) / cnt;
    ab[0] = (float)a;
    ab[1] = (float)b;
  }
  __syncthreads();
  m.k0 = k0;
  m.a = ab[0];
  m.b = ab[1];
  // error bounds against the first position of each key (lower_bound)
  int32_t elo = 0, ehi = 0;
  for (uint64_t i = first + threadIdx.x; i < last; i += blockDim.x) {
    if (i > 0 && keys[i - 1] == keys[i]) continue;  // not the first occurrence
    const int64_t pred = leaf_pred(m, (int32_t)keys[i]);
    const int64_t d = (int64_t)i - pred;
    elo = max(elo, (int32_t)min<int64_t>(d, 1 << 30));
    ehi = max(ehi, (int32_t)min<int64_t>(-d, 1 << 30));
  }
  __shared__ int32_t er[2];
  if (threadIdx.x == 0) er[0] = er[1] = 0;
  __syncthreads();
  atomicMax(&er[0], elo);
  atomicMax(&er[1], ehi);
  __syncthreads();
  if (threadIdx.x == 0) {
    m.err_lo = er[0];
    m.err_hi = er[1];
    models[l] = m;
  }
}

// ---- probes -------------------------------------------------------------------------

struct IndexView {
  int kind;               // 0 hash, 1 sorted (binary search), 2 sorted (learned)
  uint64_t n;
  int log2nb;
  const uint32_t *keys, *pays, *off;
  const int32_t *knot_key;
  const float *knot_pos;
  const LeafModel *leaves;
};

// [b, e) = index entries to compare with key k (hash: its bucket; sorted: from
// the lower bound of k).  For sorted indexes e is the end of the array and the
// caller stops at the first larger key.
template <int KIND>
__device__ __forceinline__ void locate(const IndexView &ix, const int32_t *kk, const float *kp, int32_t k,
                                       uint64_t *b, uint64_t *e) {
  if (KIND == 0) {
    const uint32_t h = bucket_hash(k, ix.log2nb);
    *b = ix.off[h];
    *e = ix.off[h + 1];
  } else if (KIND == 1) {
    *b = lower_bound_i32(ix.keys, 0, ix.n, k);
    *e = ix.n;
  } else {
    const LeafModel m = ix.leaves[leaf_of(root_pred(kk, kp, k), ix.n)];
    const int64_t pred = leaf_pred(m, k);
    const int64_t lo = max<int64_t>(0, pred - m.err_hi);
    const int64_t hi = min<int64_t>((int64_t)ix.n, pred + m.err_lo + 1);
    uint64_t pos = lo < hi ? lower_bound_i32(ix.keys, (uint64_t)lo, (uint64_t)hi, k) : (uint64_t)min<int64_t>(lo, ix.n);
    *b = pos;
    *e = ix.n;
  }
}

template <int KIND, bool WRITE>
__global__ __launch_bounds__(kIdxNT) void k_index_probe(IndexView ix, const uint2 *__restrict__ S, uint64_t nS,
                                                        uint64_t *__restrict__ tile_cnt, uint64_t *cursor,
                                                        uint2 *__restrict__ out, uint64_t out_cap,
                                                        uint64_t *__restrict__ lb_out, int wide, int outer) {
  __shared__ int32_t kk[KIND == 2 ? kKnots + 1 : 1];
  __shared__ float kp[KIND == 2 ? kKnots + 1 : 1];
  __shared__ uint64_t red[kIdxNT / kWave + 1];
  __shared__ uint64_t base_sh;
  if (KIND == 2) {
    for (int i = threadIdx.x; i <= kKnots; i += kIdxNT) {
      kk[i] = ix.knot_key[i];
      kp[i] = ix.knot_pos[i];
    }
    __syncthreads();
  }
  const uint64_t t = blockIdx.x;
  const uint64_t s0 = t * kIdxTile;
  uint2 s[kIdxU];
  uint64_t b[kIdxU];
  uint32_t m[kIdxU];  // one probe's matches: at most the index size (< 2^31)
  uint64_t total = 0;  // eight probes' matches can pass 2^32
  uint32_t nul = 0;
#pragma unroll
  for (int u = 0; u < kIdxU; ++u) {
    const uint64_t i = s0 + (uint64_t)u * kIdxNT + threadIdx.x;
    m[u] = 0;
    b[u] = 0;
    if (i < nS) s[u] = S[i];
  }
#pragma unroll
  for (int u = 0; u < kIdxU; ++u) {
    const uint64_t i = s0 + (uint64_t)u * kIdxNT + threadIdx.x;
    if (i >= nS) continue;
    uint64_t e;
    locate<KIND>(ix, kk, kp, (int32_t)s[u].x, &b[u], &e);
    if (lb_out) lb_out[i] = b[u];
    if (KIND == 0) {
      for (uint64_t j = b[u]; j < e; ++j) m[u] += ix.keys[j] == s[u].x;
    } else {
      for (uint64_t j = b[u]; j < e && ix.keys[j] == s[u].x; ++j) ++m[u];
    }
    if (outer && m[u] == 0) nul |= 1u << u;  // left outer join: one null-padded row
    total += (uint64_t)m[u] + ((nul >> u) & 1u);
  }
  if (!WRITE) {
    const uint64_t tot = block_sum<kIdxNT, uint64_t>(total, red);
    if (threadIdx.x == 0) {
      if (tile_cnt) tile_cnt[t] = tot;
      else if (cursor && tot) atomic_add_agent(cursor, tot);
    }
    return;
  }
  // output offsets: scanned tile counts, or a range reserved with one atomic
  uint64_t tile_total;
  const uint64_t excl = block_excl_scan<kIdxNT, uint64_t>(total, red, &tile_total);
  uint64_t pos;
  if (tile_cnt) {
    pos = tile_cnt[t] + excl;
  } else {
    if (threadIdx.x == 0) base_sh = tile_total ? atomic_add_agent(cursor, tile_total) : 0ull;
    __syncthreads();
    pos = base_sh + excl;
  }
#pragma unroll
  for (int u = 0; u < kIdxU; ++u) {
    if ((nul >> u) & 1u) {
      if (pos < out_cap) emit_null(out, pos, s[u].y, s[u].x, wide != 0);
      ++pos;
      continue;
    }
    if (m[u] == 0) continue;
    uint64_t j = b[u];
    for (uint32_t c = 0; c < m[u]; ++j) {
      if (ix.keys[j] != s[u].x) continue;
      if (pos < out_cap) emit_pair(out, pos, ix.pays[j], s[u].y, s[u].x, wide != 0);
      ++pos;
      ++c;
    }
  }
}

template <int KIND>
hipError_t launch_probe(const IndexView &ix, bool write, const eo_tuple *S, uint64_t nS, uint64_t *tile_cnt,
                        uint64_t *cursor, eo_pair *out, uint64_t out_cap, uint64_t *lb_out, bool wide,
                        bool outer, hipStream_t stream) {
  const uint64_t tiles = (nS + kIdxTile - 1) / kIdxTile;
  if (tiles == 0) return hipSuccess;
  if (tiles >= (1ull << 31)) return hipErrorInvalidValue;
  if (write)
    EO_TIMED("index_write", stream,
             hipLaunchKernelGGL((k_index_probe<KIND, true>), dim3((uint32_t)tiles), dim3(kIdxNT), 0, stream, ix,
                                reinterpret_cast<const uint2 *>(S), nS, tile_cnt, cursor,
                                reinterpret_cast<uint2 *>(out), out_cap, lb_out, (int)wide, (int)outer));
  else
    EO_TIMED("index_probe", stream,
             hipLaunchKernelGGL((k_index_probe<KIND, false>), dim3((uint32_t)tiles), dim3(kIdxNT), 0, stream, ix,
                                reinterpret_cast<const uint2 *>(S), nS, tile_cnt, cursor, nullptr, 0, lb_out, 0,
                                (int)outer));
  return hipGetLastError();
}

// *flag = 1 if some adjacent pair is out of (signed) key order.
__global__ void k_check_sorted(const uint2 *__restrict__ in, uint64_t n, uint32_t *flag) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i + 1 < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    if ((int32_t)in[i].x > (int32_t)in[i + 1].x) *flag = 1u;
}

uint32_t grid_for(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 16384)); }

}  // namespace

uint64_t index_probe_tiles(uint64_t nS) { return (nS + kIdxTile - 1) / kIdxTile; }
int learned_leaves() { return kLeaves; }
int learned_knots() { return kKnots; }
size_t learned_model_bytes() {
  return (kKnots + 1) * (sizeof(int32_t) + sizeof(float)) + kLeaves * sizeof(LeafModel) + 64;
}

hipError_t hidx_build(const eo_tuple *R, uint64_t n, int log2nb, uint32_t *keys, uint32_t *pays, uint32_t *off,
                      uint32_t *cursor, void *scan_scratch, uint32_t *err, hipStream_t stream) {
  const uint64_t nb = 1ull << log2nb;
  hipError_t e = fill_u32(off, 0u, nb + 1, stream);
  if (e != hipSuccess) return e;
  const uint2 *vR = reinterpret_cast<const uint2 *>(R);
  hipLaunchKernelGGL(k_hidx_count, dim3(grid_for(n)), dim3(256), 0, stream, vR, n, log2nb, off);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = scan_u32(off, off, nb + 1, nullptr, scan_scratch, err, stream)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(cursor, off, nb * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_hidx_place, dim3(grid_for(n)), dim3(256), 0, stream, vR, n, log2nb, cursor, keys, pays);
  return hipGetLastError();
}

hipError_t check_sorted(const eo_tuple *in, uint64_t n, uint32_t *flag, hipStream_t stream) {
  hipError_t e = fill_u32(flag, 0u, 1, stream);
  if (e != hipSuccess || n < 2) return e;
  hipLaunchKernelGGL(k_check_sorted, dim3(grid_for(n)), dim3(256), 0, stream, reinterpret_cast<const uint2 *>(in), n,
                     flag);
  return hipGetLastError();
}

hipError_t sidx_split(const eo_tuple *sorted, uint64_t n, uint32_t *keys, uint32_t *pays, hipStream_t stream) {
  hipLaunchKernelGGL(k_split_soa, dim3(grid_for(n)), dim3(256), 0, stream, reinterpret_cast<const uint2 *>(sorted), n,
                     keys, pays);
  return hipGetLastError();
}

hipError_t learned_fit(const uint32_t *keys, uint64_t n, void *model, uint16_t *leaf_scratch, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  char *mp = (char *)model;
  int32_t *knot_key = (int32_t *)mp;
  float *knot_pos = (float *)(mp + (kKnots + 1) * sizeof(int32_t));
  LeafModel *leaves = (LeafModel *)(mp + ((kKnots + 1) * (sizeof(int32_t) + sizeof(float)) + 63) / 64 * 64);
  hipLaunchKernelGGL(k_knots, dim3(2), dim3(256), 0, stream, keys, n, knot_key, knot_pos);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_leaf_ids, dim3(grid_for(n)), dim3(256), 0, stream, keys, n, knot_key, knot_pos, leaf_scratch);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_fit_leaves, dim3(kLeaves), dim3(256), 0, stream, keys, n, leaf_scratch, leaves);
  return hipGetLastError();
}

hipError_t index_probe(int kind, bool write, uint64_t n, int log2nb, const uint32_t *keys, const uint32_t *pays,
                       const uint32_t *off, const void *model, const eo_tuple *S, uint64_t nS,
                       uint64_t *tile_cnt, uint64_t *cursor, eo_pair *out, uint64_t out_cap, uint64_t *lb_out,
                       bool wide, bool outer, hipStream_t stream) {
  IndexView ix{};
  ix.kind = kind;
  ix.n = n;
  ix.log2nb = log2nb;
  ix.keys = keys;
  ix.pays = pays;
  ix.off = off;
  if (model) {
    const char *mp = (const char *)model;
    ix.knot_key = (const int32_t *)mp;
    ix.knot_pos = (const float *)(mp + (kKnots + 1) * sizeof(int32_t));
    ix.leaves = (const LeafModel *)(mp + ((kKnots + 1) * (sizeof(int32_t) + sizeof(float)) + 63) / 64 * 64);
  }
  switch (kind) {
    case 0: return launch_probe<0>(ix, write, S, nS, tile_cnt, cursor, out, out_cap, lb_out, wide, outer, stream);
    case 1: return launch_probe<1>(ix, write, S, nS, tile_cnt, cursor, out, out_cap, lb_out, wide, outer, stream);
    case 2: return launch_probe<2>(ix, write, S, nS, tile_cnt, cursor, out, out_cap, lb_out, wide, outer, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace eo
