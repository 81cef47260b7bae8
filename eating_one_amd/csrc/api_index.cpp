// C ABI of the resident indexes: hash index, sorted index (binary search) and
// learned sorted index.  Replaces the reference's createIndex + join of
// joins/innerJoin_gpu/main_street/inner-idx-bnlj.cc:149-673 (hash) and
// joins/innerJoin_gpu/sort_index/main_street/inner-idx-bnlj.cc:175-658
// (sorted), whose indexes were built on the CPU and copied to the GPU per join.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "eojoin.h"
#include "kernels.hpp"

struct eo_index {
  eo_ctx *ctx = nullptr;
  int kind = 0;        // EO_INDEX_HASH / EO_INDEX_SORTED
  int learned = 0;
  uint64_t n = 0;
  int log2nb = 0;
  uint32_t *keys = nullptr, *pays = nullptr, *off = nullptr;
  void *model = nullptr;
  uint64_t bytes = 0;
  double build_ms = 0.0;
  int32_t max_err = 0;
  double mean_err = 0.0;
};

namespace {

using namespace eo;

eo_status alloc_dev(void **p, size_t bytes, uint64_t *acc) {
  EO_TRY_HIP(hipMalloc(p, std::max<size_t>(bytes, 256)), "index allocation");
  *acc += bytes;
  return EO_OK;
}

void free_index(eo_index *ix) {
  if (!ix) return;
  if (ix->keys) (void)hipFree(ix->keys);
  if (ix->pays) (void)hipFree(ix->pays);
  if (ix->off) (void)hipFree(ix->off);
  if (ix->model) (void)hipFree(ix->model);
  delete ix;
}

struct LeafHost {
  float a, b;
  int32_t k0;
  uint32_t pad;
  int32_t err_lo, err_hi;
  uint32_t first, last;
};

// Common input staging: R on the device (host mode copies into the arena).
const eo_tuple *stage_input(eo_ctx *c, const eo_tuple *R, uint64_t n, bool host, char *dst, hipStream_t s,
                            eo_status *st) {
  *st = EO_OK;
  if (!host) return R;
  hipError_t e = hipMemcpyAsync(dst, R, n * sizeof(eo_tuple), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) *st = hip_fail(e, "h2d R");
  return (const eo_tuple *)dst;
}

}  // namespace

extern "C" {

eo_status eo_hidx_build(eo_ctx *c, const eo_tuple *R, uint64_t nR, const eo_join_opts *o, eo_index **out) {
  if (!c || !out || !opts_ok(o) || nR > kMaxRows || (nR && !R)) return EO_EINVAL;
  *out = nullptr;
  DeviceGuard g(c->device);
  hipStream_t s = pick_stream(c, o);
  const bool host = where_of(o) == EO_HOST;
  int log2nb = 0;
  while ((1ull << log2nb) < std::max<uint64_t>(nR, 1)) ++log2nb;
  const uint64_t nb = 1ull << log2nb;
  eo_index *ix = new eo_index();
  ix->ctx = c;
  ix->kind = EO_INDEX_HASH;
  ix->n = nR;
  ix->log2nb = log2nb;
  eo_status st = alloc_dev((void **)&ix->keys, nR * 4, &ix->bytes);
  if (st == EO_OK) st = alloc_dev((void **)&ix->pays, nR * 4, &ix->bytes);
  if (st == EO_OK) st = alloc_dev((void **)&ix->off, (nb + 1) * 4, &ix->bytes);
  if (st != EO_OK) {
    free_index(ix);
    return st;
  }
  Layout L;
  const size_t o_in = L.add(host ? nR * sizeof(eo_tuple) : 0);
  const size_t o_cur = L.add(nb * 4);
  const size_t o_scr = L.add(scan_scratch_bytes_u32(nb + 1));
  hipError_t e = hipStreamSynchronize(s);
  if (e == hipSuccess) e = c->arena.reserve(L.total);
  if (e != hipSuccess) {
    free_index(ix);
    return hip_fail(e, "arena");
  }
  char *A = c->arena.base;
  const eo_tuple *dR = stage_input(c, R, nR, host, A + o_in, s, &st);
  if (st != EO_OK) {
    free_index(ix);
    return st;
  }
  PhaseClock clk(c, s, true);
  clk.mark();
  e = hidx_build(dR, nR, log2nb, ix->keys, ix->pays, ix->off, (uint32_t *)(A + o_cur), A + o_scr, c->d_err, s);
  clk.mark();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    free_index(ix);
    return hip_fail(e, "hash index build");
  }
  ix->build_ms = clk.ms(0, 1);
  if (o && o->stats) {
    std::memset(o->stats, 0, sizeof(*o->stats));
    o->stats->ms_build = ix->build_ms;
    o->stats->ms_total = ix->build_ms;
  }
  st = check_device_errors(c, s);
  if (st != EO_OK) {
    free_index(ix);
    return st;
  }
  *out = ix;
  return EO_OK;
}

// presorted: R is already sorted by key (an imported index): checked on the
// device, no radix sort.
static eo_status sidx_build_impl(eo_ctx *c, const eo_tuple *R, uint64_t nR, int learned, const eo_join_opts *o,
                                 eo_index **out, bool presorted) {
  if (!c || !out || !opts_ok(o) || nR > kMaxRows || (nR && !R)) return EO_EINVAL;
  *out = nullptr;
  DeviceGuard g(c->device);
  hipStream_t s = pick_stream(c, o);
  const bool host = where_of(o) == EO_HOST;
  eo_index *ix = new eo_index();
  ix->ctx = c;
  ix->kind = EO_INDEX_SORTED;
  ix->learned = learned ? 1 : 0;
  ix->n = nR;
  eo_status st = alloc_dev((void **)&ix->keys, nR * 4, &ix->bytes);
  if (st == EO_OK) st = alloc_dev((void **)&ix->pays, nR * 4, &ix->bytes);
  if (st == EO_OK && learned) st = alloc_dev(&ix->model, learned_model_bytes(), &ix->bytes);
  if (st != EO_OK) {
    free_index(ix);
    return st;
  }
  // stable LSD radix sort on the sign-flipped key: 4 passes of 8 bits
  PartPassPlan pp[4];
  size_t pscr = 0;
  for (int k = 0; k < 4; ++k) {
    pp[k] = plan_partition_pass(nR, 8, 8 * k, 0x80000000u, c->props.cus);
    pscr = std::max(pscr, partition_pass_scratch_bytes(pp[k]));
  }
  Layout L;
  const size_t o_in = L.add(host ? nR * sizeof(eo_tuple) : 0);
  const size_t o_a = L.add(nR * sizeof(eo_tuple));
  const size_t o_b = L.add(nR * sizeof(eo_tuple));
  const size_t o_leaf = L.add(learned ? nR * sizeof(uint16_t) : 0);
  const size_t o_scr = L.add(pscr);
  const size_t o_flag = L.add(sizeof(uint64_t));
  hipError_t e = hipStreamSynchronize(s);
  if (e == hipSuccess) e = c->arena.reserve(L.total);
  if (e != hipSuccess) {
    free_index(ix);
    return hip_fail(e, "arena");
  }
  char *A = c->arena.base;
  const eo_tuple *dR = stage_input(c, R, nR, host, A + o_in, s, &st);
  if (st != EO_OK) {
    free_index(ix);
    return st;
  }
  PhaseClock clk(c, s, true);
  clk.mark();
  eo_tuple *bufa = (eo_tuple *)(A + o_a), *bufb = (eo_tuple *)(A + o_b);
  const eo_tuple *src = dR;
  e = hipSuccess;
  if (nR && presorted) {
    uint32_t *flag = (uint32_t *)(A + o_flag);
    e = check_sorted(dR, nR, flag, s);
    uint64_t unsorted = 0;
    if (e == hipSuccess) {
      st = read_u64(c, (const uint64_t *)flag, &unsorted, 1, s);
      if (st != EO_OK) {
        free_index(ix);
        return st;
      }
      if ((uint32_t)unsorted) {
        free_index(ix);
        set_last_error("eo_sidx_import: input is not sorted by key");
        return EO_EINVAL;
      }
    }
  }
  if (nR) {
    for (int k = 0; k < 4 && e == hipSuccess && !presorted; ++k) {
      eo_tuple *dst = (k % 2 == 0) ? bufa : bufb;
      e = partition_pass(pp[k], src, dst, nullptr, A + o_scr, c->d_err, s);
      src = dst;
    }
    if (e == hipSuccess) e = sidx_split(src, nR, ix->keys, ix->pays, s);
    if (e == hipSuccess && learned) e = learned_fit(ix->keys, nR, ix->model, (uint16_t *)(A + o_leaf), s);
  }
  clk.mark();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    free_index(ix);
    return hip_fail(e, "sorted index build");
  }
  ix->build_ms = clk.ms(0, 1);
  if (learned && nR) {
    std::vector<LeafHost> lv(learned_leaves());
    const size_t head = ((learned_knots() + 1) * (sizeof(int32_t) + sizeof(float)) + 63) / 64 * 64;
    e = hipMemcpy(lv.data(), (char *)ix->model + head, lv.size() * sizeof(LeafHost), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      free_index(ix);
      return hip_fail(e, "model readback");
    }
    double sum = 0;
    uint64_t used = 0;
    for (const auto &m : lv) {
      if (m.first == m.last) continue;
      ix->max_err = std::max(ix->max_err, std::max(m.err_lo, m.err_hi));
      sum += (double)m.err_lo + m.err_hi;
      ++used;
    }
    ix->mean_err = used ? sum / used : 0.0;
  }
  if (o && o->stats) {
    std::memset(o->stats, 0, sizeof(*o->stats));
    o->stats->ms_build = ix->build_ms;
    o->stats->ms_total = ix->build_ms;
  }
  st = check_device_errors(c, s);
  if (st != EO_OK) {
    free_index(ix);
    return st;
  }
  *out = ix;
  return EO_OK;
}

eo_status eo_sidx_build(eo_ctx *c, const eo_tuple *R, uint64_t nR, int learned, const eo_join_opts *o,
                        eo_index **out) {
  return sidx_build_impl(c, R, nR, learned, o, out, false);
}

eo_status eo_sidx_import(eo_ctx *c, const eo_tuple *sorted, uint64_t n, int learned, const eo_join_opts *o,
                         eo_index **out) {
  return sidx_build_impl(c, sorted, n, learned, o, out, true);
}

eo_status eo_index_query(const eo_index *ix, eo_index_info *info) {
  if (!ix || !info) return EO_EINVAL;
  std::memset(info, 0, sizeof(*info));
  info->kind = ix->kind;
  info->learned = ix->learned;
  info->n = ix->n;
  info->bytes = ix->bytes;
  info->leaves = ix->learned ? learned_leaves() : 0;
  info->max_err = ix->max_err;
  info->mean_err = ix->mean_err;
  info->build_ms = ix->build_ms;
  return EO_OK;
}

eo_status eo_index_free(eo_index *ix) {
  if (!ix) return EO_EINVAL;
  DeviceGuard g(ix->ctx->device);
  (void)hipStreamSynchronize(ix->ctx->stream);
  free_index(ix);
  return EO_OK;
}

static int probe_kind(const eo_index *ix, int method) {
  if (ix->kind == EO_INDEX_HASH) return method == EO_LOOKUP_DEFAULT ? 0 : -1;
  if (method == EO_LOOKUP_BINARY) return 1;
  if (method == EO_LOOKUP_LEARNED) return ix->learned ? 2 : -1;
  return ix->learned ? 2 : 1;
}

eo_status eo_index_join(eo_ctx *c, const eo_index *ix, const eo_tuple *S, uint64_t nS, eo_pair *out,
                        uint64_t out_cap, uint64_t *n_out, const eo_join_opts *o) {
  if (!c || !ix || !n_out || !opts_ok(o) || nS > kMaxRows || (nS && !S) || (out_cap && !out)) return EO_EINVAL;
  *n_out = 0;
  const int method = (o && (o->flags & EO_F_BINARY_SEARCH)) ? EO_LOOKUP_BINARY : EO_LOOKUP_DEFAULT;
  const int kind = probe_kind(ix, method);
  if (kind < 0) return EO_EINVAL;
  eo_join_stats *stats = o ? o->stats : nullptr;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  const bool outer = o && (o->flags & EO_F_LEFT_OUTER);
  if (nS == 0 || (ix->n == 0 && !outer)) return EO_OK;
  const bool wide = o && (o->flags & EO_F_RESULT16);
  const size_t row = wide ? 16 : sizeof(eo_pair);
  DeviceGuard g(c->device);
  hipStream_t s = pick_stream(c, o);
  const bool host = where_of(o) == EO_HOST;
  const uint64_t tiles = index_probe_tiles(nS);
  Layout L;
  const size_t o_in = L.add(host ? nS * sizeof(eo_tuple) : 0);
  const size_t o_cnt = L.add(tiles * sizeof(uint64_t));
  const size_t o_small = L.add(64);
  const size_t o_scr = L.add(scan_scratch_bytes_u64(tiles));
  const bool empty_outer = outer && ix->n == 0;  // every probe unmatched
  const size_t o_bits = L.add(empty_outer ? matched_words(nS) * sizeof(uint32_t) : 0);
  EO_TRY_HIP(hipStreamSynchronize(s), "sync");
  EO_TRY_HIP(c->arena.reserve(L.total), "arena");
  char *A = c->arena.base;
  eo_status st;
  const eo_tuple *dS = stage_input(c, S, nS, host, A + o_in, s, &st);
  if (st != EO_OK) return st;
  uint64_t *cnt = (uint64_t *)(A + o_cnt);
  uint64_t *small = (uint64_t *)(A + o_small);
  PhaseClock clk(c, s, stats != nullptr);
  ScopedTimer tmr(c, stats);
  clk.mark();
  if (empty_outer && (o->flags & EO_F_DETERMINISTIC)) {
    set_last_error("EO_F_LEFT_OUTER on an empty index: default order only");
    return EO_EINVAL;
  }
  if (!(o && (o->flags & EO_F_DETERMINISTIC))) {
    // one lookup pass; each tile reserves its output range with one atomic
    // (with out == NULL, or a too small buffer, it counts every match)
    eo_pair *dout = out;
    bool staged = false;
    if (host && out && out_cap) {
      EO_TRY_HIP(hipMallocAsync((void **)&dout, out_cap * row, s), "output staging");
      staged = true;
    }
    EO_TRY_HIP(fill_u64(small, 0, 1, s), "cursor reset");
    hipError_t e = hipSuccess;
    if (empty_outer) {
      uint32_t *none = (uint32_t *)(A + o_bits);
      e = fill_u32(none, 0u, matched_words(nS), s);
      if (e == hipSuccess) e = emit_unmatched(dS, nS, none, small, out ? dout : nullptr, out ? out_cap : 0, wide, s);
    } else {
      e = index_probe(kind, out != nullptr, ix->n, ix->log2nb, ix->keys, ix->pays, ix->off, ix->model, dS, nS,
                      nullptr, small, out ? dout : nullptr, out ? out_cap : 0, nullptr, wide, outer, s);
    }
    clk.mark();
    uint64_t total = 0;
    if (e == hipSuccess) {
      st = read_u64(c, small, &total, 1, s);
      if (st == EO_OK && staged && total)  // on overflow: the out_cap rows that were written
        e = hipMemcpyAsync(out, dout, std::min(total, out_cap) * row, hipMemcpyDeviceToHost, s);
    }
    if (staged) (void)hipFreeAsync(dout, s);
    if (e != hipSuccess) return hip_fail(e, "index join");
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
  EO_TRY_HIP(index_probe(kind, false, ix->n, ix->log2nb, ix->keys, ix->pays, ix->off, ix->model, dS, nS, cnt, nullptr,
                         nullptr, 0, nullptr, false, outer, s),
             "index count");
  EO_TRY_HIP(scan_u64(cnt, cnt, tiles, small, A + o_scr, c->d_err, s), "index scan");
  clk.mark();
  uint64_t total = 0;
  st = read_u64(c, small, &total, 1, s);
  if (st == EO_OK) st = check_device_errors(c, s);
  if (st != EO_OK) return st;
  *n_out = total;
  if (stats) stats->n_out = total;
  if (!out) return EO_OK;
  if (total > out_cap) return EO_ECAPACITY;
  if (total) {
    eo_pair *dout = out;
    if (host) EO_TRY_HIP(hipMallocAsync((void **)&dout, total * row, s), "output staging");
    hipError_t e = index_probe(kind, true, ix->n, ix->log2nb, ix->keys, ix->pays, ix->off, ix->model, dS, nS, cnt,
                               nullptr, dout, total, nullptr, wide, outer, s);
    clk.mark();
    if (e == hipSuccess && host) e = hipMemcpyAsync(out, dout, total * row, hipMemcpyDeviceToHost, s);
    if (host) (void)hipFreeAsync(dout, s);
    if (e != hipSuccess) return hip_fail(e, "index write");
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

eo_status eo_index_lookup(eo_ctx *c, const eo_index *ix, const eo_tuple *S, uint64_t nS, int method,
                          uint64_t *pos, const eo_join_opts *o) {
  if (!c || !ix || !opts_ok(o) || nS > kMaxRows || (nS && (!S || !pos))) return EO_EINVAL;
  const int kind = probe_kind(ix, method);
  if (kind < 0) return EO_EINVAL;
  if (nS == 0 || ix->n == 0) {
    if (nS && where_of(o) == EO_HOST) std::fill(pos, pos + nS, 0ull);
    return EO_OK;
  }
  DeviceGuard g(c->device);
  hipStream_t s = pick_stream(c, o);
  const bool host = where_of(o) == EO_HOST;
  Layout L;
  const size_t o_in = L.add(host ? nS * sizeof(eo_tuple) : 0);
  const size_t o_pos = L.add(host ? nS * sizeof(uint64_t) : 0);
  EO_TRY_HIP(hipStreamSynchronize(s), "sync");
  EO_TRY_HIP(c->arena.reserve(L.total), "arena");
  char *A = c->arena.base;
  eo_status st;
  const eo_tuple *dS = stage_input(c, S, nS, host, A + o_in, s, &st);
  if (st != EO_OK) return st;
  uint64_t *dpos = host ? (uint64_t *)(A + o_pos) : pos;
  PhaseClock clk(c, s, o && o->stats);
  ScopedTimer tmr(c, o ? o->stats : nullptr);
  clk.mark();
  EO_TRY_HIP(index_probe(kind, false, ix->n, ix->log2nb, ix->keys, ix->pays, ix->off, ix->model, dS, nS, nullptr, nullptr,
                         nullptr, 0, dpos, false, false, s),
             "index lookup");
  clk.mark();
  if (host) EO_TRY_HIP(hipMemcpyAsync(pos, dpos, nS * sizeof(uint64_t), hipMemcpyDeviceToHost, s), "d2h");
  EO_TRY_HIP(hipStreamSynchronize(s), "sync");
  if (o && o->stats) {
    std::memset(o->stats, 0, sizeof(*o->stats));
    o->stats->ms_count = clk.ms(0, 1);
    o->stats->ms_total = clk.ms(0, 1);
    tmr.finish();
  }
  return check_device_errors(c, s);
}

eo_status eo_index_export(eo_ctx *c, const eo_index *ix, uint32_t *keys, uint32_t *pays) {
  if (!c || !ix || (ix->n && (!keys || !pays))) return EO_EINVAL;
  DeviceGuard g(c->device);
  if (ix->n) {
    EO_TRY_HIP(hipMemcpy(keys, ix->keys, ix->n * 4, hipMemcpyDeviceToHost), "export keys");
    EO_TRY_HIP(hipMemcpy(pays, ix->pays, ix->n * 4, hipMemcpyDeviceToHost), "export payloads");
  }
  return EO_OK;
}

}  // extern "C"
