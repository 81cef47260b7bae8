/*
 * eo_join — command-line driver with the reference binaries' argument order.
 *
 *   eo_join <op> <|S|> <|R|> [seed] [selectivity]
 *     op: hash | nlj | hidx | sidx | learned | stream-hash | stream-nlj
 *
 * stream-*: the inputs are held in pinned host memory and joined by
 * eo_stream_join tile by tile (R and S tiles of 2^18 tuples, the reference
 * NLJ's PART of joins/blockJoin_gpu/main_street/tuple.h:13), the rows copied
 * back to host memory.
 *
 * Replaces `./HashExe <left=|S|> <right=|R|>` (joins/hashJoin_gpu/main_street/
 * hj.cc:1277-1290) and the NLJ / index executables of the same form
 * (joins/blockJoin_gpu/main_street/nestLoopJoin.cc:663-679,
 * joins/innerJoin_gpu/main_street/inner-idx-bnlj.cc:675-695).  Like the
 * reference it generates its own inputs (here seeded and on the device:
 * R keys a permutation of [0, |R|), a fraction `selectivity` of the S keys
 * drawn from them), runs one join, and prints the match count, the timings
 * and the first three result pairs (the reference prints them at
 * hj.cc:1245-1249).  Plain C over the library's C ABI.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "eojoin.h"

static int fail(const char *what, eo_status st) {
  fprintf(stderr, "eo_join: %s: %s (%s)\n", what, eo_status_string(st), eo_last_error());
  return 1;
}

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s <hash|nlj|hidx|sidx|learned|stream-hash|stream-nlj> <|S|> <|R|> [seed] [selectivity]\n",
            argv[0]);
    return 2;
  }
  const char *op = argv[1];
  const uint64_t nS = strtoull(argv[2], NULL, 10);
  const uint64_t nR = strtoull(argv[3], NULL, 10);
  const uint64_t seed = argc > 4 ? strtoull(argv[4], NULL, 0) : 0xE01C1ull;
  const double sel = argc > 5 ? atof(argv[5]) : 1.0;
  int kind = -1;
  if (!strcmp(op, "hash")) kind = 0;
  else if (!strcmp(op, "nlj")) kind = 1;
  else if (!strcmp(op, "hidx")) kind = 2;
  else if (!strcmp(op, "sidx")) kind = 3;
  else if (!strcmp(op, "learned")) kind = 4;
  else if (!strcmp(op, "stream-hash")) kind = 5;
  else if (!strcmp(op, "stream-nlj")) kind = 6;
  if (kind < 0) {
    fprintf(stderr, "eo_join: unknown op '%s'\n", op);
    return 2;
  }
  eo_ctx *ctx = NULL;
  eo_status st = eo_ctx_create(&ctx, 0, NULL);
  if (st != EO_OK) return fail("eo_ctx_create", st);
  eo_tuple *R = NULL, *S = NULL;
  eo_pair *out = NULL;
  if ((st = eo_dev_alloc(ctx, nR * sizeof(eo_tuple), (void **)&R)) != EO_OK) return fail("alloc R", st);
  if ((st = eo_dev_alloc(ctx, nS * sizeof(eo_tuple), (void **)&S)) != EO_OK) return fail("alloc S", st);
  eo_join_opts o;
  memset(&o, 0, sizeof(o));
  o.struct_size = sizeof(o);
  o.where = EO_DEVICE;
  if ((st = eo_gen_build(ctx, R, nR, seed, &o)) != EO_OK) return fail("eo_gen_build", st);
  if ((st = eo_gen_probe(ctx, S, nS, nR, sel, seed + 1, &o)) != EO_OK) return fail("eo_gen_probe", st);

  if (kind >= 5) {  /* host-resident inputs, streamed tile by tile */
    const uint64_t part = 1ull << 18;
    eo_tuple *hR = NULL, *hS = NULL;
    eo_pair *hO = NULL;
    if ((st = eo_host_alloc(ctx, (nR ? nR : 1) * sizeof(eo_tuple), (void **)&hR)) != EO_OK) return fail("host R", st);
    if ((st = eo_host_alloc(ctx, (nS ? nS : 1) * sizeof(eo_tuple), (void **)&hS)) != EO_OK) return fail("host S", st);
    if ((st = eo_copy_d2h(ctx, hR, R, nR * sizeof(eo_tuple))) != EO_OK) return fail("copy R", st);
    if ((st = eo_copy_d2h(ctx, hS, S, nS * sizeof(eo_tuple))) != EO_OK) return fail("copy S", st);
    eo_stream_stats ss;
    eo_stream_opts so;
    memset(&so, 0, sizeof(so));
    so.struct_size = sizeof(so);
    so.op = kind == 5 ? EO_STREAM_HASH : EO_STREAM_NLJ;
    so.r_tile = part;
    so.s_chunk = part;
    uint64_t n = 0, n2 = 0;
    if ((st = eo_stream_join(ctx, hR, nR, hS, nS, NULL, 0, &n, &so)) != EO_OK) return fail("count", st);
    if ((st = eo_host_alloc(ctx, (n ? n : 1) * sizeof(eo_pair), (void **)&hO)) != EO_OK) return fail("host out", st);
    so.stats = &ss;
    if ((st = eo_stream_join(ctx, hR, nR, hS, nS, hO, n, &n2, &so)) != EO_OK) return fail("join", st);
    printf("op=%s |S|=%llu |R|=%llu seed=0x%llx selectivity=%.3f\n", op, (unsigned long long)nS,
           (unsigned long long)nR, (unsigned long long)seed, sel);
    printf("count: %llu\n", (unsigned long long)n2);
    printf("wall time: %.3f ms (%u R tiles x %u S chunks, %.1f MB in, %.1f MB out)\n", ss.ms_total, ss.r_tiles,
           ss.r_tiles ? ss.s_chunks / ss.r_tiles : 0, ss.h2d_bytes / 1e6, ss.d2h_bytes / 1e6);
    for (uint64_t i = 0; i < (n2 < 3 ? n2 : 3); ++i)
      printf("pair %llu: R payload %d, S payload %d\n", (unsigned long long)i, hO[i].r_payload, hO[i].s_payload);
    eo_host_free(ctx, hO);
    eo_host_free(ctx, hS);
    eo_host_free(ctx, hR);
    eo_dev_free(ctx, S);
    eo_dev_free(ctx, R);
    eo_ctx_destroy(ctx);
    return 0;
  }

  eo_index *ix = NULL;
  eo_join_stats stats;
  memset(&stats, 0, sizeof(stats));
  double build_ms = 0.0;
  if (kind >= 2) {
    eo_index_info info;
    st = kind == 2 ? eo_hidx_build(ctx, R, nR, &o, &ix) : eo_sidx_build(ctx, R, nR, kind == 4, &o, &ix);
    if (st != EO_OK) return fail("index build", st);
    eo_index_query(ix, &info);
    build_ms = info.build_ms;
  }
  uint64_t n = 0;
  /* two-phase sizing: count, allocate, write */
  if (kind == 0) st = eo_hash_join(ctx, R, nR, S, nS, NULL, 0, &n, &o);
  else if (kind == 1) st = eo_nlj_join(ctx, R, nR, S, nS, NULL, 0, &n, &o);
  else st = eo_index_join(ctx, ix, S, nS, NULL, 0, &n, &o);
  if (st != EO_OK) return fail("count", st);
  if ((st = eo_dev_alloc(ctx, (n ? n : 1) * sizeof(eo_pair), (void **)&out)) != EO_OK) return fail("alloc out", st);
  o.stats = &stats;
  uint64_t n2 = 0;
  if (kind == 0) st = eo_hash_join(ctx, R, nR, S, nS, out, n, &n2, &o);
  else if (kind == 1) st = eo_nlj_join(ctx, R, nR, S, nS, out, n, &n2, &o);
  else st = eo_index_join(ctx, ix, S, nS, out, n, &n2, &o);
  if (st != EO_OK) return fail("join", st);

  printf("op=%s |S|=%llu |R|=%llu seed=0x%llx selectivity=%.3f\n", op, (unsigned long long)nS,
         (unsigned long long)nR, (unsigned long long)seed, sel);
  printf("count: %llu\n", (unsigned long long)n2);
  printf("device time: %.3f ms (partition %.3f, count %.3f, write %.3f", stats.ms_total, stats.ms_partition,
         stats.ms_count, stats.ms_write);
  if (kind >= 2) printf(", index build %.3f", build_ms);
  printf(")\n");
  for (uint32_t k = 0; k < stats.n_kernels; ++k)
    printf("  %-20s %9.3f ms  %u launches\n", stats.kernels[k].name, stats.kernels[k].ms, stats.kernels[k].launches);
  eo_pair first[3];
  const uint64_t show = n2 < 3 ? n2 : 3;
  if (show && eo_copy_d2h(ctx, first, out, show * sizeof(eo_pair)) == EO_OK)
    for (uint64_t i = 0; i < show; ++i)
      printf("pair %llu: R payload %d, S payload %d\n", (unsigned long long)i, first[i].r_payload,
             first[i].s_payload);
  if (ix) eo_index_free(ix);
  eo_dev_free(ctx, out);
  eo_dev_free(ctx, S);
  eo_dev_free(ctx, R);
  eo_ctx_destroy(ctx);
  return 0;
}
