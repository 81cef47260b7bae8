/*
 * oracle/cpu_join.c — CPU restatement of the reference's CPU_JOIN algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the
 * smoke() entry and bench.py's cpu_baseline leg load it.  The library
 * (eating_one_amd/libeojoin.so) never links, loads or calls it.
 *
 * Tuple mapping: the reference TUPLE is {key = row id, val = join attribute};
 * here otuple.key is the join attribute and otuple.payload the row id, and a
 * result pair is (R payload, S payload) = (RESULT.rkey, RESULT.lkey).
 *
 * Each function cites the reference lines it restates.  Known defects of the
 * reference are fixed and listed at the function; every fix only removes
 * undefined behaviour or a field mix-up, never changes which pairs join.
 * Partition/bucket numbers use (uint32_t)key % n: identical to the
 * reference's int `val % n` for every key >= 0 (the reference has undefined
 * behaviour for negative keys: negative array index).
 *
 * Every join returns the total number of matches and writes the first
 * min(total, cap) pairs to out (out may be NULL with cap 0 to only count).
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int32_t key;
  int32_t payload;
} otuple;

typedef struct {
  int32_t r_payload;
  int32_t s_payload;
} opair;

static inline void emit(opair *out, uint64_t cap, uint64_t *n, int32_t r_payload, int32_t s_payload) {
  if (*n < cap) {
    out[*n].r_payload = r_payload;
    out[*n].s_payload = s_payload;
  }
  ++*n;
}

/* Sequential exclusive prefix sum: the definition the reference's scan test
 * checks against (scanExclusiveHost, joins/hashJoin_gpu/other/scan_test/
 * scan_gold.cpp:29-40).  Returns the total. */
uint64_t oracle_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n) {
  uint64_t run = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t v = in[i];
    out[i] = (uint32_t)run;
    run += v;
  }
  return run;
}

uint64_t oracle_scan_u64(const uint64_t *in, uint64_t *out, uint64_t n) {
  uint64_t run = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t v = in[i];
    out[i] = run;
    run += v;
  }
  return run;
}

/* Stable partition by (uint32)key % nparts: histogram, exclusive prefix sum,
 * scatter in input order — joins/CPU_JOIN/hash_cpu/hj.cc:132-153 (R) and
 * :160-181 (S); the same order the reference's on-disk partitioner writes
 * (joins/join_ori/hj.cc:60-87, its hash-part-* files).  part_start gets
 * nparts + 1 entries.  Returns 0, or -1 on allocation failure. */
int oracle_partition(const otuple *in, uint64_t n, uint32_t nparts, otuple *out, uint64_t *part_start) {
  uint64_t *count = (uint64_t *)calloc(nparts, sizeof(uint64_t));
  if (!count) return -1;
  for (uint64_t i = 0; i < n; ++i) count[(uint32_t)in[i].key % nparts]++;
  uint64_t run = 0;
  for (uint32_t p = 0; p < nparts; ++p) {
    part_start[p] = run;
    run += count[p];
    count[p] = 0;
  }
  part_start[nparts] = run;
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t p = (uint32_t)in[i].key % nparts;
    out[part_start[p] + count[p]++] = in[i];
  }
  free(count);
  return 0;
}

/* Radix-partition digit variant used by the GPU API: digit = ((uint32)key >>
 * shift) & (2^bits - 1).  For shift 0 this equals oracle_partition with
 * nparts = 2^bits. */
int oracle_partition_bits(const otuple *in, uint64_t n, int bits, int shift, otuple *out, uint64_t *part_start) {
  const uint32_t F = 1u << bits;
  uint64_t *count = (uint64_t *)calloc(F, sizeof(uint64_t));
  if (!count) return -1;
  for (uint64_t i = 0; i < n; ++i) count[((uint32_t)in[i].key >> shift) & (F - 1)]++;
  uint64_t run = 0;
  for (uint32_t p = 0; p < F; ++p) {
    part_start[p] = run;
    run += count[p];
    count[p] = 0;
  }
  part_start[F] = run;
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t p = ((uint32_t)in[i].key >> shift) & (F - 1);
    out[part_start[p] + count[p]++] = in[i];
  }
  free(count);
  return 0;
}

/* Hash join of joins/CPU_JOIN/hash_cpu/hj.cc:116-228: stable partition of R
 * and S into nparts (NB_BKTENT = 2^20, tuple.h:1), then per partition a
 * nested loop, S outer, R inner (hj.cc:216-228).
 * Fixes: the reference keeps result[JT_SIZE] (19.2 GB) on the stack (hj.cc:189)
 * and reads hrt[] with S's offsets and hlt[] with R's offsets when recording
 * a match (hj.cc:220-223); here the matched R and S tuples themselves are
 * recorded. */
uint64_t oracle_hash_join(const otuple *R, uint64_t nR, const otuple *S, uint64_t nS, uint32_t nparts, opair *out,
                          uint64_t cap) {
  if (nparts == 0) nparts = 1u << 20;
  otuple *hr = (otuple *)malloc((nR ? nR : 1) * sizeof(otuple));
  otuple *hs = (otuple *)malloc((nS ? nS : 1) * sizeof(otuple));
  uint64_t *rloc = (uint64_t *)malloc(((uint64_t)nparts + 1) * sizeof(uint64_t));
  uint64_t *sloc = (uint64_t *)malloc(((uint64_t)nparts + 1) * sizeof(uint64_t));
  uint64_t n = 0;
  if (!hr || !hs || !rloc || !sloc) {
    n = UINT64_MAX;
    goto done;
  }
  if (oracle_partition(R, nR, nparts, hr, rloc) || oracle_partition(S, nS, nparts, hs, sloc)) {
    n = UINT64_MAX;
    goto done;
  }
  for (uint32_t p = 0; p < nparts; ++p)
    for (uint64_t j = sloc[p]; j < sloc[p + 1]; ++j)
      for (uint64_t k = rloc[p]; k < rloc[p + 1]; ++k)
        if (hs[j].key == hr[k].key) emit(out, cap, &n, hr[k].payload, hs[j].payload);
done:
  free(hr);
  free(hs);
  free(rloc);
  free(sloc);
  return n;
}

/* Hash-index join of joins/CPU_JOIN/inner_cpu/inner-idx-bnlj.cc:142-185:
 * index build by counting sort of R into nbuckets buckets (NB_BUCKET = 2^27,
 * tuple.h:1): idxcount[] histogram, Buck_array[] bucket start (-1 if empty),
 * Bucket[] = {val, adr} in R order within a bucket; probe each S tuple in
 * order against its bucket. */
uint64_t oracle_hidx_join(const otuple *R, uint64_t nR, const otuple *S, uint64_t nS, uint32_t nbuckets, opair *out,
                          uint64_t cap) {
  if (nbuckets == 0) nbuckets = 1u << 27;
  uint32_t *idxcount = (uint32_t *)calloc(nbuckets, sizeof(uint32_t));
  int64_t *buck = (int64_t *)malloc((uint64_t)nbuckets * sizeof(int64_t));
  otuple *bucket = (otuple *)malloc((nR ? nR : 1) * sizeof(otuple)); /* {key = val, payload = adr} */
  uint64_t n = 0;
  if (!idxcount || !buck || !bucket) {
    n = UINT64_MAX;
    goto done;
  }
  for (uint64_t i = 0; i < nR; ++i) idxcount[(uint32_t)R[i].key % nbuckets]++;
  {
    int64_t count = 0;
    for (uint32_t b = 0; b < nbuckets; ++b) {
      buck[b] = idxcount[b] == 0 ? -1 : count;
      count += idxcount[b];
      idxcount[b] = 0;
    }
  }
  for (uint64_t i = 0; i < nR; ++i) {
    const uint32_t b = (uint32_t)R[i].key % nbuckets;
    bucket[buck[b] + idxcount[b]].key = R[i].key;
    bucket[buck[b] + idxcount[b]].payload = (int32_t)i;
    idxcount[b]++;
  }
  for (uint64_t j = 0; j < nS; ++j) {
    const uint32_t h = (uint32_t)S[j].key % nbuckets;
    for (uint32_t i = 0; i < idxcount[h]; ++i) {
      const otuple *e = &bucket[buck[h] + i];
      if (e->key == S[j].key) emit(out, cap, &n, R[e->payload].payload, S[j].payload);
    }
  }
done:
  free(idxcount);
  free(buck);
  free(bucket);
  return n;
}

/* Quicksort of {val, adr} by val: the Hoare scheme with middle pivot of
 * joins/CPU_JOIN/sortIndex_cpu/inner-idx-bnlj.cc:106-127, with the recursion
 * replaced by an explicit stack (the larger side is pushed, the smaller is
 * handled first) so that 2^26 keys cannot exhaust the call stack. */
static void hoare_qsort(otuple *b, int64_t lo0, int64_t hi0) {
  int64_t stack[128][2];
  int sp = 0;
  stack[sp][0] = lo0;
  stack[sp][1] = hi0;
  ++sp;
  while (sp) {
    --sp;
    int64_t p = stack[sp][0], q = stack[sp][1];
    while (p < q) {
      int64_t i = p, j = q;
      const int32_t pivot = b[(p + q) / 2].key;
      while (1) {
        while (b[i].key < pivot) i++;
        while (pivot < b[j].key) j--;
        if (i >= j) break;
        otuple t = b[i];
        b[i] = b[j];
        b[j] = t;
        i++;
        j--;
      }
      /* sub-ranges [p, i-1] and [j+1, q] as in the reference (:124-125) */
      int64_t l1 = p, h1 = i - 1, l2 = j + 1, h2 = q;
      if (h1 - l1 > h2 - l2) {
        int64_t tl = l1, th = h1;
        l1 = l2; h1 = h2; l2 = tl; h2 = th;
      }
      if (l2 < h2) {
        stack[sp][0] = l2;
        stack[sp][1] = h2;
        ++sp;
      }
      p = l1;
      q = h1;
    }
  }
}

/* Binary search of inner-idx-bnlj.cc:163-174: returns the last midpoint. */
static uint64_t ref_search(const otuple *b, int32_t num, uint64_t n) {
  int64_t l = 0, r = (int64_t)n - 1, m = 0;
  do {
    m = (l + r) / 2;
    if (num < b[m].key) r = m - 1; else l = m + 1;
  } while (l <= r && num != b[m].key);
  return (uint64_t)m;
}

/* Sorted-index join of joins/CPU_JOIN/sortIndex_cpu/inner-idx-bnlj.cc:
 * createIndex (:129-146: copy {val, adr}, quicksort, verify order), then per S
 * tuple a binary search and an expansion over equal keys, left from the hit
 * then right of it (:212-237).
 * Fixes: the right expansion read Bucket[right] past the end when the hit is
 * the last entry (:226-227), and an empty index read Bucket[0]; both are
 * bounds-checked here.  Returns UINT64_MAX if the sort check fails. */
uint64_t oracle_sidx_join(const otuple *R, uint64_t nR, const otuple *S, uint64_t nS, opair *out, uint64_t cap) {
  if (nR == 0) return 0;
  otuple *b = (otuple *)malloc(nR * sizeof(otuple)); /* {key = val, payload = adr} */
  if (!b) return UINT64_MAX;
  for (uint64_t i = 0; i < nR; ++i) {
    b[i].key = R[i].key;
    b[i].payload = (int32_t)i;
  }
  hoare_qsort(b, 0, (int64_t)nR - 1);
  for (uint64_t i = 1; i < nR; ++i)
    if (b[i - 1].key > b[i].key) {
      free(b);
      return UINT64_MAX;
    }
  uint64_t n = 0;
  for (uint64_t j = 0; j < nS; ++j) {
    const int32_t v = S[j].key;
    const uint64_t bidx = ref_search(b, v, nR);
    uint64_t x = bidx;
    while (b[x].key == v) {
      emit(out, cap, &n, R[b[x].payload].payload, S[j].payload);
      if (x == 0) break;
      x--;
    }
    x = bidx + 1;
    while (x < nR && b[x].key == v) {
      emit(out, cap, &n, R[b[x].payload].payload, S[j].payload);
      x++;
    }
  }
  free(b);
  return n;
}

/* Sorted-index build only: the sorted {key, row} index of createIndex
 * (sortIndex_cpu/inner-idx-bnlj.cc:129-146), for checking the GPU's sorted
 * index as a multiset per key. */
int oracle_sidx_build(const otuple *R, uint64_t nR, otuple *sorted) {
  for (uint64_t i = 0; i < nR; ++i) {
    sorted[i].key = R[i].key;
    sorted[i].payload = (int32_t)i;
  }
  if (nR) hoare_qsort(sorted, 0, (int64_t)nR - 1);
  for (uint64_t i = 1; i < nR; ++i)
    if (sorted[i - 1].key > sorted[i].key) return -1;
  return 0;
}

/* Nested-loop join of joins/CPU_JOIN/block_cpu/nestLoopJoin.cc:61-77:
 * S (left) outer, R (right) inner. */
uint64_t oracle_nlj_join(const otuple *R, uint64_t nR, const otuple *S, uint64_t nS, opair *out, uint64_t cap) {
  uint64_t n = 0;
  for (uint64_t i = 0; i < nS; ++i)
    for (uint64_t j = 0; j < nR; ++j)
      if (S[i].key == R[j].key) emit(out, cap, &n, R[j].payload, S[i].payload);
  return n;
}

/* ---- distance-predicate nested loop (FP64) -----------------------------------
 * joins/CPU_JOIN/3d_cpu/nestLoopJoin.cc: TUPLE {int id; double val[3]}
 * (:17-21 of its tuple section), predicate eval() (:73-81): with
 * d = r.val[i] - l.val[i], the sum d0*d0 + d1*d1 + d2*d2 accumulated left to
 * right, each product and sum rounded separately (this file is built with
 * -ffp-contract=off, as the reference's x86-64 build has no fused
 * multiply-add), compared with DISTANCE^2 by `<`.  join() (:84-101) loops
 * left (S) outside, right (R) inside and emits (lid, rid); here the pair is
 * (R id, S id) like every other join of this file. */
typedef struct {
  int32_t id;
  int32_t pad;
  double v[3];
} opoint3;

static int dist_match(const opoint3 *l, const opoint3 *r, double d2) {
  double acc = 0.0;
  for (int i = 0; i < 3; ++i) {
    const double t = r->v[i] - l->v[i];
    acc += t * t;
  }
  return acc < d2;
}

uint64_t oracle_nlj_dist_join(const opoint3 *R, uint64_t nR, const opoint3 *S, uint64_t nS, double distance,
                              opair *out, uint64_t cap) {
  const double d2 = distance * distance;
  uint64_t n = 0;
  for (uint64_t i = 0; i < nS; ++i)
    for (uint64_t j = 0; j < nR; ++j)
      if (dist_match(&S[i], &R[j], d2)) emit(out, cap, &n, R[j].id, S[i].id);
  return n;
}

/* ---- N-core hash join ------------------------------------------------------
 * The pthread hash join of joins/multi_joins/hash/thread-hash-bnlj.cc:146-501:
 * per-thread histograms over static slices (lcount_partitioning, :146), a
 * serial prefix (:393), per-thread scatter (lpartitioning, :181), then workers
 * pulling partitions from a shared counter (RTGetter, :292; executor, :326).
 * Fixes: results go to per-thread buffers instead of a mutex-guarded global
 * append (:343), and the scatter slice end is clamped to the relation's own
 * size (the reference clamps S slices to |R|, :195, :201). */

typedef struct {
  const otuple *in;
  otuple *out;
  uint64_t beg, end;
  uint32_t nparts;
  uint64_t *hist; /* nparts counts of this thread, then its cursors */
} part_task;

static void *part_count_thread(void *arg) {
  part_task *t = (part_task *)arg;
  for (uint64_t i = t->beg; i < t->end; ++i) t->hist[(uint32_t)t->in[i].key % t->nparts]++;
  return NULL;
}

static void *part_scatter_thread(void *arg) {
  part_task *t = (part_task *)arg;
  for (uint64_t i = t->beg; i < t->end; ++i) {
    const uint32_t p = (uint32_t)t->in[i].key % t->nparts;
    t->out[t->hist[p]++] = t->in[i];
  }
  return NULL;
}

static int partition_mt(const otuple *in, uint64_t n, uint32_t nparts, int nt, otuple *out, uint64_t *start) {
  pthread_t th[256];
  part_task tk[256];
  uint64_t *hist = (uint64_t *)calloc((uint64_t)nparts * nt, sizeof(uint64_t));
  if (!hist) return -1;
  for (int t = 0; t < nt; ++t) {
    tk[t].in = in;
    tk[t].out = out;
    tk[t].beg = n * t / nt;
    tk[t].end = n * (t + 1) / nt;
    tk[t].nparts = nparts;
    tk[t].hist = hist + (uint64_t)t * nparts;
    pthread_create(&th[t], NULL, part_count_thread, &tk[t]);
  }
  for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
  uint64_t run = 0;
  for (uint32_t p = 0; p < nparts; ++p) {
    start[p] = run;
    for (int t = 0; t < nt; ++t) {
      const uint64_t c = tk[t].hist[p];
      tk[t].hist[p] = run;
      run += c;
    }
  }
  start[nparts] = run;
  for (int t = 0; t < nt; ++t) pthread_create(&th[t], NULL, part_scatter_thread, &tk[t]);
  for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
  free(hist);
  return 0;
}

typedef struct {
  const otuple *hr, *hs;
  const uint64_t *rloc, *sloc;
  uint32_t nparts;
  uint32_t *next; /* shared partition counter */
  opair *buf;
  uint64_t n, cap;
} probe_task;

static void *probe_thread(void *arg) {
  probe_task *t = (probe_task *)arg;
  const uint32_t chunk = 256;
  for (;;) {
    const uint32_t p0 = __atomic_fetch_add(t->next, chunk, __ATOMIC_RELAXED);
    if (p0 >= t->nparts) break;
    const uint32_t p1 = p0 + chunk < t->nparts ? p0 + chunk : t->nparts;
    for (uint32_t p = p0; p < p1; ++p)
      for (uint64_t j = t->sloc[p]; j < t->sloc[p + 1]; ++j)
        for (uint64_t k = t->rloc[p]; k < t->rloc[p + 1]; ++k)
          if (t->hs[j].key == t->hr[k].key) {
            if (t->n == t->cap) {
              uint64_t nc = t->cap ? t->cap * 2 : 4096;
              opair *nb = (opair *)realloc(t->buf, nc * sizeof(opair));
              if (!nb) return (void *)1;
              t->buf = nb;
              t->cap = nc;
            }
            t->buf[t->n].r_payload = t->hr[k].payload;
            t->buf[t->n].s_payload = t->hs[j].payload;
            t->n++;
          }
  }
  return NULL;
}

uint64_t oracle_hash_join_mt(const otuple *R, uint64_t nR, const otuple *S, uint64_t nS, uint32_t nparts,
                             int nthreads, opair *out, uint64_t cap) {
  if (nparts == 0) nparts = 1u << 20;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  otuple *hr = (otuple *)malloc((nR ? nR : 1) * sizeof(otuple));
  otuple *hs = (otuple *)malloc((nS ? nS : 1) * sizeof(otuple));
  uint64_t *rloc = (uint64_t *)malloc(((uint64_t)nparts + 1) * sizeof(uint64_t));
  uint64_t *sloc = (uint64_t *)malloc(((uint64_t)nparts + 1) * sizeof(uint64_t));
  uint64_t total = UINT64_MAX;
  if (hr && hs && rloc && sloc && !partition_mt(R, nR, nparts, nthreads, hr, rloc) &&
      !partition_mt(S, nS, nparts, nthreads, hs, sloc)) {
    pthread_t th[256];
    probe_task tk[256];
    uint32_t next = 0;
    for (int t = 0; t < nthreads; ++t) {
      tk[t] = (probe_task){hr, hs, rloc, sloc, nparts, &next, NULL, 0, 0};
      pthread_create(&th[t], NULL, probe_thread, &tk[t]);
    }
    int failed = 0;
    for (int t = 0; t < nthreads; ++t) {
      void *rv = NULL;
      pthread_join(th[t], &rv);
      failed |= rv != NULL;
    }
    if (!failed) {
      total = 0;
      for (int t = 0; t < nthreads; ++t) {
        for (uint64_t i = 0; i < tk[t].n && total + i < cap; ++i) out[total + i] = tk[t].buf[i];
        total += tk[t].n;
      }
    }
    for (int t = 0; t < nthreads; ++t) free(tk[t].buf);
  }
  free(hr);
  free(hs);
  free(rloc);
  free(sloc);
  return total;
}

/* ---- N-core sorted-index join ----------------------------------------------
 * The pthread sorted-index join of joins/multi_joins/sortindex/
 * thread-hash-bnlj.cc:180-335: the index is built as in sortIndex_cpu (copy
 * {val, adr}, quicksort, order check), then workers take LT_BUF-sized chunks
 * of S from a shared cursor (LTGetter, :210) and binary-search + expand each
 * probe (executor, :254).  Fixes: a shared atomic counter instead of the mutex
 * around the cursor, per-thread result buffers instead of a mutex per result
 * row (:276, :292), and the expansions are bounds-checked as in
 * oracle_sidx_join. */

typedef struct {
  const otuple *R, *S, *b;
  uint64_t nR, nS;
  uint64_t *next; /* shared S cursor */
  opair *buf;
  uint64_t n, cap;
  int count_only;
} sidx_task;

static int sidx_push(sidx_task *t, int32_t rp, int32_t sp) {
  if (!t->count_only) {
    if (t->n == t->cap) {
      uint64_t nc = t->cap ? t->cap * 2 : 4096;
      opair *nb = (opair *)realloc(t->buf, nc * sizeof(opair));
      if (!nb) return -1;
      t->buf = nb;
      t->cap = nc;
    }
    t->buf[t->n].r_payload = rp;
    t->buf[t->n].s_payload = sp;
  }
  t->n++;
  return 0;
}

static void *sidx_probe_thread(void *arg) {
  sidx_task *t = (sidx_task *)arg;
  const uint64_t chunk = 4096; /* the reference's LT_BUF */
  for (;;) {
    const uint64_t j0 = __atomic_fetch_add(t->next, chunk, __ATOMIC_RELAXED);
    if (j0 >= t->nS) break;
    const uint64_t j1 = j0 + chunk < t->nS ? j0 + chunk : t->nS;
    for (uint64_t j = j0; j < j1; ++j) {
      const int32_t v = t->S[j].key;
      const uint64_t bidx = ref_search(t->b, v, t->nR);
      uint64_t x = bidx;
      while (t->b[x].key == v) {
        if (sidx_push(t, t->R[t->b[x].payload].payload, t->S[j].payload)) return (void *)1;
        if (x == 0) break;
        x--;
      }
      for (x = bidx + 1; x < t->nR && t->b[x].key == v; ++x)
        if (sidx_push(t, t->R[t->b[x].payload].payload, t->S[j].payload)) return (void *)1;
    }
  }
  return NULL;
}

uint64_t oracle_sidx_join_mt(const otuple *R, uint64_t nR, const otuple *S, uint64_t nS, int nthreads, opair *out,
                             uint64_t cap) {
  if (nR == 0) return 0;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  otuple *b = (otuple *)malloc(nR * sizeof(otuple));
  if (!b || oracle_sidx_build(R, nR, b)) {
    free(b);
    return UINT64_MAX;
  }
  pthread_t th[256];
  sidx_task tk[256];
  uint64_t next = 0;
  for (int t = 0; t < nthreads; ++t) {
    tk[t] = (sidx_task){R, S, b, nR, nS, &next, NULL, 0, 0, out == NULL};
    pthread_create(&th[t], NULL, sidx_probe_thread, &tk[t]);
  }
  int failed = 0;
  for (int t = 0; t < nthreads; ++t) {
    void *rv = NULL;
    pthread_join(th[t], &rv);
    failed |= rv != NULL;
  }
  uint64_t total = 0;
  for (int t = 0; t < nthreads; ++t) {
    if (out)
      for (uint64_t i = 0; i < tk[t].n && total + i < cap; ++i) out[total + i] = tk[t].buf[i];
    total += tk[t].n;
    free(tk[t].buf);
  }
  free(b);
  return failed ? UINT64_MAX : total;
}
