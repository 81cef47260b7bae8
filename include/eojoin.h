/*
 * eojoin.h — C ABI of the MI355X inner-join library (libeojoin.so).
 *
 * Data model (reference: joins/hashJoin_gpu/main_street/tuple.h:43-53,
 * joins/CPU_JOIN/hash_cpu/tuple.h:16-26):
 *   the reference TUPLE is {int key (row id), int val (join attribute)};
 *   here eo_tuple.key is the join attribute (reference .val) and
 *   eo_tuple.payload is the row id (reference .key).  An output eo_pair is
 *   (R payload, S payload) = (RESULT.rkey, RESULT.lkey) of the reference.
 *   R is the build side (reference "right"/rt), S the probe side ("left"/lt).
 *
 * Conventions (replacing the reference's printf+exit(1) error handling,
 * joins/hashJoin_gpu/main_street/hj.cc:214-227):
 *   - every entry point returns an eo_status; nothing calls exit();
 *   - inputs are caller-owned; opts->where says whether R/S/out are host or
 *     device pointers (EO_DEVICE: inputs already resident in HBM);
 *   - two-phase sizing, mirroring count→scan→alloc→write of hj.cc:1040-1069:
 *     out == NULL returns the match count in *n_out and writes nothing;
 *     out_cap < count returns EO_ECAPACITY with *n_out set; the contents of
 *     out are then unspecified (with EO_F_DETERMINISTIC nothing is written);
 *   - output order is unspecified: compare results as multisets;
 *   - one eo_ctx per host thread; calls on one ctx are serialised by the caller;
 *   - indexes, tables and batches belong to the ctx they were made on: free
 *     them before eo_ctx_destroy of that ctx;
 *   - scratch memory comes from a ctx-owned device arena that only grows.
 */
#ifndef EOJOIN_H
#define EOJOIN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EOJOIN_ABI_VERSION 1u

typedef struct eo_tuple {
  int32_t key;     /* join attribute (reference TUPLE.val) */
  int32_t payload; /* row id        (reference TUPLE.key) */
} eo_tuple;

typedef struct eo_pair {
  int32_t r_payload; /* reference RESULT.rkey */
  int32_t s_payload; /* reference RESULT.lkey */
} eo_pair;

typedef enum eo_status {
  EO_OK = 0,
  EO_EINVAL = 1,     /* bad argument */
  EO_ENOMEM = 2,     /* device or host allocation failed */
  EO_ECAPACITY = 3,  /* out_cap < match count; *n_out holds the count */
  EO_EHIP = 4,       /* HIP runtime error or device-side fault flag */
  EO_EBUDGET = 5,    /* request exceeds the HBM admission budget */
  EO_ENODEV = 6,     /* no usable GPU */
  EO_EAGAIN = 7,     /* not ready yet (non-blocking poll) */
  EO_ECANCELED = 8,  /* request cancelled */
  EO_ETIMEDOUT = 9,  /* request deadline expired */
  EO_EINTERNAL = 10  /* internal invariant violated */
} eo_status;

typedef enum eo_where { EO_HOST = 0, EO_DEVICE = 1 } eo_where;

typedef struct eo_ctx eo_ctx;
typedef struct eo_index eo_index;

typedef struct eo_ctx_opts {
  uint32_t struct_size;  /* sizeof(eo_ctx_opts) */
  uint32_t flags;        /* reserved, 0 */
  uint64_t arena_bytes;  /* initial scratch arena size (0 = grow on demand) */
  uint64_t hbm_budget;   /* admission budget for batch/serve paths (0 = 200 GB) */
} eo_ctx_opts;

#define EO_MAX_KERNEL_STATS 12

/* Device time of one kernel family during a call (HIP events around each
 * launch on the call's stream), summed over its launches. */
typedef struct eo_kernel_stat {
  char name[32];
  double ms;
  uint32_t launches;
  uint32_t reserved;
} eo_kernel_stat;

/* Per-phase device times of the last call (hipEvent based), in ms. */
typedef struct eo_join_stats {
  double ms_total;
  double ms_partition;   /* histogram + scan + scatter, both relations */
  double ms_count;       /* count pass (incl. per-tile scan) */
  double ms_write;       /* write pass */
  double ms_build;       /* index build (index joins) */
  uint64_t n_out;        /* match count */
  uint32_t radix_bits;   /* partition bits used */
  uint32_t passes;       /* partition passes used */
  uint32_t tiles;        /* join work tiles (partition x R-chunk x S-range) */
  uint32_t lds_tuples;   /* LDS table capacity used */
  uint32_t n_kernels;    /* valid entries of kernels[] */
  uint32_t sampled;      /* hash join partitions: 0 exact (histogram pass), 1 sized
                            from a sample, 2 sized from a sample that proved too
                            small, so the join was redone exactly */
  eo_kernel_stat kernels[EO_MAX_KERNEL_STATS];
} eo_join_stats;

typedef struct eo_join_opts {
  uint32_t struct_size;  /* sizeof(eo_join_opts) */
  int32_t where;         /* EO_HOST or EO_DEVICE for R, S and out */
  int32_t radix_bits;    /* total radix bits of the partition (0 = autotune) */
  int32_t passes;        /* partition passes, 1 or 2 (0 = autotune) */
  int32_t lds_tuples;    /* R tuples per LDS table: 4096, 8192 or 16384; 0 = the
                            tuned entry, else the planner's (16384, or 8192 for a
                            join whose |R| needs two partition passes) */
  int32_t s_tile;        /* S tuples per join tile (0 = autotune) */
  int32_t flags;         /* EO_F_* bits */
  int32_t reserved;
  void *stream;          /* hipStream_t to run on; NULL = the ctx stream */
  eo_join_stats *stats;  /* optional out: per-phase times of this call */
} eo_join_opts;

#define EO_F_NO_SYNC_STATS 1  /* do not record per-phase events */
/* Deterministic output order: separate count and write passes with a device
 * scan of per-tile counts between them (the reference's count -> presum ->
 * join sequence, hj.cc:1011-1091): the same rows in the same order on every
 * run.  Default: one fused pass per tile that reserves its output range with
 * an atomic (order varies run to run). */
#define EO_F_DETERMINISTIC 2
/* Output rows in the reference's 16-byte RESULT layout {rkey, rval, lkey,
 * lval} = (R payload, key, S payload, key) (joins/hashJoin_gpu/main_street/
 * tuple.h:48-53) instead of 8-byte eo_pairs: `out` then points to 16-byte
 * rows and out_cap counts rows.  eo_hash_join, eo_nlj_join, eo_index_join. */
#define EO_F_RESULT16 8
/* Left outer join: every probe-side (S, the reference's "left") tuple with
 * no match also yields one row whose R side is EO_NULL_PAYLOAD (RESULT16:
 * rkey = rval = EO_NULL_PAYLOAD).  The reference names such a join
 * (joins/outerJoin_gpu) but implements none.  eo_hash_join, eo_nlj_join and
 * eo_index_join; with EO_F_DETERMINISTIC the null rows follow the inner rows
 * in S order (the hash join's S order is its stable partition's). */
#define EO_F_LEFT_OUTER 16
/* Hash join: always partition from a full histogram pass.  By default a
 * fused inner join of two passes and at least 2^20 tuples per relation sizes
 * its partitions from a 1/16 sample instead (one read of the input less) and
 * redoes the join exactly if the sample underestimated a partition; the rows
 * are the same either way. */
#define EO_F_EXACT_PARTITIONS 32
#define EO_NULL_PAYLOAD ((int32_t)0x80000000)

typedef struct eo_device_info {
  char name[64];
  char arch[32];          /* e.g. "gfx950" */
  int32_t cus;            /* compute units */
  int32_t xcds;           /* accelerator complex dies (8 on MI355X) */
  int32_t wave_size;      /* 64 */
  int32_t clock_khz;
  uint64_t lds_per_wg;    /* max LDS bytes per workgroup */
  uint64_t hbm_bytes;     /* total device memory */
  uint64_t l2_bytes;
} eo_device_info;

/* ---- library / context --------------------------------------------------- */

uint32_t eo_abi_version(void);
/* Hash of the sources the library was built from (16 hex digits), so that a
   caller can tell a stale build from a current one.  No reference counterpart
   (the reference rebuilds its cubins with make). */
const char *eo_build_source_hash(void);
const char *eo_status_string(eo_status st);
/* Message of the last failure on this host thread (HIP error text etc.). */
const char *eo_last_error(void);

/* Replaces cuInit/cuDeviceGet/cuCtxCreate + cuModuleLoad of hj.cc:213-289. */
eo_status eo_ctx_create(eo_ctx **ctx, int device, const eo_ctx_opts *opts);
eo_status eo_ctx_destroy(eo_ctx *ctx);
eo_status eo_ctx_sync(eo_ctx *ctx);
eo_status eo_device_query(eo_ctx *ctx, eo_device_info *info);

/* ---- device memory (replaces cuMemAlloc / cuMemcpyHtoD / cuMemcpyDtoH of
 *      hj.cc:324-354 and hj.cc:1124) -------------------------------------- */

eo_status eo_dev_alloc(eo_ctx *ctx, uint64_t bytes, void **dptr);
eo_status eo_dev_free(eo_ctx *ctx, void *dptr);
/* Pinned, device-mapped host memory (zero copy).  Passed as `out` with
 * where = EO_DEVICE, the join kernels write their rows straight into host
 * memory over PCIe, as the reference's cuMemHostAlloc(DEVICEMAP) result
 * buffers (joins/hashJoin_gpu/main_street/hj.cc:152); inputs may live there
 * too (read over PCIe on every pass). */
eo_status eo_host_alloc(eo_ctx *ctx, uint64_t bytes, void **hptr);
eo_status eo_host_free(eo_ctx *ctx, void *hptr);
eo_status eo_copy_h2d(eo_ctx *ctx, void *dst, const void *src, uint64_t bytes);
eo_status eo_copy_d2h(eo_ctx *ctx, void *dst, const void *src, uint64_t bytes);
eo_status eo_copy_d2d(eo_ctx *ctx, void *dst, const void *src, uint64_t bytes);

/* Streaming-copy bandwidth probe: copies `bytes` (multiple of 16) between two
 * scratch buffers `iters` times with the library's copy kernel and returns the
 * mean device bandwidth in GB/s (read + write bytes / time).  The measured
 * HBM ceiling reported beside every roofline fraction. */
eo_status eo_bandwidth_probe(eo_ctx *ctx, uint64_t bytes, int iters, double *gbs);

/* Bandwidth probe with an explicit kernel shape, for measuring the HBM
 * ceiling the join kernels are compared against: a copy, a read-only or a
 * write-only pass over `bytes` (multiple of 16) with `unroll` 16-byte words
 * in flight per thread (1, 2, 4, 8), `threads` per workgroup (64..1024,
 * 0 = 256), `blocks_per_cu` workgroups per CU (0 = one thread per word, no
 * grid-stride loop) and nontemporal loads / stores; EO_PROBE_SCATTER is a
 * copy with run-scattered writes (a radix scatter's write pattern; bytes / 16
 * / run words must be a power of two).  *gbs = bytes moved (read
 * + write) per second, mean over `iters` launches after one untimed launch.
 * No reference counterpart (the reference reports no bandwidth). */
enum { EO_PROBE_COPY = 0, EO_PROBE_READ = 1, EO_PROBE_WRITE = 2, EO_PROBE_SCATTER = 3 };
enum { EO_PROBE_NT_LOAD = 1, EO_PROBE_NT_STORE = 2 };
typedef struct eo_probe_opts {
  uint32_t struct_size;  /* sizeof(eo_probe_opts) */
  int32_t mode;          /* EO_PROBE_COPY / READ / WRITE */
  uint64_t bytes;
  int32_t unroll;
  int32_t threads;
  int32_t blocks_per_cu;
  uint32_t flags;        /* EO_PROBE_NT_LOAD | EO_PROBE_NT_STORE */
  int32_t iters;
  int32_t run_bytes;     /* EO_PROBE_SCATTER: a copy whose writes go in runs of this
                            many bytes (power of two, 16..65536) to scattered places */
} eo_probe_opts;
eo_status eo_bandwidth_probe_ex(eo_ctx *ctx, const eo_probe_opts *opts, double *gbs);

/* Integer VALU issue-rate probe, the ceiling of the nested-loop join's
 * compare roofline: EO_VALU_COMPARE runs the NLJ's inner loop (keys in
 * registers against a key tile in LDS, each compare's wave mask
 * OR-accumulated) with no matches, EO_VALU_ADD independent 32-bit add
 * chains.  `blocks_per_cu` 256-thread workgroups per CU (1..32), `reps`
 * passes; *per_s = compares (or adds) per second over `iters` launches after
 * one untimed launch.  No reference counterpart. */
enum { EO_VALU_COMPARE = 0, EO_VALU_ADD = 1 };
eo_status eo_valu_probe(eo_ctx *ctx, int kind, int blocks_per_cu, int reps, int iters, double *per_s);

/* Seeded synthetic relations written straight into device memory (R, S are
 * device pointers), for request streams too large to build on the host.
 * Counterpart of createTuple (joins/hashJoin_gpu/main_street/hj.cc:62-160),
 * which fills host arrays with clock-seeded rand().
 *   eo_gen_build: key = a seeded permutation of [0, n), payload = row;
 *   eo_gen_probe: with probability `selectivity` a key uniform in [0, nR)
 *     (exactly one match in a build side from eo_gen_build(nR)), otherwise
 *     uniform in [nR, 2^31 - 1); payload = row. */
eo_status eo_gen_build(eo_ctx *ctx, eo_tuple *R, uint64_t n, uint64_t seed,
                       const eo_join_opts *opts);
eo_status eo_gen_probe(eo_ctx *ctx, eo_tuple *S, uint64_t n, uint64_t nR,
                       double selectivity, uint64_t seed, const eo_join_opts *opts);

/* ---- primitives ---------------------------------------------------------- */

/* Exclusive prefix sum of n uint32 values; *total (optional) = sum.
 * Replaces presum() (joins/hashJoin_gpu/main_street/scan_main.cpp:38-111,
 * API scan_common.h:100-103) and getValue() (scan_common.h:112-115), without
 * its allocate-and-free-the-input ownership transfer: out may equal in. */
eo_status eo_exclusive_scan_u32(eo_ctx *ctx, const uint32_t *in, uint32_t *out,
                                uint64_t n, uint64_t *total,
                                const eo_join_opts *opts);
eo_status eo_exclusive_scan_u64(eo_ctx *ctx, const uint64_t *in, uint64_t *out,
                                uint64_t n, uint64_t *total,
                                const eo_join_opts *opts);

/* Stable radix partition of n tuples by digit ((uint32)key >> shift) & (2^bits-1),
 * 1 <= bits <= 16.  part_off (2^bits + 1 entries, optional) receives the
 * partition start offsets.  Replaces lcount_partitioning + presum +
 * lpartitioning + countPartition (joins/hashJoin_gpu/main_street/
 * partitioning.cu:15-131, :273-293; hj.cc:430-656); its output equals the
 * stable partition of joins/CPU_JOIN/hash_cpu/hj.cc:132-181 for key >= 0. */
eo_status eo_radix_partition(eo_ctx *ctx, const eo_tuple *in, uint64_t n,
                             int bits, int shift, eo_tuple *out,
                             uint64_t *part_off, const eo_join_opts *opts);

/* ---- joins --------------------------------------------------------------- */

/* Radix-partitioned hash join: replaces join() of
 * joins/hashJoin_gpu/main_street/hj.cc:171-1269 (CPU oracle
 * joins/CPU_JOIN/hash_cpu/hj.cc:116-228). */
eo_status eo_hash_join(eo_ctx *ctx, const eo_tuple *R, uint64_t nR,
                       const eo_tuple *S, uint64_t nS, eo_pair *out,
                       uint64_t out_cap, uint64_t *n_out,
                       const eo_join_opts *opts);

/* Hash join that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops that must not call hipMalloc).  Same results as
 * eo_hash_join with the same arguments; R, S and out are device pointers
 * (opts->where = EO_DEVICE, required) and EO_F_DETERMINISTIC is refused.
 * The join's scratch comes from one fixed device buffer of
 * EO_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends (about
 * 8 (|R| + |S|) bytes plus small tables: a one-pass join of |R| + |S| up to
 * about 7.5 million tuples); a join that needs more returns EO_EINVAL with
 * the sizes in eo_last_error().  All calls on one GPU share the buffer:
 * calls that overlap on one GPU (several host threads or contexts) race. */
#define EO_NOALLOC_SCRATCH_BYTES (64ull << 20)
eo_status eo_hash_join_noalloc(eo_ctx *ctx, const eo_tuple *R, uint64_t nR,
                               const eo_tuple *S, uint64_t nS, eo_pair *out,
                               uint64_t out_cap, uint64_t *n_out,
                               const eo_join_opts *opts);

/* Block nested-loop join: replaces join() of
 * joins/blockJoin_gpu/main_street/nestLoopJoin.cc:180-655 (CPU oracle
 * joins/CPU_JOIN/block_cpu/nestLoopJoin.cc:61-77). */
eo_status eo_nlj_join(eo_ctx *ctx, const eo_tuple *R, uint64_t nR,
                      const eo_tuple *S, uint64_t nS, eo_pair *out,
                      uint64_t out_cap, uint64_t *n_out,
                      const eo_join_opts *opts);

/* A 3-D point of the reference's distance join (joins/CPU_JOIN/3d_cpu/
 * nestLoopJoin.cc, TUPLE {int id; double val[VAL_NUM = 3]}, 32 bytes). */
typedef struct eo_point3 {
  int32_t id;
  int32_t pad;
  double v[3];
} eo_point3;

/* Nested-loop join with the reference's multi-attribute distance predicate
 * (nestLoopJoin.cc:73-81): a pair matches when
 * (r.v0-s.v0)^2 + (r.v1-s.v1)^2 + (r.v2-s.v2)^2 < distance^2, evaluated in FP64
 * in that order without fused multiply-add (decisions identical to the
 * reference's).  Rows are eo_pairs (R id, S id).  Same sizing and error
 * conventions as eo_hash_join; the output order is unspecified. */
eo_status eo_nlj_dist_join(eo_ctx *ctx, const eo_point3 *R, uint64_t nR, const eo_point3 *S,
                           uint64_t nS, double distance, eo_pair *out, uint64_t out_cap,
                           uint64_t *n_out, const eo_join_opts *opts);

/* ---- out-of-core streaming join ------------------------------------------- */

/* Host-resident R and S of any size, also beyond device memory: R is taken
 * r_tile tuples at a time (uploaded, and partitioned once per tile for the
 * hash join); S streams through the device s_chunk tuples at a time, chunk
 * i + 1 crossing PCIe while chunk i is joined and the rows of chunk i - 1
 * travel back.  Replaces the host tile loop of
 * joins/blockJoin_gpu/main_street/nestLoopJoin.cc:323-519 (HtoD of a tile
 * pair, count, presum, join, DtoH, append; every step synchronous). */
typedef enum eo_stream_op { EO_STREAM_HASH = 0, EO_STREAM_NLJ = 1 } eo_stream_op;

typedef struct eo_stream_stats {
  double ms_total;      /* wall time of the call */
  uint64_t h2d_bytes;   /* bytes copied host -> device (R tiles and S chunks) */
  uint64_t d2h_bytes;   /* bytes of rows copied device -> host */
  uint64_t n_out;       /* match count */
  uint32_t r_tiles;     /* R tiles */
  uint32_t s_chunks;    /* S chunks joined, over all R tiles */
  uint32_t regrown;     /* chunks joined again after their rows outgrew the staging */
  uint32_t pinned;      /* bit 0: R, bit 1: S, bit 2: out were eo_host_alloc memory */
} eo_stream_stats;

typedef struct eo_stream_opts {
  uint32_t struct_size;    /* sizeof(eo_stream_opts) */
  int32_t op;              /* eo_stream_op */
  uint64_t r_tile;         /* R tuples resident at once (0: all of R; at most 2^31 - 1) */
  uint64_t s_chunk;        /* S tuples per chunk (0: 2^24; at most 2^31 - 1) */
  uint64_t out_chunk;      /* device rows staged per chunk (0: s_chunk); grows when outgrown */
  int32_t flags;           /* EO_F_RESULT16; EO_F_LEFT_OUTER when R fits one tile */
  int32_t reserved;
  const eo_join_opts *join; /* optional plan overrides of the hash join per R tile (where ignored) */
  eo_stream_stats *stats;  /* optional out */
} eo_stream_opts;

/* R, S and out are host memory; blocks from eo_host_alloc are read and
 * written by DMA directly, other memory through pinned staging buffers that
 * host threads fill and drain.  nR and nS may exceed 2^31 (payloads stay the
 * caller's int32 values).  out == NULL counts only; out_cap < count returns
 * EO_ECAPACITY with *n_out = count and out holding the first out_cap rows
 * produced.  Rows are R-tile by R-tile, chunk by chunk; the order inside a
 * chunk is unspecified. */
eo_status eo_stream_join(eo_ctx *ctx, const eo_tuple *R, uint64_t nR, const eo_tuple *S, uint64_t nS,
                         eo_pair *out, uint64_t out_cap, uint64_t *n_out, const eo_stream_opts *opts);

/* ---- resident indexes ------------------------------------------------------ */

typedef enum eo_index_kind { EO_INDEX_HASH = 0, EO_INDEX_SORTED = 1 } eo_index_kind;

/* Lookup methods of eo_index_lookup. */
enum { EO_LOOKUP_DEFAULT = 0, EO_LOOKUP_BINARY = 1, EO_LOOKUP_LEARNED = 2 };

/* eo_index_join on a learned sorted index: use plain binary search instead. */
#define EO_F_BINARY_SEARCH 4

typedef struct eo_index_info {
  int32_t kind;      /* eo_index_kind */
  int32_t learned;   /* sorted index carries the learned model */
  uint64_t n;        /* indexed R tuples */
  uint64_t bytes;    /* device bytes held by the index */
  int32_t leaves;    /* learned model: stage-2 linear models */
  int32_t max_err;   /* learned model: largest last-mile window half-width */
  double mean_err;   /* learned model: mean of err_lo + err_hi over leaves */
  double build_ms;   /* device time of the build */
} eo_index_info;

/* Hash index on R, built on the device and kept resident until
 * eo_index_free.  Replaces createIndex of joins/innerJoin_gpu/main_street/
 * inner-idx-bnlj.cc:149-180 (a CPU counting sort into NB_BUCKET buckets) and
 * its per-join HtoD copies (:279-330). */
eo_status eo_hidx_build(eo_ctx *ctx, const eo_tuple *R, uint64_t nR,
                        const eo_join_opts *opts, eo_index **index);

/* Sorted index on R (device radix sort), optionally with the learned model
 * (learned != 0).  Replaces qsort + createIndex of
 * joins/innerJoin_gpu/sort_index/main_street/inner-idx-bnlj.cc:175-216. */
eo_status eo_sidx_build(eo_ctx *ctx, const eo_tuple *R, uint64_t nR, int learned,
                        const eo_join_opts *opts, eo_index **index);

/* Index join of S against a resident index: replaces join() of
 * joins/innerJoin_gpu/main_street/inner-idx-bnlj.cc:183-673 (hash) and
 * joins/innerJoin_gpu/sort_index/main_street/inner-idx-bnlj.cc:219-658
 * (sorted: binary search, or the learned model unless EO_F_BINARY_SEARCH).
 * Same sizing and error conventions as eo_hash_join. */
eo_status eo_index_join(eo_ctx *ctx, const eo_index *index, const eo_tuple *S,
                        uint64_t nS, eo_pair *out, uint64_t out_cap,
                        uint64_t *n_out, const eo_join_opts *opts);

/* Per probe, the index position its search starts from: the lower bound of
 * S[i].key (sorted index; method BINARY or LEARNED — equal for every key
 * present in R) or the bucket start (hash index).  pos: nS uint64, host or
 * device per opts->where.  Used to check the learned model against binary
 * search probe by probe. */
eo_status eo_index_lookup(eo_ctx *ctx, const eo_index *index, const eo_tuple *S,
                          uint64_t nS, int method, uint64_t *pos,
                          const eo_join_opts *opts);

/* Sorted (optionally learned) index from pairs already sorted by key, e.g.
 * the reference's BUCKET index file read by eo_ref_index_read: no sort, the
 * order is verified on the device (EO_EINVAL if the input is not sorted by
 * signed key).  Same index as eo_sidx_build of the same pairs. */
eo_status eo_sidx_import(eo_ctx *ctx, const eo_tuple *sorted, uint64_t n, int learned,
                         const eo_join_opts *opts, eo_index **index);

/* Copies the index arrays (keys, payloads; index order) to host memory. */
eo_status eo_index_export(eo_ctx *ctx, const eo_index *index, uint32_t *keys,
                          uint32_t *pays);
eo_status eo_index_query(const eo_index *index, eo_index_info *info);
eo_status eo_index_free(eo_index *index);

/* ---- the reference's on-disk table and index files (host only) ----------- */

/* Table file of joins/hashJoin_gpu/main_street/TableBuild.cpp:113-141: an
 * int32 row count, then that many TUPLEs {int32 key (row id), int32 val}.
 * EO_REF_FILE_RAW: the same tuples without the count (joins/join_ori/R, S).
 * Read: with out == NULL only *n is returned; cap < n gives EO_ECAPACITY.
 * Tuples come back as eo_tuple {key = val, payload = row id}; write converts
 * back.  Neither needs a device. */
#define EO_REF_FILE_RAW 1
eo_status eo_ref_table_read(const char *path, int flags, eo_tuple *out, uint64_t cap,
                            uint64_t *n);
eo_status eo_ref_table_write(const char *path, int flags, const eo_tuple *tuples, uint64_t n);
/* Sorted-index file of joins/innerJoin_gpu/sort_index/main_street/
 * TableBuild.cpp:180-218: BUCKETs {int32 val, int32 adr} sorted by val, no
 * header (n = file size / 8); read as eo_tuple {key = val, payload = adr}. */
eo_status eo_ref_index_read(const char *path, eo_tuple *out, uint64_t cap, uint64_t *n);
eo_status eo_ref_index_write(const char *path, const eo_tuple *sorted, uint64_t n);

/* ---- resident build tables and batched joins ------------------------------ */

typedef struct eo_table eo_table;
typedef struct eo_batch eo_batch;

/* A build relation partitioned once on the device (radix bits chosen from
 * |R| as eo_hash_join would, or opts->radix_bits) and kept resident until
 * eo_table_free, so that later joins against it partition only S.  The
 * reference rebuilds and re-partitions R for every join
 * (joins/hashJoin_gpu/main_street/hj.cc:498-602). */
eo_status eo_table_create(eo_ctx *ctx, const eo_tuple *R, uint64_t nR,
                          const eo_join_opts *opts, eo_table **table);
eo_status eo_table_query(const eo_table *table, uint64_t *n, int32_t *radix_bits,
                         uint64_t *bytes);
eo_status eo_table_free(eo_table *table);

/* One join of a batch.  All pointers are device pointers. */
typedef struct eo_request {
  const eo_tuple *R;      /* build side; ignored when table != NULL */
  uint64_t nR;
  const eo_table *table;  /* resident build side, or NULL */
  const eo_tuple *S;      /* probe side */
  uint64_t nS;
  eo_pair *out;           /* NULL: count only */
  uint64_t out_cap;
  uint64_t tag;           /* caller's id, echoed in the result */
} eo_request;

typedef struct eo_result {
  uint64_t tag;
  uint64_t n_out;   /* match count (all of it, also when above out_cap) */
  int32_t status;   /* EO_OK, or EO_ECAPACITY when out != NULL and n_out > out_cap
                       (out then holds out_cap of the pairs, which ones unspecified) */
  uint32_t reserved;
} eo_result;

/* A batch runs many joins of any sizes as one host-sync-free sequence on one
 * stream: the relations that need the same radix bits are partitioned
 * together (one histogram, scan and scatter launch per bit count), then one
 * persistent kernel drains a device queue of (join, partition, R chunk,
 * S span) probe tiles.  It replaces the reference's one-join-per-process
 * driver with a host synchronisation after every launch (hj.cc:454-1105).
 * Results equal those of eo_hash_join on each request alone.  A batch object
 * owns its scratch arena and stream; one batch is in flight per object. */
eo_status eo_batch_create(eo_ctx *ctx, eo_batch **batch);
eo_status eo_batch_free(eo_batch *batch);
/* Device scratch bytes a batch of these requests needs (outputs excluded). */
eo_status eo_batch_footprint(eo_ctx *ctx, const eo_request *reqs, uint32_t n,
                             uint64_t *bytes);
/* Enqueues the batch and returns without waiting.  opts (optional): stream,
 * stats (filled by eo_batch_wait), radix_bits / passes / lds_tuples / s_tile
 * overrides for requests without a table, and flags for every request:
 * EO_F_RESULT16 (16-byte rows), EO_F_LEFT_OUTER (a null row per unmatched S
 * tuple; |S|/8 more scratch bytes per request, not in eo_batch_footprint). */
eo_status eo_batch_submit(eo_batch *batch, const eo_request *reqs, uint32_t n,
                          const eo_join_opts *opts);
/* EO_OK when the batch in flight has finished, EO_EAGAIN while it runs. */
eo_status eo_batch_poll(eo_batch *batch);
/* Waits for the batch in flight and writes one result per request (n must
 * equal the submitted count). */
eo_status eo_batch_wait(eo_batch *batch, eo_result *results, uint32_t n);
/* Runs the last submitted batch again over the same buffers (whose contents
 * may have changed) from a hipGraph captured on first use; complete it with
 * eo_batch_wait.  A new eo_batch_submit discards the graph. */
eo_status eo_batch_replay(eo_batch *batch);

/* ---- autotuner ---------------------------------------------------------------- */

typedef struct eo_tune_result {
  int32_t radix_bits;   /* chosen plan (0 bits = no partitioning) */
  int32_t passes;
  int32_t lds_tuples;   /* 16384 or 8192 (0 in a table saved before the LDS axis: the planner's) */
  int32_t s_tile;       /* 0: the planner's S span */
  double ms;            /* best device time of the chosen plan */
  double default_ms;    /* device time of the untuned planner's plan */
  uint32_t candidates;  /* plans timed */
  uint32_t reserved;
} eo_tune_result;

/* Times candidate plans (the LDS table size — a full table with one
 * workgroup per CU, or a half table with two — radix bits around the
 * planner's choice for it, one or two passes, three S spans) for the size
 * class of (|R|, |S|) = (ceil log2 nR,
 * ceil log2 nS) on the given device inputs, best of `reps` runs each, and
 * stores the fastest in the ctx's table.  eo_hash_join and eo_batch_submit
 * use the table for requests whose opts leave radix_bits, passes,
 * lds_tuples and s_tile at 0.  Replaces the reference's compile-time plan
 * constants (joins/hashJoin_gpu/main_street/hj.cc:378-436, tuple.h). */
eo_status eo_tune_hash_join(eo_ctx *ctx, const eo_tuple *R, uint64_t nR,
                            const eo_tuple *S, uint64_t nS, eo_pair *out,
                            uint64_t out_cap, int reps, eo_tune_result *result);
/* The table entry for a size class: EO_OK, or EO_EAGAIN when untuned. */
eo_status eo_tune_lookup(eo_ctx *ctx, uint64_t nR, uint64_t nS, eo_tune_result *result);
eo_status eo_tune_clear(eo_ctx *ctx);
/* Text file, one size class per line (validated on load). */
eo_status eo_tune_save(eo_ctx *ctx, const char *path);
eo_status eo_tune_load(eo_ctx *ctx, const char *path);

/* ---- request loop ------------------------------------------------------------ */

enum { EO_SERVE_BATCHED = 0, EO_SERVE_PER_JOIN = 1 };

typedef struct eo_serve_opts {
  uint32_t struct_size;  /* sizeof(eo_serve_opts); a struct that ends at hbm_budget
                            (no deadline_ms / cancel fields) is accepted as well */
  int32_t mode;          /* EO_SERVE_BATCHED, or EO_SERVE_PER_JOIN: one eo_hash_join per request */
  uint32_t max_batch;    /* requests per batch (0 = 512) */
  uint32_t slots;        /* batches in flight at once (0 = 2, at most 8) */
  uint64_t hbm_budget;   /* bytes admitted at once (0 = the ctx budget, 200 GB by default) */
  /* Per request, optional (NULL = none): a request not admitted within
     deadline_ms[i] (> 0) of its arrival completes with EO_ETIMEDOUT without
     running; a request whose cancel[i] is non-zero before it is admitted
     completes with EO_ECANCELED without running (the flags may be set from
     another thread while the loop runs).  Admitted requests run to the end. */
  const double *deadline_ms;
  const volatile uint32_t *cancel;
} eo_serve_opts;

typedef struct eo_serve_stats {
  double wall_ms;          /* loop start to last completion */
  double p50_ms, p99_ms;   /* latency = completion - arrival, nearest rank */
  double max_ms, mean_ms;
  uint64_t tuples;         /* sum of |R| + |S| over the requests */
  uint64_t n_out;          /* sum of match counts */
  uint32_t batches;
  uint32_t largest_batch;
  uint64_t peak_admitted;  /* largest admitted bytes in flight */
  uint32_t refused;        /* requests that completed with EO_EBUDGET */
  uint32_t expired;        /* ... EO_ETIMEDOUT */
  uint32_t cancelled;      /* ... EO_ECANCELED */
  uint32_t reserved;
} eo_serve_stats;

/* Serves a stream of requests: request i arrives arrival_ms[i] after the call
 * starts (non-decreasing; NULL = all at once).  Admission by HBM budget (a
 * request's inputs, partition scratch and output buffer count while it is in
 * flight), continuous batching into eo_batch objects (one per slot, each on
 * its own stream).  results[i] and latency_ms[i] (optional) per request; a
 * request larger than the budget gets EO_EBUDGET, an expired or cancelled one
 * EO_ETIMEDOUT or EO_ECANCELED (n_out 0, its output untouched; latency = the
 * time it left the queue).  The latency percentiles cover the requests that
 * ran.  No reference counterpart:
 * the reference runs one join per process (hj.cc:1189-1269). */
eo_status eo_serve_run(eo_ctx *ctx, const eo_request *reqs, const double *arrival_ms,
                       uint32_t n, const eo_serve_opts *opts, eo_result *results,
                       double *latency_ms, eo_serve_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* EOJOIN_H */
