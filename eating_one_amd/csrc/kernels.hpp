// Host-side launchers of the library's HIP kernels.  Every launcher enqueues
// on `stream` and returns the hipError_t of the launch; nothing here
// synchronises or allocates.  Scratch buffers are carved by the caller from
// the context arena; their sizes come from the *_scratch_bytes() helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "eojoin.h"

namespace eo {

// ---- per-launch timing --------------------------------------------------------
// When a timer is installed on the calling host thread, timed launch sites
// record a HIP event pair around the kernel on its own stream; the caller
// reads the elapsed times after synchronising.  No timer: no events.
constexpr int kMaxTimedLaunches = 64;
struct LaunchTimer {
  hipEvent_t ev[2 * kMaxTimedLaunches] = {};
  const char *name[kMaxTimedLaunches] = {};
  int n = 0;
  bool overflow = false;
};
LaunchTimer *set_launch_timer(LaunchTimer *t);  // returns the previous one
void timer_begin(const char *name, hipStream_t s);
void timer_end(hipStream_t s);
#define EO_TIMED(name, stream, ...)  \
  do {                               \
    ::eo::timer_begin(name, stream); \
    __VA_ARGS__;                     \
    ::eo::timer_end(stream);         \
  } while (0)

// ---- scan.hip ---------------------------------------------------------------
// Exclusive scan with decoupled look-back.  scratch: scan_scratch_bytes(n).
// total_dev (optional, device) receives the sum.  out may equal in.
size_t scan_scratch_bytes_u32(uint64_t n);
size_t scan_scratch_bytes_u64(uint64_t n);
hipError_t scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint64_t *total_dev,
                    void *scratch, uint32_t *err, hipStream_t stream);
hipError_t scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *total_dev,
                    void *scratch, uint32_t *err, hipStream_t stream);
hipError_t fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t stream);
hipError_t fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t stream);

// ---- util.hip ---------------------------------------------------------------
// Streaming 16-byte copy of `bytes` (a multiple of 16), one word per thread.
hipError_t copy_bytes(const void *src, void *dst, uint64_t bytes, int cus, hipStream_t stream);
// One launch of the bandwidth probe kernel that `o` describes (copy, read or
// write of o.bytes); sink: a device word array of o.threads entries.
hipError_t bandwidth_kernel(const eo_probe_opts &o, const void *src, void *dst, uint32_t *sink, int cus,
                            hipStream_t s);

// ---- partition.hip ----------------------------------------------------------
// One stable radix-partition pass over `bits` (1..kMaxPassBits) digit bits.
constexpr int kMaxPassBits = 11;
struct PartPassPlan {
  uint64_t n;
  int bits;
  int shift;
  uint32_t xor_mask;   // applied to (uint32)key before the shift (sign flip for sorting)
  uint32_t stripes;    // G: scatter workgroups, each owning one contiguous stripe
  uint64_t stripe_len; // tuples per stripe (even)
  uint32_t hist_split; // K: histogram workgroups per stripe
  bool fast;           // unstable LDS-atomic-ranked scatter (needs 16-byte aligned input)
  bool ordered = false;  // fast, <= 7 bits: ranks fixed by the input (the same positions on every run)
};
// fast = false: stable scatter (ballot-ranked), equal to the CPU_JOIN partition.
PartPassPlan plan_partition_pass(uint64_t n, int bits, int shift, uint32_t xor_mask, int cus, bool fast = false);
size_t partition_pass_scratch_bytes(const PartPassPlan &p);
// counts -> scan -> scatter.  part_start (optional, 2^bits+1 u32 entries, device)
// receives the partition start offsets of this pass.
hipError_t partition_pass(const PartPassPlan &p, const eo_tuple *in, eo_tuple *out,
                          uint32_t *part_start, void *scratch, uint32_t *err,
                          hipStream_t stream);
// Second pass of an MSD partition: the input is split into nseg segments
// (seg_start: nseg+1 u32, device); each segment is partitioned by digit
// ((uint32)key >> shift) & (2^bits - 1) with the fast scatter.  part_start
// (nseg * 2^bits + 1 u32) receives the starts of the nseg * 2^bits partitions,
// numbered seg * 2^bits + digit.  Input must be 16-byte aligned.
struct SegPassPlan {
  uint64_t n;
  uint32_t nseg;
  int bits, shift;
  uint64_t stripe_len;
  uint32_t max_stripes;
};
SegPassPlan plan_seg_pass(uint64_t n, uint32_t nseg, int bits, int shift, int cus);
size_t seg_pass_scratch_bytes(const SegPassPlan &p);
hipError_t partition_seg_pass(const SegPassPlan &p, const eo_tuple *in, eo_tuple *out, const uint32_t *seg_start,
                              uint32_t *part_start, void *scratch, uint32_t *err, hipStream_t stream);

// Partition start offsets of a tuple array already sorted by digit
// ((uint32)key ^ xor_mask) >> shift & (2^bits - 1):  part_start[d] = first index
// with digit >= d, part_start[2^bits] = n.
hipError_t partition_bounds(const eo_tuple *sorted, uint64_t n, int bits, int shift,
                            uint32_t xor_mask, uint32_t *part_start, hipStream_t stream);

// ---- hashjoin.hip -----------------------------------------------------------
// Scratch of join_fused_plan's scan over the P + 1 tile counts.
size_t join_plan_scratch_bytes(uint32_t parts);
// Fused single-pass probe.  Tiles are (partition, R-chunk of join_fused_cap()
// tuples, S span of s_span tuples); join_fused_plan counts and scans them
// (tile_start: P+1 u32, *n_tiles_dev = total).  Partition p of R is
// [r_start[p], r_end[p]) (contiguous partitions: r_end = r_start + 1), of S
// [s_start[p], s_end[p]).  *out_cursor must be zero on
// entry and holds the match count afterwards; pairs beyond out_cap are
// counted, not written.
// The probe's LDS table capacity for a planned lds_cap (8192 or less: the
// half-size table, two workgroups per CU; otherwise the full-size table).
uint32_t join_fused_cap(uint32_t lds_cap);
hipError_t join_fused_plan(uint32_t parts, uint32_t s_span, uint32_t cap, const uint32_t *r_start,
                           const uint32_t *r_end, const uint32_t *s_start, const uint32_t *s_end,
                           uint32_t *tile_start, uint64_t *n_tiles_dev, void *scan_scratch, uint32_t *err,
                           hipStream_t stream);
// Grid of join_fused for `parts` partitions of nR x nS tuples, sized without
// the tile count (tiles_est: the host's estimate of it).
uint32_t join_fused_grid(uint32_t parts, uint64_t nR, uint64_t nS, uint64_t tiles_est, int cus);
// The probe runs the *n_tiles_dev planned tiles over `grid` blocks (any grid
// is correct; a tile count of 2^31 or more raises kErrOverflow in *err and
// probes nothing).
// wide: out holds 16-byte RESULT rows {R payload, key, S payload, key};
// matched (optional): bitmap over S positions of the tuples that matched.
// swapped: R is the caller's S and S its R (the smaller relation built into
// the tables); the rows are written with the payloads exchanged.
hipError_t join_fused(uint32_t parts, uint32_t s_span, uint32_t cap, uint32_t grid, const uint64_t *n_tiles_dev,
                      const eo_tuple *R, const uint32_t *r_start, const uint32_t *r_end, const eo_tuple *S,
                      const uint32_t *s_start, const uint32_t *s_end, const uint32_t *tile_start, eo_pair *out,
                      uint64_t out_cap, uint64_t *out_cursor, bool wide, uint32_t *matched, uint32_t *err,
                      hipStream_t stream, bool swapped = false);
// One launch of the integer VALU probe `kind` (EO_VALU_COMPARE / EO_VALU_ADD)
// over `blocks` 256-thread blocks; *ops = the operations it performs.
hipError_t valu_kernel(int kind, uint32_t blocks, uint32_t reps, uint64_t *sink, hipStream_t s, double *ops);
// Copies the device error flags into *dst (one u64 word) and clears them, so
// that one readback returns results and flags together.
hipError_t take_error_flags(uint32_t *err, uint64_t *dst, hipStream_t stream);
// Deterministic order on the fused tiles (join_fused_plan's tile_start):
// span_count writes every tile's match count (and sets matched bits);
// span_write writes every tile's rows from tile_off[t] on, in an order fixed
// by the inputs.
hipError_t span_count(uint32_t parts, uint32_t s_span, uint32_t cap, uint32_t n_tiles, const eo_tuple *R,
                      const uint32_t *r_start, const eo_tuple *S, const uint32_t *s_start, const uint32_t *tile_start,
                      uint32_t *matched, uint64_t *tile_count, hipStream_t stream);
hipError_t span_write(uint32_t parts, uint32_t s_span, uint32_t cap, uint32_t n_tiles, const eo_tuple *R,
                      const uint32_t *r_start, const eo_tuple *S, const uint32_t *s_start, const uint32_t *tile_start,
                      const uint64_t *tile_off, eo_pair *out, uint64_t out_cap, bool wide, hipStream_t stream);
// Left outer join: bitmap words over n S positions, and the pass that emits
// one null-padded row per S tuple whose bit is clear (rows reserved from
// *cursor; out == NULL counts only).
size_t matched_words(uint64_t n);
// Deterministic variant: count_unmatched writes unmatched_blocks(n) block
// counts; after their exclusive scan, emit_unmatched_at writes each block's
// rows at base + block_off[block], in S order.
uint64_t unmatched_blocks(uint64_t n);
hipError_t count_unmatched(uint64_t n, const uint32_t *matched, uint64_t *block_cnt, hipStream_t stream);
hipError_t emit_unmatched_at(const eo_tuple *S, uint64_t n, const uint32_t *matched, const uint64_t *block_off,
                             uint64_t base, eo_pair *out, uint64_t out_cap, bool wide, hipStream_t stream);
hipError_t emit_unmatched(const eo_tuple *S, uint64_t n, const uint32_t *matched, uint64_t *cursor, eo_pair *out,
                          uint64_t out_cap, bool wide, hipStream_t stream);

// ---- index.hip --------------------------------------------------------------
// Resident indexes.  kind 0: hash (keys/pays grouped by bucket, off[NB+1]);
// kind 1: sorted (binary search); kind 2: sorted with the learned model.
uint64_t index_probe_tiles(uint64_t nS);
int learned_leaves();
int learned_knots();
size_t learned_model_bytes();
hipError_t hidx_build(const eo_tuple *R, uint64_t n, int log2nb, uint32_t *keys, uint32_t *pays, uint32_t *off,
                      uint32_t *cursor, void *scan_scratch, uint32_t *err, hipStream_t stream);
hipError_t sidx_split(const eo_tuple *sorted, uint64_t n, uint32_t *keys, uint32_t *pays, hipStream_t stream);
// *flag (device) = 1 if in[] is not sorted by signed key, else 0.
hipError_t check_sorted(const eo_tuple *in, uint64_t n, uint32_t *flag, hipStream_t stream);
// Fits the learned model over sorted keys; leaf_scratch: n u16.
hipError_t learned_fit(const uint32_t *keys, uint64_t n, void *model, uint16_t *leaf_scratch, hipStream_t stream);
// Count (write = false: per-tile counts into tile_cnt, or, with tile_cnt ==
// NULL, the total added into *cursor) or write (write = true: tile_cnt holds
// exclusive-scanned tile offsets, or, with tile_cnt == NULL, every tile
// reserves its range from *cursor and positions >= out_cap are counted only).
// lb_out (optional, nS u64): the index position each probe starts from
// (lower bound for sorted kinds).  outer: a probe without a match counts
// (and writes) one null-padded row (left outer join).
hipError_t index_probe(int kind, bool write, uint64_t n, int log2nb, const uint32_t *keys, const uint32_t *pays,
                       const uint32_t *off, const void *model, const eo_tuple *S, uint64_t nS,
                       uint64_t *tile_cnt, uint64_t *cursor, eo_pair *out, uint64_t out_cap, uint64_t *lb_out,
                       bool wide, bool outer, hipStream_t stream);

// Two-pass MSD partition (high digit bits_hi over the whole array, then the
// low digit bits_lo inside each first-pass partition) from a single
// histogram pass; needs 16-byte aligned in and tmp and bits_hi + bits_lo <=
// kMsd2MaxBits (the combined-digit histogram lives in LDS).
constexpr int kMsd2MaxBits = 14;
bool msd2_supported(int bits_hi, int bits_lo);
size_t msd2_scratch_bytes(uint64_t n, int bits_hi, int bits_lo, int cus);
// ordered: every tuple lands at a position fixed by the input (the same on
// every run; not the input's order within a partition).
hipError_t partition_msd2(uint64_t n, int bits_hi, int bits_lo, const eo_tuple *in, eo_tuple *tmp, eo_tuple *out,
                          uint32_t *part_start, void *scratch, uint32_t *err, int cus, hipStream_t stream,
                          bool ordered = false);

// Two-pass MSD partition into regions sized from a 1/16 sample of the input
// instead of a histogram pass: every tile reserves its runs from per-region
// cursors.  The final partitions are [part_start[p], part_end[p]) with gaps
// (part_start: 2^(bits_hi+bits_lo) + 1 entries); tmp holds gap_tmp_rows(n)
// tuples, out gap_out_rows(n).  A region the sample underestimated raises
// kErrCapacity in *err (its tuples are then incomplete: redo the join with
// partition_msd2).  Needs n >= kGapMinRows and 16-byte aligned in, tmp, out.
constexpr uint64_t kGapMinRows = 1ull << 20;
bool gap_supported(uint64_t n, int bits_hi, int bits_lo);
// Tuples of a buffer that holds `parts` sampled regions of a relation of n
// (for n >= kGapMinRows: enough for any sample).
uint64_t gap_rows(uint64_t n, uint64_t parts);
uint64_t gap_tmp_rows(uint64_t n, int bits_hi, int bits_lo);
uint64_t gap_out_rows(uint64_t n, int bits_hi, int bits_lo);
size_t gap_scratch_bytes(uint64_t n, int bits_hi, int bits_lo);
// phases: kGapPrepare (sample, reduce, region layout: writes part_start and
// the scratch) then kGapScatter (both passes and the partition ends, from the
// same scratch), in two calls or one.  Preparing both relations of a join
// before scattering either lets a layout that finds its relation clumped
// (kErrCapacity) stop both relations' scatters.
constexpr unsigned kGapPrepare = 1u, kGapScatter = 2u;
hipError_t partition_gapped(uint64_t n, int bits_hi, int bits_lo, const eo_tuple *in, eo_tuple *tmp, eo_tuple *out,
                            uint32_t *part_start, uint32_t *part_end, void *scratch, uint32_t *err, int cus,
                            hipStream_t stream, unsigned phases = kGapPrepare | kGapScatter);

// ---- batch.hip --------------------------------------------------------------
// One relation of one join to partition in a batch group (all of a group's
// relations use the same number of radix bits, digit = key & (2^bits - 1)).
struct BatchSeg {
  const uint2 *in;           // 16-byte aligned
  uint2 *out;
  uint64_t n;
  uint32_t stripe_base;      // first stripe of this relation in the group
  uint32_t nstripes;
  uint64_t count_base;       // first count of this relation: counts laid out [rel][digit][stripe]
  uint32_t *part_start;      // 2^bits + 1 partition starts (output)
  // sampled groups: partitions sized from a 1/16 sample (gaps between them)
  uint32_t samp_base;        // first sample workgroup of this relation in the group
  uint32_t nsamp;
  uint32_t *hs;              // batch_hs_words(bits): 2^bits sampled counts, then the clumping sum and the
                             // lines it was measured on (zero on entry)
  uint32_t *re, *cur;        // 2^bits region ends and cursors
  uint32_t *part_end;        // 2^bits partition ends (output)
  uint64_t budget;           // tuples of out
};
struct BatchGroup {
  int bits;
  int sampled;               // partitions sized from a sample, runs reserved with atomics
  const BatchSeg *segs;      // device table, ordered by stripe_base (and samp_base)
  uint32_t nseg;
  uint32_t total_stripes;
  uint64_t stripe_len;
  uint64_t ncounts;          // sum over relations of 2^bits * nstripes (exact groups)
  uint32_t total_samp;       // sample workgroups (sampled groups)
  uint32_t *hs_all;          // the group's sampled counts, nseg * 2^bits, contiguous
};
// Tuples per partition tile (the fast scatter's LDS staging tile), and the
// tiles one sample workgroup of a sampled batch relation covers.
constexpr uint64_t kPartTile = 16384;
constexpr uint32_t kBatchSampleTiles = 16;
constexpr int kBatchSampleMaxBits = 8;  // sampled batch groups: up to 8 radix bits
// words of a sampled batch relation's counts: the 2^bits digits, the sum of
// same-digit line counts and the number of lines it was measured on
inline constexpr uint32_t batch_hs_words(int bits) { return (1u << bits) + 2u; }
// One join of a batch, as the persistent probe sees it.
struct BatchJoin {
  const uint2 *R, *S;        // partitioned (or raw when P == 1)
  const uint32_t *r_start, *s_start;  // P + 1 each
  const uint32_t *r_end, *s_end;      // partition ends: r_start + 1 when contiguous
  uint32_t P, s_span;
  uint2 *out;
  uint64_t out_cap;
  uint64_t *cursor;          // match count / output reservation (zero on entry)
  uint32_t pair_base;        // first (join, partition) pair
  uint32_t wide;             // 16-byte RESULT rows
  uint32_t *matched;         // left outer join: bitmap over S positions (zero on entry), or NULL
  uint64_t nS;
  uint32_t swapped;          // R and S are the request's S and R (rows written with the payloads exchanged)
  uint32_t pad;
};
// One probe tile of a batch, resolved ahead of the probe so that a workgroup
// starts a tile's loads after a single descriptor read: R chunk (already
// offset), the join's S and output, and the S span.
// R tuples per probe tile: the batch probe's LDS table (the half-size table,
// two 1024-thread workgroups per CU)
constexpr uint32_t kBatchChunk = 8704;
struct BatchTile {
  const uint2 *R;            // first tuple of the R chunk
  const uint2 *S;            // the join's S (partitioned)
  uint32_t j;                // join
  uint32_t r_n;              // R tuples of the chunk
  uint32_t s_beg, s_n;       // S span
};
uint64_t batch_stripe_len(int bits);  // tuples per partition workgroup of a `bits`-bit group
// err: the device flags; a sampled group raises kErrCapacity there when a
// region proved too small.
hipError_t batch_partition_group(const BatchGroup &g, uint64_t *counts, void *scan_scratch, uint32_t *err,
                                 hipStream_t stream);
// Writes pair_join (npairs u32: the join of every (join, partition) pair),
// tile_start (npairs + 1 u32: scanned probe tile numbers of each pair) and
// the descriptors of the first min(total, tiles_cap) tiles.
hipError_t batch_plan(const BatchJoin *joins, uint32_t njoins, uint32_t *pair_join, uint32_t npairs,
                      uint32_t *tile_start, BatchTile *tiles, uint32_t tiles_cap, void *scan_scratch, uint32_t *err,
                      hipStream_t stream);
// Left outer joins of a batch: one null row per unmatched S tuple of every
// join with a bitmap; emit_start (njoins + 1 u32) is the prefix of each
// join's unmatched_blocks(nS), emit_start[njoins] the grid.
hipError_t batch_emit_unmatched(const BatchJoin *joins, uint32_t njoins, const uint32_t *emit_start,
                                uint32_t blocks, hipStream_t stream);
// *queue_head must be zero on entry; grid = one workgroup per CU.  Tiles
// below tiles_cap read their descriptor; later ones resolve their pair by
// search.
hipError_t batch_probe(const BatchJoin *joins, const uint32_t *pair_join, const uint32_t *tile_start, uint32_t npairs,
                       const BatchTile *tiles, uint32_t tiles_cap, uint32_t *queue_head, int cus,
                       hipStream_t stream);

// ---- gen.hip ----------------------------------------------------------------
hipError_t gen_build(eo_tuple *out, uint64_t n, uint64_t seed, hipStream_t s);
hipError_t gen_probe(eo_tuple *out, uint64_t n, uint64_t nR, double selectivity, uint64_t seed, hipStream_t s);

// ---- nlj.hip ----------------------------------------------------------------
struct NljPlan {
  uint64_t nR, nS;
  uint32_t r_tiles, s_tiles;  // grid = r_tiles x s_tiles workgroups
  uint32_t r_tile, s_tile;    // tuples per tile
};
NljPlan plan_nlj(uint64_t nR, uint64_t nS);
uint64_t nlj_count_slots(const NljPlan &p);  // per-wave count entries
// matched (optional): bitmap over S positions, bits set for S tuples with a match.
hipError_t nlj_count(const NljPlan &p, const eo_tuple *R, const eo_tuple *S, uint64_t *wave_count,
                     hipStream_t stream, uint32_t *matched = nullptr);
// One pass: every match counted into *cursor (zero on entry); with out != NULL
// the pairs are written at positions reserved per wave (order unspecified,
// positions >= out_cap counted only).
hipError_t nlj_fused(const NljPlan &p, const eo_tuple *R, const eo_tuple *S, uint64_t *cursor, eo_pair *out,
                     uint64_t out_cap, bool wide, uint32_t *matched, hipStream_t stream);
// Distance-predicate NLJ over reference 3-D points: pairs (R id, S id) with
// squared distance < d2, counted into *cursor (zero on entry), written when
// out != NULL (positions >= out_cap counted only).
hipError_t nlj_dist(const eo_point3 *R, uint64_t nR, const eo_point3 *S, uint64_t nS, double d2, uint64_t *cursor,
                    eo_pair *out, uint64_t out_cap, hipStream_t stream);
hipError_t nlj_write(const NljPlan &p, const eo_tuple *R, const eo_tuple *S, const uint64_t *wave_off,
                     eo_pair *out, uint64_t out_cap, bool wide, hipStream_t stream);

}  // namespace eo
