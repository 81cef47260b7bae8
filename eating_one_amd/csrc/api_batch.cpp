// C ABI of resident build tables and batched joins.
//
// The reference runs one join per process: it re-creates the CUDA context,
// re-partitions both relations and synchronises after every launch
// (joins/hashJoin_gpu/main_street/hj.cc:171-1269).  A batch here plans all of
// its joins on the host in one pass, uploads one descriptor blob, and then
// enqueues a fixed launch sequence with no host synchronisation:
//   fill (cursors, queue head) -> per bit count: group hist, scan, scatter,
//   starts -> individual (two-pass or misaligned) partitions -> pair table and
//   tile plan -> persistent probe -> copy of the per-join counts to the host.
// Because that sequence depends only on the descriptors, it is also captured
// once into a hipGraph and replayed by eo_batch_replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "device.hpp"
#include "eojoin.h"
#include "kernels.hpp"
#include "plan.hpp"
#include "table.hpp"


namespace eo {
namespace {

constexpr int kMaxGroupBits = kMaxPassBits;  // one fast pass per group
constexpr uint64_t kMaxTileDescs = 1ull << 21;  // 64 MiB of probe tile descriptors per batch
constexpr uint32_t kBatchSampleMinJoins = 256;  // and in batches of at least 256 requests

// Where one relation of one join comes from and goes to.
struct SideInfo {
  const eo_tuple *in = nullptr;  // raw input
  uint64_t n = 0;
  int mode = 0;                  // 0: as is (P == 1), 1: group, 2: individual, 3: resident table
  size_t o_part = 0, o_start = 0;  // arena offsets of the partitioned copy and its starts
  // mode 1 with partitions sized from a sample: the copy holds `budget`
  // tuples, partitions [start, end) with gaps; region ends and cursors
  bool sampled = false;
  uint64_t budget = 0;
  size_t o_end = 0, o_re = 0, o_cur = 0;
  size_t blob_start = 0;         // blob offset of {0, n} when mode == 0
  const eo_tuple *table_part = nullptr;
  const uint32_t *table_start = nullptr;
};

struct JoinInfo {
  uint32_t req = 0;  // request index
  HashPlan hp;
  SideInfo r, s;
  uint64_t bits_word = 0;  // left outer: first word of this join's bitmap in the batch's bitmap region
  uint32_t eblocks = 0;    // left outer: blocks of the null-row kernel
  bool swapped = false;    // r is the request's S and s its R (built on the smaller relation)
};

}  // namespace
}  // namespace eo

struct eo_batch {
  eo_ctx *ctx = nullptr;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;  // stream of the batch in flight
  eo::Arena arena;
  char *h_blob = nullptr;  // pinned descriptor staging
  size_t h_blob_cap = 0;
  uint64_t *h_counts = nullptr;  // pinned per-join counts
  size_t h_counts_cap = 0;
  hipEvent_t done = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;  // batch start / end (stats only)
  eo::LaunchTimer timer;
  bool timed = false;
  eo_join_stats *stats = nullptr;
  bool in_flight = false;
  // set once a submit has enqueued its whole batch; cleared when a submit
  // starts, so a batch whose planning or enqueue failed is never replayed
  bool replayable = false;
  // last submitted batch
  std::vector<eo_request> reqs;
  std::vector<eo::JoinInfo> joins;
  std::vector<int> req_join;  // join index of a request, -1 if it has no work
  // enqueue parameters of the last batch (device addresses inside arena)
  struct Group {
    eo::BatchGroup g;
    size_t o_hs = 0;  // sampled groups: the relations' sampled counts (arena)
  };
  std::vector<Group> groups;
  size_t o_blob = 0, blob_bytes = 0, o_cursor = 0, o_head = 0, o_counts = 0, o_cscr = 0, o_tmp = 0, o_pscr = 0,
         o_pair = 0, o_tiles = 0, o_tscr = 0, o_tdesc = 0;
  uint32_t tiles_cap = 0;  // probe tile descriptors written ahead of the probe
  size_t o_joins_blob = 0;
  uint32_t npairs = 0;
  bool wide = false, outer = false;  // batch flags: EO_F_RESULT16, EO_F_LEFT_OUTER
  // partitions: exact (histogram) only, or sampled for large one-pass relations;
  // any_sampled: some group of the batch in flight is sampled (its capacity
  // flag follows the per-join counts in the cursor array and the readback)
  bool exact = false, any_sampled = false;
  eo_join_opts sub_opts{};  // the last submit's options (for an exact redo)
  bool has_opts = false;
  size_t o_bits = 0, o_emit = 0;     // bitmap region (arena), emit_start (blob)
  uint64_t bits_words = 0;
  uint32_t emit_blocks = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace eo {
namespace {

uint64_t part_bytes(uint64_t n) { return n * sizeof(eo_tuple); }

// Host-only plan of a batch: which relation is partitioned how, the arena
// layout and the blob layout.  Fills b->joins, b->req_join, b->groups (without
// device addresses), and the arena offsets.
eo_status plan_batch(eo_batch *b, const eo_request *reqs, uint32_t n, const eo_join_opts *o, Layout &L,
                     size_t &blob_bytes, std::vector<std::vector<std::pair<uint32_t, int>>> &group_members) {
  eo_ctx *c = b->ctx;
  const int cus = c->props.cus;
  eo_join_opts po{};
  if (o) po = *o;
  b->wide = o && (o->flags & EO_F_RESULT16);
  b->outer = o && (o->flags & EO_F_LEFT_OUTER);
  b->bits_words = 0;
  b->emit_blocks = 0;
  po.stats = nullptr;
  po.stream = nullptr;
  po.flags = 0;
  po.where = EO_DEVICE;
  b->joins.clear();
  b->req_join.assign(n, -1);
  group_members.assign(2 * (kMaxGroupBits + 1), {});  // [bits]: exact groups, [kMaxGroupBits + 1 + bits]: sampled
  b->any_sampled = false;
  uint64_t max_tmp = 0;
  size_t max_pscr = 0;
  uint64_t npairs = 0;
  uint64_t tiles_est = 0;
  // S tuples of the whole batch: the probe tiles of all joins share one
  // persistent launch, so each join's S spans are planned against this pool
  // (long spans, few LDS table builds) rather than as if it ran alone
  uint64_t s_pool = 0;  // probe-side tuples (the larger relation of a join built on its smaller one)
  for (uint32_t i = 0; i < n; ++i) {
    const uint64_t nr = reqs[i].table ? reqs[i].table->n : reqs[i].nR;
    const uint64_t np = (!reqs[i].table && !b->outer && reqs[i].nS < nr) ? nr : reqs[i].nS;
    s_pool += std::min<uint64_t>(np, kMaxRows);
  }
  for (uint32_t i = 0; i < n; ++i) {
    const eo_request &q = reqs[i];
    const uint64_t nR = q.table ? q.table->n : q.nR;
    if (nR > kMaxRows || q.nS > kMaxRows) return EO_EINVAL;
    if ((!q.table && nR && !q.R) || (q.nS && !q.S) || (q.out_cap && !q.out)) return EO_EINVAL;
    if (q.table && q.table->ctx != c) return EO_EINVAL;
    if (q.nS == 0 || (nR == 0 && !b->outer)) continue;  // no rows (an outer join has a null row per S tuple)
    JoinInfo J;
    J.req = i;
    // build side (LDS tables) and probe side: an inner join without a
    // resident table builds on its smaller relation (the probe exchanges the
    // payloads of its rows back), so that the partition bits follow it
    const eo_tuple *bld = q.R, *prb = q.S;
    uint64_t nB = nR, nP = q.nS;
    if (!q.table && !b->outer && q.nS < nR) {
      std::swap(bld, prb);
      std::swap(nB, nP);
      J.swapped = true;
    }
    if (b->outer) {
      J.bits_word = b->bits_words;
      b->bits_words += (matched_words(q.nS) + 63) / 64 * 64;  // 256-byte aligned per join
      J.eblocks = (uint32_t)unmatched_blocks(q.nS);
      if ((uint64_t)b->emit_blocks + J.eblocks >= (1ull << 31)) return EO_EINVAL;
      b->emit_blocks += J.eblocks;
    }
    if (q.table) {
      eo_join_opts to = po;
      to.radix_bits = q.table->hp.bits ? q.table->hp.bits : -1;
      to.passes = q.table->hp.passes;
      eo_status st = plan_hash_join(nR, q.nS, &to, &J.hp, cus);
      if (st != EO_OK) return st;
      if (to.s_tile == 0) J.hp.s_span = plan_span(nR, q.nS, J.hp.lds_cap, J.hp.bits, s_pool, cus);
    } else {
      eo_join_opts to = po;
      apply_tuned(c, nB, nP, &to);
      // partitions sized for the batch probe's half-size LDS table
      if (!to.lds_tuples) to.lds_tuples = 8192;
      eo_status st = plan_hash_join(nB, nP, &to, &J.hp, cus);
      if (st != EO_OK) return st;
      if (to.s_tile == 0) J.hp.s_span = plan_span(nB, nP, J.hp.lds_cap, J.hp.bits, s_pool, cus);
    }
    const uint32_t P = 1u << J.hp.bits;
    auto side = [&](SideInfo &sd, const eo_tuple *in, uint64_t cnt, int which) {
      sd.in = in;
      sd.n = cnt;
      if (J.hp.bits == 0) {
        sd.mode = 0;
        return;
      }
      const bool group = J.hp.passes == 1 && J.hp.bits <= kMaxGroupBits && aligned16(in);
      // large relations of a large inner batch with few digits: partitions
      // sized from a sample (no histogram pass; a batch whose sample proves
      // too small is redone exactly when it is waited for).  Every tile
      // reserves its runs with one atomic per digit on cursors that all tiles
      // of its relation share: from 9 bits on (512 atomics per tile) that
      // costs more than the histogram it replaces, and in a batch of few
      // joins the tiles of one relation reserve all at once (measured, C5 in
      // batches of 30 requests: 227 ms sampled up to 11 bits, 200 ms up to 8
      // bits, 180 ms exact; in batches of 1024: 83.7 ms sampled up to 8 bits,
      // 90.8 ms exact).
      sd.sampled = group && n >= kBatchSampleMinJoins && J.hp.bits <= kBatchSampleMaxBits && !b->outer &&
                   !b->exact && cnt >= kGapMinRows && gap_rows(cnt, P) + cnt < (1ull << 32);
      sd.budget = sd.sampled ? gap_rows(cnt, P) : cnt;
      sd.o_part = L.add(part_bytes(sd.budget));
      sd.o_start = L.add((P + 1) * sizeof(uint32_t));
      if (sd.sampled) {
        sd.o_end = L.add(P * sizeof(uint32_t));
        sd.o_re = L.add(P * sizeof(uint32_t));
        sd.o_cur = L.add(P * sizeof(uint32_t));
        b->any_sampled = true;
      }
      if (group) {
        sd.mode = 1;
        group_members[J.hp.bits + (sd.sampled ? kMaxGroupBits + 1 : 0)].push_back({(uint32_t)b->joins.size(), which});
      } else {
        sd.mode = 2;
        if (J.hp.passes == 2) max_tmp = std::max(max_tmp, cnt);
        max_pscr = std::max(max_pscr, partition_scratch(cnt, J.hp, cus));
      }
    };
    if (q.table) {
      J.r.mode = 3;
      J.r.n = nR;
      J.r.table_part = q.table->part;
      J.r.table_start = q.table->start;
    } else {
      side(J.r, bld, nB, 0);
    }
    side(J.s, prb, nP, 1);
    npairs += P;
    if (npairs >= (1ull << 31)) return EO_EINVAL;
    if (nB && nP) {
      // probe tiles if the keys spread evenly, with a spare S span per partition
      const uint64_t chunks = std::max<uint64_t>(1, (nB / P + kBatchChunk - 1) / kBatchChunk);
      tiles_est += P * chunks * ((nP / P + J.hp.s_span - 1) / J.hp.s_span + 1);
    }
    b->req_join[i] = (int)b->joins.size();
    b->joins.push_back(J);
  }
  const uint32_t nj = (uint32_t)b->joins.size();
  // blob: joins, then the segment table of every group, then {0, n} starts
  size_t blob = (size_t)nj * sizeof(BatchJoin);
  b->o_joins_blob = 0;
  b->groups.clear();
  uint64_t max_counts = 0;
  for (int gi = 1; gi < 2 * (kMaxGroupBits + 1); ++gi) {
    const int bits = gi % (kMaxGroupBits + 1);
    const bool sampled = gi > kMaxGroupBits;
    auto &mem = group_members[gi];
    if (bits == 0 || mem.empty()) continue;
    const uint64_t len = batch_stripe_len(bits);
    eo_batch::Group G{};
    G.g.bits = bits;
    G.g.sampled = sampled ? 1 : 0;
    G.g.nseg = (uint32_t)mem.size();
    G.g.stripe_len = len;
    blob = (blob + 15) / 16 * 16;
    G.g.segs = reinterpret_cast<const BatchSeg *>(blob);  // blob offset until addresses are known
    blob += mem.size() * sizeof(BatchSeg);
    uint64_t stripes = 0, counts = 0, samp = 0;
    for (auto &m : mem) {
      const SideInfo &sd = m.second ? b->joins[m.first].s : b->joins[m.first].r;
      const uint64_t ns = (sd.n + len - 1) / len;
      stripes += ns;
      if (sampled) samp += (sd.n + kPartTile * kBatchSampleTiles - 1) / (kPartTile * kBatchSampleTiles);
      else counts += ns << bits;
    }
    if (stripes >= (1ull << 31) || samp >= (1ull << 31)) return EO_EINVAL;
    G.g.total_stripes = (uint32_t)stripes;
    G.g.ncounts = counts;
    G.g.total_samp = (uint32_t)samp;
    if (sampled) G.o_hs = L.add(mem.size() * sizeof(uint32_t) * batch_hs_words(bits));
    max_counts = std::max(max_counts, counts);
    b->groups.push_back(G);
  }
  for (auto &J : b->joins) {
    for (SideInfo *sd : {&J.r, &J.s}) {
      if (sd->mode != 0) continue;
      sd->blob_start = blob;
      blob += 2 * sizeof(uint32_t);
    }
  }
  if (b->outer) {
    b->o_emit = blob;
    blob += ((size_t)nj + 1) * sizeof(uint32_t);
  }
  blob_bytes = (blob + 15) / 16 * 16;
  b->o_blob = L.add(blob_bytes);
  b->o_bits = L.add(b->bits_words * sizeof(uint32_t));
  b->o_cursor = L.add(((size_t)nj + 1) * sizeof(uint64_t));  // per-join counts, then the capacity flag
  b->o_head = L.add(sizeof(uint64_t));
  b->o_counts = L.add(max_counts * sizeof(uint64_t));
  b->o_cscr = L.add(max_counts ? scan_scratch_bytes_u64(max_counts) : 0);
  b->o_tmp = L.add(part_bytes(max_tmp));
  b->o_pscr = L.add(max_pscr);
  b->o_pair = L.add(npairs * sizeof(uint32_t));
  b->o_tiles = L.add((npairs + 1) * sizeof(uint32_t));
  b->o_tscr = L.add(scan_scratch_bytes_u32(npairs + 1));
  // descriptors for the expected tiles (at most 64 MiB); tiles beyond them,
  // e.g. of skewed partitions, resolve themselves in the probe
  b->tiles_cap = (uint32_t)std::min<uint64_t>(tiles_est, kMaxTileDescs);
  b->o_tdesc = L.add((size_t)b->tiles_cap * sizeof(BatchTile));
  b->npairs = (uint32_t)npairs;
  return EO_OK;
}

// Fills the pinned blob with device addresses (arena base known).
void fill_blob(eo_batch *b, const std::vector<std::vector<std::pair<uint32_t, int>>> &group_members) {
  char *A = b->arena.base;
  char *H = b->h_blob;
  char *D = A + b->o_blob;
  BatchJoin *hj = reinterpret_cast<BatchJoin *>(H + b->o_joins_blob);
  uint64_t *cursor = reinterpret_cast<uint64_t *>(A + b->o_cursor);
  uint32_t pair_base = 0;
  for (size_t j = 0; j < b->joins.size(); ++j) {
    const JoinInfo &J = b->joins[j];
    const eo_request &q = b->reqs[J.req];
    BatchJoin bj{};
    auto src = [&](const SideInfo &sd, const uint2 **tup, const uint32_t **st, const uint32_t **en) {
      if (sd.mode == 0) {
        *tup = reinterpret_cast<const uint2 *>(sd.in);
        uint32_t *h = reinterpret_cast<uint32_t *>(H + sd.blob_start);
        h[0] = 0;
        h[1] = (uint32_t)sd.n;
        *st = reinterpret_cast<const uint32_t *>(D + sd.blob_start);
      } else if (sd.mode == 3) {
        *tup = reinterpret_cast<const uint2 *>(sd.table_part);
        *st = sd.table_start;
      } else {
        *tup = reinterpret_cast<const uint2 *>(A + sd.o_part);
        *st = reinterpret_cast<const uint32_t *>(A + sd.o_start);
      }
      *en = sd.sampled ? reinterpret_cast<const uint32_t *>(A + sd.o_end) : *st + 1;  // gapped or contiguous
    };
    src(J.r, &bj.R, &bj.r_start, &bj.r_end);
    src(J.s, &bj.S, &bj.s_start, &bj.s_end);
    bj.P = 1u << J.hp.bits;
    bj.s_span = J.hp.s_span;
    bj.out = reinterpret_cast<uint2 *>(q.out);
    bj.out_cap = q.out ? q.out_cap : 0;
    bj.cursor = cursor + j;
    bj.pair_base = pair_base;
    bj.wide = b->wide ? 1u : 0u;
    bj.matched = b->outer ? reinterpret_cast<uint32_t *>(A + b->o_bits) + J.bits_word : nullptr;
    bj.nS = J.s.n;
    bj.swapped = J.swapped ? 1u : 0u;
    pair_base += bj.P;
    hj[j] = bj;
  }
  if (b->outer) {
    uint32_t *es = reinterpret_cast<uint32_t *>(H + b->o_emit);
    uint32_t acc = 0;
    for (size_t j = 0; j < b->joins.size(); ++j) {
      es[j] = acc;
      acc += b->joins[j].eblocks;
    }
    es[b->joins.size()] = acc;
  }
  size_t gi = 0;
  for (int gk = 1; gk < 2 * (kMaxGroupBits + 1); ++gk) {
    const int bits = gk % (kMaxGroupBits + 1);
    auto &mem = group_members[gk];
    if (bits == 0 || mem.empty()) continue;
    const uint64_t len = batch_stripe_len(bits);
    eo_batch::Group &G = b->groups[gi++];
    const size_t off = reinterpret_cast<size_t>(G.g.segs);
    BatchSeg *hs = reinterpret_cast<BatchSeg *>(H + off);
    uint32_t stripe_base = 0, samp_base = 0;
    uint64_t count_base = 0;
    uint32_t *hs_all = G.g.sampled ? reinterpret_cast<uint32_t *>(A + G.o_hs) : nullptr;
    for (size_t k = 0; k < mem.size(); ++k) {
      const SideInfo &sd = mem[k].second ? b->joins[mem[k].first].s : b->joins[mem[k].first].r;
      BatchSeg sg{};
      sg.in = reinterpret_cast<const uint2 *>(sd.in);
      sg.out = reinterpret_cast<uint2 *>(A + sd.o_part);
      sg.n = sd.n;
      sg.stripe_base = stripe_base;
      sg.nstripes = (uint32_t)((sd.n + len - 1) / len);
      sg.count_base = count_base;
      sg.part_start = reinterpret_cast<uint32_t *>(A + sd.o_start);
      if (G.g.sampled) {
        const uint64_t unit = kPartTile * kBatchSampleTiles;
        sg.samp_base = samp_base;
        sg.nsamp = (uint32_t)((sd.n + unit - 1) / unit);
        sg.hs = hs_all + (uint64_t)k * batch_hs_words(bits);
        sg.re = reinterpret_cast<uint32_t *>(A + sd.o_re);
        sg.cur = reinterpret_cast<uint32_t *>(A + sd.o_cur);
        sg.part_end = reinterpret_cast<uint32_t *>(A + sd.o_end);
        sg.budget = sd.budget;
        samp_base += sg.nsamp;
      } else {
        count_base += (uint64_t)sg.nstripes << bits;
      }
      stripe_base += sg.nstripes;
      hs[k] = sg;
    }
    G.g.segs = reinterpret_cast<const BatchSeg *>(D + off);
    G.g.hs_all = hs_all;
  }
}

// The device work of a planned batch: no host synchronisation, no memset
// nodes, so that it can be captured into a graph.
eo_status enqueue_batch(eo_batch *b, hipStream_t s) {
  eo_ctx *c = b->ctx;
  char *A = b->arena.base;
  const uint32_t nj = (uint32_t)b->joins.size();
  uint64_t *cursor = reinterpret_cast<uint64_t *>(A + b->o_cursor);
  uint32_t *head = reinterpret_cast<uint32_t *>(A + b->o_head);
  if (nj == 0) return EO_OK;
  EO_TRY_HIP(fill_u64(cursor, 0, (uint64_t)nj + 1, s), "cursor reset");  // counts and the capacity flag
  uint32_t *cap_flag = reinterpret_cast<uint32_t *>(cursor + nj);
  EO_TRY_HIP(fill_u32(head, 0u, 1, s), "queue reset");
  if (b->outer)
    EO_TRY_HIP(fill_u32(reinterpret_cast<uint32_t *>(A + b->o_bits), 0u, b->bits_words, s), "match bitmap reset");
  uint64_t *counts = reinterpret_cast<uint64_t *>(A + b->o_counts);
  for (auto &G : b->groups)
    EO_TRY_HIP(batch_partition_group(G.g, counts, A + b->o_cscr, G.g.sampled ? cap_flag : c->d_err, s),
               "batch partition");
  for (auto &J : b->joins) {
    for (const SideInfo *sd : {&J.r, &J.s}) {
      if (sd->mode != 2) continue;
      eo_status st = partition_relation(c, sd->in, sd->n, J.hp, reinterpret_cast<eo_tuple *>(A + b->o_tmp),
                                        reinterpret_cast<eo_tuple *>(A + sd->o_part),
                                        reinterpret_cast<uint32_t *>(A + sd->o_start), A + b->o_pscr, s);
      if (st != EO_OK) return st;
    }
  }
  const BatchJoin *dj = reinterpret_cast<const BatchJoin *>(A + b->o_blob + b->o_joins_blob);
  uint32_t *pair_join = reinterpret_cast<uint32_t *>(A + b->o_pair);
  uint32_t *tile_start = reinterpret_cast<uint32_t *>(A + b->o_tiles);
  BatchTile *tdesc = reinterpret_cast<BatchTile *>(A + b->o_tdesc);
  EO_TRY_HIP(batch_plan(dj, nj, pair_join, b->npairs, tile_start, tdesc, b->tiles_cap, A + b->o_tscr, c->d_err, s),
             "batch plan");
  EO_TRY_HIP(batch_probe(dj, pair_join, tile_start, b->npairs, tdesc, b->tiles_cap, head, c->props.cus, s),
             "batch probe");
  if (b->outer)
    EO_TRY_HIP(batch_emit_unmatched(dj, nj, reinterpret_cast<const uint32_t *>(A + b->o_blob + b->o_emit),
                                    b->emit_blocks, s),
               "outer rows");
  EO_TRY_HIP(hipMemcpyAsync(b->h_counts, cursor, ((size_t)nj + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, s),
             "count readback");
  return EO_OK;
}

void drop_graph(eo_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

}  // namespace
}  // namespace eo

using namespace eo;

extern "C" {

eo_status eo_table_create(eo_ctx *c, const eo_tuple *R, uint64_t nR, const eo_join_opts *o, eo_table **out) {
  if (!c || !out || !opts_ok(o) || nR > kMaxRows || (nR && !R)) return EO_EINVAL;
  *out = nullptr;
  DeviceGuard g(c->device);
  HashPlan hp;
  // partitions sized for the batch probe's half-size LDS table, unless the
  // caller chose a table size
  eo_join_opts to{};
  if (o) to = *o;
  if (!to.lds_tuples) to.lds_tuples = 8192;
  eo_status st = plan_hash_join(nR, nR, &to, &hp, c->props.cus);
  if (st != EO_OK) return st;
  hipStream_t s = pick_stream(c, o);
  const bool host = where_of(o) == EO_HOST;
  const uint32_t P = 1u << hp.bits;
  eo_table *t = new eo_table();
  t->ctx = c;
  t->n = nR;
  t->hp = hp;
  Layout T;
  const size_t o_part = T.add(part_bytes(nR));
  const size_t o_start = T.add((P + 1) * sizeof(uint32_t));
  t->bytes = T.total;
  hipError_t e = hipMalloc(&t->mem, std::max<size_t>(T.total, 256));
  if (e != hipSuccess) {
    delete t;
    return hip_fail(e, "table allocation");
  }
  t->part = reinterpret_cast<eo_tuple *>((char *)t->mem + o_part);
  t->start = reinterpret_cast<uint32_t *>((char *)t->mem + o_start);
  auto fail = [&](eo_status code) {
    (void)hipStreamSynchronize(s);
    (void)hipFree(t->mem);
    delete t;
    return code;
  };
  Layout L;
  const size_t o_in = L.add(host && hp.bits ? part_bytes(nR) : 0);
  const size_t o_tmp = L.add(hp.passes == 2 ? part_bytes(nR) : 0);
  const size_t o_scr = L.add(partition_scratch(nR, hp, c->props.cus));
  e = hipStreamSynchronize(s);
  if (e == hipSuccess) e = c->arena.reserve(L.total);
  if (e != hipSuccess) return fail(hip_fail(e, "table scratch"));
  char *A = c->arena.base;
  if (hp.bits == 0) {
    e = hipMemcpyAsync(t->part, R, part_bytes(nR), host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = fill_u32(t->start, 0u, 1, s);
    if (e == hipSuccess) e = fill_u32(t->start + 1, (uint32_t)nR, 1, s);
    if (e != hipSuccess) return fail(hip_fail(e, "table copy"));
  } else {
    const eo_tuple *dR = R;
    if (host) {
      e = hipMemcpyAsync(A + o_in, R, part_bytes(nR), hipMemcpyHostToDevice, s);
      if (e != hipSuccess) return fail(hip_fail(e, "h2d R"));
      dR = reinterpret_cast<const eo_tuple *>(A + o_in);
    }
    st = partition_relation(c, dR, nR, hp, reinterpret_cast<eo_tuple *>(A + o_tmp), t->part, t->start, A + o_scr, s);
    if (st != EO_OK) return fail(st);
  }
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(hip_fail(e, "table build"));
  st = check_device_errors(c, s);
  if (st != EO_OK) return fail(st);
  *out = t;
  return EO_OK;
}

eo_status eo_table_query(const eo_table *t, uint64_t *n, int32_t *radix_bits, uint64_t *bytes) {
  if (!t) return EO_EINVAL;
  if (n) *n = t->n;
  if (radix_bits) *radix_bits = t->hp.bits;
  if (bytes) *bytes = t->bytes;
  return EO_OK;
}

eo_status eo_table_free(eo_table *t) {
  if (!t) return EO_EINVAL;
  DeviceGuard g(t->ctx->device);
  if (t->mem) (void)hipFree(t->mem);
  delete t;
  return EO_OK;
}

eo_status eo_batch_create(eo_ctx *c, eo_batch **out) {
  if (!c || !out) return EO_EINVAL;
  DeviceGuard g(c->device);
  eo_batch *b = new eo_batch();
  b->ctx = c;
  hipError_t e = hipStreamCreateWithFlags(&b->own_stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreate(&b->t0);
  if (e == hipSuccess) e = hipEventCreate(&b->t1);
  for (int i = 0; i < 2 * kMaxTimedLaunches && e == hipSuccess; ++i) e = hipEventCreate(&b->timer.ev[i]);
  if (e != hipSuccess) {
    eo_status st = hip_fail(e, "eo_batch_create");
    eo_batch_free(b);
    return st;
  }
  b->stream = b->own_stream;
  *out = b;
  return EO_OK;
}

eo_status eo_batch_free(eo_batch *b) {
  if (!b) return EO_EINVAL;
  DeviceGuard g(b->ctx->device);
  if (b->in_flight && b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  b->arena.release();
  if (b->h_blob) (void)hipHostFree(b->h_blob);
  if (b->h_counts) (void)hipHostFree(b->h_counts);
  for (auto &ev : b->timer.ev)
    if (ev) (void)hipEventDestroy(ev);
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->t0) (void)hipEventDestroy(b->t0);
  if (b->t1) (void)hipEventDestroy(b->t1);
  if (b->own_stream) (void)hipStreamDestroy(b->own_stream);
  delete b;
  return EO_OK;
}

eo_status eo_batch_footprint(eo_ctx *c, const eo_request *reqs, uint32_t n, uint64_t *bytes) {
  if (!c || !bytes || (n && !reqs)) return EO_EINVAL;
  eo_batch tmp;
  tmp.ctx = c;
  Layout L;
  size_t blob = 0;
  std::vector<std::vector<std::pair<uint32_t, int>>> members;
  eo_status st = plan_batch(&tmp, reqs, n, nullptr, L, blob, members);
  if (st != EO_OK) return st;
  *bytes = L.total;
  return EO_OK;
}

}  // extern "C"

namespace eo {
namespace {

// eo_batch_submit; exact: every partition from a histogram pass (the redo of
// a batch whose sampled partitions proved too small).
eo_status submit_batch(eo_batch *b, const eo_request *reqs, uint32_t n, const eo_join_opts *o, bool exact) {
  if (b->in_flight) {
    set_last_error("eo_batch_submit: a batch is still in flight on this object");
    return EO_EINVAL;
  }
  eo_ctx *c = b->ctx;
  DeviceGuard g(c->device);
  drop_graph(b);
  b->replayable = false;
  b->exact = exact || (o && (o->flags & EO_F_EXACT_PARTITIONS));
  b->has_opts = o != nullptr;
  if (o) b->sub_opts = *o;
  if (reqs != b->reqs.data()) b->reqs.assign(reqs, reqs + n);
  Layout L;
  size_t blob_bytes = 0;
  std::vector<std::vector<std::pair<uint32_t, int>>> members;
  eo_status st = plan_batch(b, reqs, n, o, L, blob_bytes, members);
  if (st != EO_OK) return st;
  if (L.total > c->hbm_budget) {
    set_last_error("eo_batch_submit: batch scratch exceeds the HBM budget");
    return EO_EBUDGET;
  }
  b->blob_bytes = blob_bytes;
  hipStream_t s = (o && o->stream) ? (hipStream_t)o->stream : b->own_stream;
  b->stream = s;
  // grow the pinned staging and the arena (nothing of this batch is in flight)
  if (L.total > b->arena.cap) {
    EO_TRY_HIP(hipStreamSynchronize(s), "sync");
    EO_TRY_HIP(b->arena.reserve(L.total), "batch arena");
  }
  if (blob_bytes > b->h_blob_cap) {
    if (b->h_blob) (void)hipHostFree(b->h_blob);
    b->h_blob = nullptr;
    b->h_blob_cap = 0;
    EO_TRY_HIP(hipHostMalloc((void **)&b->h_blob, std::max<size_t>(blob_bytes, 4096), hipHostMallocDefault),
               "blob staging");
    b->h_blob_cap = std::max<size_t>(blob_bytes, 4096);
  }
  const size_t ncnt = b->joins.size() + 1;  // the counts and the capacity flag
  if (ncnt > b->h_counts_cap) {
    if (b->h_counts) (void)hipHostFree(b->h_counts);
    b->h_counts = nullptr;
    b->h_counts_cap = 0;
    EO_TRY_HIP(hipHostMalloc((void **)&b->h_counts, ncnt * sizeof(uint64_t), hipHostMallocDefault), "count staging");
    b->h_counts_cap = ncnt;
  }
  fill_blob(b, members);
  b->stats = o ? o->stats : nullptr;
  b->timed = b->stats != nullptr;
  if (b->stats) std::memset(b->stats, 0, sizeof(*b->stats));
  b->timer.n = 0;
  b->timer.overflow = false;
  LaunchTimer *prev = set_launch_timer(b->timed ? &b->timer : nullptr);
  hipError_t e = hipSuccess;
  if (blob_bytes) e = hipMemcpyAsync(b->arena.base + b->o_blob, b->h_blob, blob_bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && b->timed) e = hipEventRecord(b->t0, s);
  st = e == hipSuccess ? enqueue_batch(b, s) : hip_fail(e, "blob upload");
  set_launch_timer(prev);
  if (st == EO_OK && b->timed) {
    e = hipEventRecord(b->t1, s);
    if (e != hipSuccess) st = hip_fail(e, "event");
  }
  if (st != EO_OK) {
    (void)hipStreamSynchronize(s);
    return st;
  }
  EO_TRY_HIP(hipEventRecord(b->done, s), "event");
  b->in_flight = true;
  b->replayable = true;
  return EO_OK;
}

}  // namespace
}  // namespace eo

extern "C" {

eo_status eo_batch_submit(eo_batch *b, const eo_request *reqs, uint32_t n, const eo_join_opts *o) {
  if (!b || (n && !reqs) || !opts_ok(o)) return EO_EINVAL;
  return submit_batch(b, reqs, n, o, false);
}

eo_status eo_batch_replay(eo_batch *b) {
  if (!b) return EO_EINVAL;
  if (b->in_flight) {
    set_last_error("eo_batch_replay: a batch is still in flight on this object");
    return EO_EINVAL;
  }
  if (!b->replayable || (b->reqs.empty() && b->joins.empty())) {
    set_last_error("eo_batch_replay: no batch was submitted successfully on this object");
    return EO_EINVAL;
  }
  eo_ctx *c = b->ctx;
  DeviceGuard g(c->device);
  hipStream_t s = b->stream;
  LaunchTimer *prev = set_launch_timer(nullptr);
  if (!b->exec) {
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      set_launch_timer(prev);
      return hip_fail(e, "begin capture");
    }
    eo_status st = enqueue_batch(b, s);
    hipGraph_t gr = nullptr;
    e = hipStreamEndCapture(s, &gr);
    if (st != EO_OK || e != hipSuccess) {
      if (gr) (void)hipGraphDestroy(gr);
      set_launch_timer(prev);
      return st != EO_OK ? st : hip_fail(e, "end capture");
    }
    b->graph = gr;
    e = hipGraphInstantiate(&b->exec, gr, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      drop_graph(b);
      set_launch_timer(prev);
      return hip_fail(e, "graph instantiate");
    }
  }
  set_launch_timer(prev);
  EO_TRY_HIP(hipGraphLaunch(b->exec, s), "graph launch");
  EO_TRY_HIP(hipEventRecord(b->done, s), "event");
  b->stats = nullptr;
  b->timed = false;
  b->in_flight = true;
  return EO_OK;
}

eo_status eo_batch_poll(eo_batch *b) {
  if (!b) return EO_EINVAL;
  if (!b->in_flight) return EO_OK;
  DeviceGuard g(b->ctx->device);
  const hipError_t e = hipEventQuery(b->done);
  if (e == hipErrorNotReady) return EO_EAGAIN;
  if (e != hipSuccess) return hip_fail(e, "batch poll");
  return EO_OK;
}

eo_status eo_batch_wait(eo_batch *b, eo_result *res, uint32_t n) {
  if (!b || (n && !res)) return EO_EINVAL;
  if (!b->in_flight) return EO_EINVAL;
  if (n != b->reqs.size()) return EO_EINVAL;
  eo_ctx *c = b->ctx;
  DeviceGuard g(c->device);
  bool redone = false;
  for (;;) {
    b->in_flight = false;
    EO_TRY_HIP(hipEventSynchronize(b->done), "batch wait");
    eo_status st = check_device_errors(c, b->stream);
    if (st != EO_OK) return st;
    const size_t nj = b->joins.size();
    if (!(b->any_sampled && nj && (b->h_counts[nj] & kErrCapacity))) break;
    // a sampled partition was too small: its join's rows are incomplete, so
    // the whole batch runs again with exact partitions (same requests and
    // options, stats included)
    const eo_join_opts so = b->sub_opts;
    eo_join_stats *stats = b->stats;
    eo_join_opts ro = so;
    ro.stats = stats;
    st = submit_batch(b, b->reqs.data(), n, b->has_opts || stats ? &ro : nullptr, true);
    if (st != EO_OK) return st;
    redone = true;
  }
  uint64_t total = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const eo_request &q = b->reqs[i];
    const int j = b->req_join[i];
    eo_result r{};
    r.tag = q.tag;
    r.n_out = j >= 0 ? b->h_counts[j] : 0;
    r.status = (q.out && r.n_out > q.out_cap) ? EO_ECAPACITY : EO_OK;
    total += r.n_out;
    res[i] = r;
  }
  if (b->stats) {
    fold_timer(b->timer, b->stats);
    b->stats->n_out = total;
    b->stats->sampled = redone ? 2u : (b->any_sampled ? 1u : 0u);
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, b->t0, b->t1) == hipSuccess) b->stats->ms_total = ms;
    b->stats = nullptr;
  }
  return EO_OK;
}

}  // extern "C"
