// Host-only checks built with AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/test_host_sanitized.py compiles and runs this file with g++):
//   - the oracle (oracle/cpu_join.c) on seeded random inputs, every join
//     against brute force, partitions against their definition;
//   - the library's host-only code: the reference table / index file readers
//     and writers (well-formed, truncated and inconsistent files), the
//     autotuner's table file (round trip, malformed rows), the hash join
//     planner over a sweep of sizes and overrides, and the request loop's
//     admission policy (FIFO batches, budget, deadlines, cancellation) on
//     synthetic clocks.
// Exit status 0 and "host checks ok" on success; any sanitizer report aborts.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <utility>
#include <vector>

#include "ctx.hpp"
#include "eojoin.h"
#include "plan.hpp"
#include "serve_queue.hpp"

extern "C" {
typedef struct {
  int32_t key, payload;
} otuple;
typedef struct {
  int32_t r_payload, s_payload;
} opair;
typedef struct {
  int32_t id, pad;
  double v[3];
} opoint3;
uint64_t oracle_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n);
uint64_t oracle_scan_u64(const uint64_t *in, uint64_t *out, uint64_t n);
int oracle_partition(const otuple *in, uint64_t n, uint32_t nparts, otuple *out, uint64_t *part_start);
int oracle_partition_bits(const otuple *in, uint64_t n, int bits, int shift, otuple *out, uint64_t *part_start);
uint64_t oracle_hash_join(const otuple *R, uint64_t nR, const otuple *S, uint64_t nS, uint32_t nparts, opair *out,
                          uint64_t cap);
uint64_t oracle_hidx_join(const otuple *R, uint64_t nR, const otuple *S, uint64_t nS, uint32_t nbuckets, opair *out,
                          uint64_t cap);
uint64_t oracle_sidx_join(const otuple *R, uint64_t nR, const otuple *S, uint64_t nS, opair *out, uint64_t cap);
int oracle_sidx_build(const otuple *R, uint64_t nR, otuple *sorted);
uint64_t oracle_nlj_join(const otuple *R, uint64_t nR, const otuple *S, uint64_t nS, opair *out, uint64_t cap);
uint64_t oracle_nlj_dist_join(const opoint3 *R, uint64_t nR, const opoint3 *S, uint64_t nS, double distance,
                              opair *out, uint64_t cap);
uint64_t oracle_hash_join_mt(const otuple *R, uint64_t nR, const otuple *S, uint64_t nS, uint32_t nparts,
                             int nthreads, opair *out, uint64_t cap);
uint64_t oracle_sidx_join_mt(const otuple *R, uint64_t nR, const otuple *S, uint64_t nS, int nthreads, opair *out,
                             uint64_t cap);
}

static int g_checks = 0;
#define CHECK(cond)                                                            \
  do {                                                                         \
    ++g_checks;                                                                \
    if (!(cond)) {                                                             \
      std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
      std::exit(1);                                                            \
    }                                                                          \
  } while (0)

using Pairs = std::vector<std::pair<int32_t, int32_t>>;

static std::vector<otuple> rand_tuples(std::mt19937_64 &g, size_t n, int32_t key_range) {
  std::vector<otuple> t(n);
  for (size_t i = 0; i < n; ++i) t[i] = {(int32_t)(g() % (uint64_t)key_range), (int32_t)i};
  return t;
}

static Pairs sorted_pairs(const std::vector<opair> &v, uint64_t n) {
  Pairs p;
  for (uint64_t i = 0; i < n; ++i) p.push_back({v[i].r_payload, v[i].s_payload});
  std::sort(p.begin(), p.end());
  return p;
}

static Pairs brute(const std::vector<otuple> &R, const std::vector<otuple> &S) {
  Pairs p;
  for (const otuple &s : S)
    for (const otuple &r : R)
      if (r.key == s.key) p.push_back({r.payload, s.payload});
  std::sort(p.begin(), p.end());
  return p;
}

template <typename F>
static Pairs run_join(F f, uint64_t expect_hint) {
  std::vector<opair> out(expect_hint + 1);
  uint64_t n = f(out.data(), out.size());
  if (n > out.size()) {
    out.resize(n);
    n = f(out.data(), out.size());
  }
  return sorted_pairs(out, n);
}

static void check_oracle() {
  std::mt19937_64 g(0xE01C);
  for (int round = 0; round < 6; ++round) {
    const size_t nR = 1 + g() % 3000, nS = 1 + g() % 5000;
    const int32_t range = 1 + (int32_t)(g() % 4000);
    const auto R = rand_tuples(g, nR, range), S = rand_tuples(g, nS, range);
    const Pairs want = brute(R, S);
    const otuple *r = R.data(), *s = S.data();
    CHECK(run_join([&](opair *o, uint64_t c) { return oracle_hash_join(r, nR, s, nS, 64, o, c); }, 16) == want);
    CHECK(run_join([&](opair *o, uint64_t c) { return oracle_hidx_join(r, nR, s, nS, 1024, o, c); }, 16) == want);
    CHECK(run_join([&](opair *o, uint64_t c) { return oracle_nlj_join(r, nR, s, nS, o, c); }, 16) == want);
    CHECK(run_join([&](opair *o, uint64_t c) { return oracle_sidx_join(r, nR, s, nS, o, c); }, 16) == want);
    CHECK(run_join([&](opair *o, uint64_t c) { return oracle_hash_join_mt(r, nR, s, nS, 64, 4, o, c); }, 16) ==
          want);
    CHECK(run_join([&](opair *o, uint64_t c) { return oracle_sidx_join_mt(r, nR, s, nS, 3, o, c); }, 16) == want);
    // count only
    CHECK(oracle_hash_join(r, nR, s, nS, 64, nullptr, 0) == want.size());
    // sorted index: sorted by key and a permutation of R
    std::vector<otuple> sorted(nR);
    CHECK(oracle_sidx_build(r, nR, sorted.data()) == 0);
    for (size_t i = 1; i < nR; ++i) CHECK(sorted[i - 1].key <= sorted[i].key);
    // stable partition: digit order, input order inside a digit
    const uint32_t parts = 1u + (uint32_t)(g() % 37);
    std::vector<otuple> out(nR);
    std::vector<uint64_t> start(parts + 1);
    CHECK(oracle_partition(r, nR, parts, out.data(), start.data()) == 0);
    CHECK(start[0] == 0 && start[parts] == nR);
    for (uint32_t p = 0; p < parts; ++p)
      for (uint64_t i = start[p]; i < start[p + 1]; ++i) {
        CHECK((uint32_t)out[i].key % parts == p);
        if (i > start[p]) CHECK(out[i - 1].payload < out[i].payload);
      }
    const int bits = 1 + (int)(g() % 8), shift = (int)(g() % 4);
    std::vector<uint64_t> start2((1u << bits) + 1);
    CHECK(oracle_partition_bits(r, nR, bits, shift, out.data(), start2.data()) == 0);
    for (uint32_t p = 0; p < (1u << bits); ++p)
      for (uint64_t i = start2[p]; i < start2[p + 1]; ++i)
        CHECK((((uint32_t)out[i].key >> shift) & ((1u << bits) - 1)) == p);
  }
  // scans against the sequential definition
  for (uint64_t n : {0ull, 1ull, 7ull, 1000ull, 65537ull}) {
    std::vector<uint32_t> a(n), b(n);
    std::vector<uint64_t> a64(n), b64(n);
    uint64_t run = 0;
    for (uint64_t i = 0; i < n; ++i) {
      a[i] = (uint32_t)(i * 2654435761u % 1000);
      a64[i] = a[i] * 1000003ull;
    }
    const uint64_t t = oracle_scan_u32(a.data(), b.data(), n);
    const uint64_t t64 = oracle_scan_u64(a64.data(), b64.data(), n);
    uint64_t run64 = 0;
    for (uint64_t i = 0; i < n; ++i) {
      CHECK(b[i] == (uint32_t)run && b64[i] == run64);
      run += a[i];
      run64 += a64[i];
    }
    CHECK(t == run && t64 == run64);
  }
  // distance join against a plain double loop
  std::vector<opoint3> P(300), Q(400);
  for (size_t i = 0; i < P.size(); ++i) P[i] = {(int32_t)i, 0, {double(g() % 20), double(g() % 20), double(g() % 20)}};
  for (size_t i = 0; i < Q.size(); ++i) Q[i] = {(int32_t)i, 0, {double(g() % 20), double(g() % 20), double(g() % 20)}};
  uint64_t want = 0;
  for (const auto &q : Q)
    for (const auto &p : P) {
      const double dx = q.v[0] - p.v[0], dy = q.v[1] - p.v[1], dz = q.v[2] - p.v[2];
      want += dx * dx + dy * dy + dz * dz < 9.0;
    }
  CHECK(oracle_nlj_dist_join(P.data(), P.size(), Q.data(), Q.size(), 3.0, nullptr, 0) == want);
}

static std::string tmp_path(const char *name) {
  const char *d = std::getenv("EO_CHECK_TMP");
  return std::string(d ? d : "/tmp") + "/" + name;
}

static void write_bytes(const std::string &path, const void *p, size_t n) {
  FILE *f = std::fopen(path.c_str(), "wb");
  CHECK(f != nullptr);
  if (n) CHECK(std::fwrite(p, 1, n, f) == n);
  std::fclose(f);
}

static void check_ref_files() {
  std::mt19937_64 g(7);
  std::vector<eo_tuple> t(1000);
  for (size_t i = 0; i < t.size(); ++i) t[i] = {(int32_t)(g() % 100), (int32_t)i};
  const std::string p = tmp_path("eo_table.bin");
  for (int flags : {0, (int)EO_REF_FILE_RAW}) {
    CHECK(eo_ref_table_write(p.c_str(), flags, t.data(), t.size()) == EO_OK);
    uint64_t n = 0;
    CHECK(eo_ref_table_read(p.c_str(), flags, nullptr, 0, &n) == EO_OK && n == t.size());
    std::vector<eo_tuple> back(n);
    CHECK(eo_ref_table_read(p.c_str(), flags, back.data(), n, &n) == EO_OK);
    for (size_t i = 0; i < n; ++i) CHECK(back[i].key == t[i].key && back[i].payload == t[i].payload);
    std::vector<eo_tuple> small(10);
    const eo_status st = eo_ref_table_read(p.c_str(), flags, small.data(), small.size(), &n);
    CHECK(st == EO_ECAPACITY && n == t.size());
  }
  // a count that promises more tuples than the file holds
  std::vector<int32_t> bad = {100, 1, 2, 3, 4};
  write_bytes(p, bad.data(), bad.size() * 4);
  uint64_t n = 0;
  CHECK(eo_ref_table_read(p.c_str(), 0, nullptr, 0, &n) != EO_OK);
  // negative count, a truncated tuple, an empty file
  bad = {-5};
  write_bytes(p, bad.data(), 4);
  CHECK(eo_ref_table_read(p.c_str(), 0, nullptr, 0, &n) != EO_OK);
  write_bytes(p, bad.data(), 3);
  CHECK(eo_ref_table_read(p.c_str(), 0, nullptr, 0, &n) != EO_OK);
  write_bytes(p, bad.data(), 0);
  CHECK(eo_ref_table_read(p.c_str(), 0, nullptr, 0, &n) != EO_OK);
  CHECK(eo_ref_table_read(p.c_str(), EO_REF_FILE_RAW, nullptr, 0, &n) == EO_OK && n == 0);
  CHECK(eo_ref_table_read(tmp_path("no_such_file").c_str(), 0, nullptr, 0, &n) == EO_EINVAL);
  // index files: sorted BUCKETs
  std::sort(t.begin(), t.end(), [](const eo_tuple &a, const eo_tuple &b) { return a.key < b.key; });
  CHECK(eo_ref_index_write(p.c_str(), t.data(), t.size()) == EO_OK);
  CHECK(eo_ref_index_read(p.c_str(), nullptr, 0, &n) == EO_OK && n == t.size());
  std::vector<eo_tuple> back(n);
  CHECK(eo_ref_index_read(p.c_str(), back.data(), n, &n) == EO_OK);
  for (size_t i = 0; i < n; ++i) CHECK(back[i].key == t[i].key && back[i].payload == t[i].payload);
  write_bytes(p, bad.data(), 5);  // not a whole number of BUCKETs
  CHECK(eo_ref_index_read(p.c_str(), nullptr, 0, &n) != EO_OK);
  std::remove(p.c_str());
}

static void check_tune_table() {
  eo_ctx *c = new eo_ctx();
  const std::string p = tmp_path("eo_tune.txt");
  auto write_text = [&](const char *s) { write_bytes(p, s, std::strlen(s)); };
  write_text("eojoin-tune 2\n20 21 7 1 16384 0 1.5\n27 28 14 2 8192 0 4.6\n");
  CHECK(eo_tune_load(c, p.c_str()) == EO_OK && c->tuned.size() == 2);
  eo_tune_result r{};
  CHECK(eo_tune_lookup(c, 1u << 27, 1u << 28, &r) == EO_OK && r.radix_bits == 14 && r.passes == 2 &&
        r.lds_tuples == 8192);
  CHECK(eo_tune_lookup(c, 1u << 10, 1u << 10, &r) == EO_EAGAIN);
  // apply_tuned fills the plan of a call that left it open, and never one that set it
  eo_join_opts o{};
  eo::apply_tuned(c, 1u << 27, 1u << 28, &o);
  CHECK(o.radix_bits == 14 && o.passes == 2 && o.lds_tuples == 8192);
  eo_join_opts o2{};
  o2.radix_bits = 9;
  eo::apply_tuned(c, 1u << 27, 1u << 28, &o2);
  CHECK(o2.radix_bits == 9 && o2.passes == 0);
  // save -> load round trip
  CHECK(eo_tune_save(c, p.c_str()) == EO_OK);
  eo_ctx *c2 = new eo_ctx();
  CHECK(eo_tune_load(c2, p.c_str()) == EO_OK && c2->tuned.size() == 2);
  // version 1 rows (no LDS column) still load
  write_text("eojoin-tune 1\n16 20 2 1 0 0.25\n");
  CHECK(eo_tune_load(c2, p.c_str()) == EO_OK);
  // malformed tables are refused and leave the table unchanged
  const size_t before = c2->tuned.size();
  const char *bad[] = {"", "garbage\n", "eojoin-tune 3\n", "eojoin-tune 2\n20 21 7 1 16384\n",
                       "eojoin-tune 2\n20 21 40 1 16384 0 1.0\n", "eojoin-tune 2\n-1 21 7 1 16384 0 1.0\n",
                       "eojoin-tune 2\n20 21 7 3 16384 0 1.0\n", "eojoin-tune 2\n20 21 7 1 1234 0 1.0\n",
                       "eojoin-tune 2\n20 99999999999999999999 7 1 16384 0 1.0\n"};
  for (const char *b : bad) {
    write_text(b);
    CHECK(eo_tune_load(c2, p.c_str()) == EO_EINVAL);
    CHECK(c2->tuned.size() == before);
  }
  CHECK(eo_tune_load(c2, tmp_path("no_such_file").c_str()) == EO_EINVAL);
  CHECK(eo_tune_clear(c2) == EO_OK && c2->tuned.empty());
  std::remove(p.c_str());
  delete c;
  delete c2;
}

static void check_planner() {
  for (int lr = 0; lr <= 31; ++lr)
    for (int ls = 0; ls <= 31; ls += 3) {
      const uint64_t nR = lr ? (1ull << lr) - 1 : 0, nS = ls ? (1ull << ls) + 1 : 0;
      eo::HashPlan hp;
      CHECK(eo::plan_hash_join(nR, nS, nullptr, &hp, 256) == EO_OK);
      CHECK(hp.bits >= 0 && hp.bits <= eo::kMaxJoinBits);
      CHECK(hp.passes == (hp.bits == 0 ? 0 : (hp.bits <= eo::kMaxPassBits ? 1 : 2)));
      if (hp.passes == 2) CHECK(hp.bits_pass1 + hp.bits_pass2 == hp.bits);
      CHECK(hp.lds_cap == 8192 || hp.lds_cap == 16384);
      CHECK(hp.s_span >= 8192 && hp.s_span <= (1u << 18) && (hp.s_span & (hp.s_span - 1)) == 0);
      // mean R partition within one LDS table
      if (hp.bits > 0 && hp.bits < eo::kMaxJoinBits) CHECK((nR >> hp.bits) <= hp.lds_cap);
      // unpartitioned: R in one table, or in a few tables beside a small S
      if (hp.bits == 0 && nR > hp.lds_cap) {
        const uint64_t chunks = (nR + eo::fused_table_tuples(hp.lds_cap) - 1) / eo::fused_table_tuples(hp.lds_cap);
        CHECK(chunks <= eo::kUnpartitionedChunks && chunks * nS <= eo::kUnpartitionedSReads);
      }
    }
  // overrides: accepted when consistent, refused otherwise
  eo_join_opts o{};
  eo::HashPlan hp;
  o.radix_bits = 12;
  o.passes = 1;
  CHECK(eo::plan_hash_join(1 << 20, 1 << 20, &o, &hp) == EO_EINVAL);  // 12 bits do not fit one pass
  o.passes = 2;
  CHECK(eo::plan_hash_join(1 << 20, 1 << 20, &o, &hp) == EO_OK && hp.bits_pass1 + hp.bits_pass2 == 12);
  o = eo_join_opts{};
  o.lds_tuples = 1000;
  CHECK(eo::plan_hash_join(1 << 20, 1 << 20, &o, &hp) == EO_EINVAL);
  o = eo_join_opts{};
  o.radix_bits = 23;
  CHECK(eo::plan_hash_join(1 << 20, 1 << 20, &o, &hp) == EO_EINVAL);
  o = eo_join_opts{};
  o.radix_bits = -1;
  CHECK(eo::plan_hash_join(1 << 20, 1 << 20, &o, &hp) == EO_OK && hp.bits == 0 && hp.passes == 0);
}

// One admission decision sequence of the request loop on a synthetic clock.
static void check_serve_queue() {
  using eo::ServeQueue;
  using eo::ServeQueueParams;
  std::vector<std::pair<uint32_t, eo_status>> gone;
  std::vector<uint32_t> idx;
  uint64_t bytes = 0;
  bool excl = false;
  ServeQueueParams p;
  p.budget = 1000;
  p.batch_cap = 500;
  p.max_batch = 3;
  {  // FIFO batches: count cap, byte cap, stop at the first request that does not fit
    const double arr[] = {0, 0, 0, 0, 0, 5};
    const uint64_t b[] = {100, 100, 100, 100, 450, 10};
    ServeQueue q(6, arr, b, nullptr, nullptr, p);
    CHECK(q.take(0.0, false, 0, &idx, &bytes, &excl) && idx == std::vector<uint32_t>({0, 1, 2}) && bytes == 300);
    CHECK(q.take(0.0, true, 300, &idx, &bytes, &excl) && idx == std::vector<uint32_t>({3}) && !excl);
    CHECK(!q.take(0.0, true, 800, &idx, &bytes, &excl));  // 450 would pass the budget with 800 in flight
    CHECK(q.take(0.0, false, 0, &idx, &bytes, &excl) && idx == std::vector<uint32_t>({4}));
    CHECK(!q.take(4.9, false, 0, &idx, &bytes, &excl));  // not arrived yet
    CHECK(q.next_arrival() == 5.0);
    CHECK(q.take(5.0, false, 0, &idx, &bytes, &excl) && idx == std::vector<uint32_t>({5}) && q.drained());
  }
  {  // a request above the batch share runs alone on an idle device; one above the budget is refused
    const uint64_t b[] = {600, 1001, 10};
    ServeQueue q(3, nullptr, b, nullptr, nullptr, p);
    CHECK(!q.take(0.0, true, 100, &idx, &bytes, &excl));
    CHECK(q.take(0.0, false, 0, &idx, &bytes, &excl) && idx == std::vector<uint32_t>({0}) && excl);
    CHECK(!q.take(0.0, false, 0, &idx, &bytes, &excl));  // head above the budget: never admitted
    gone.clear();
    q.retire(0.0, &gone);
    CHECK(gone.size() == 1 && gone[0].first == 1 && gone[0].second == EO_EBUDGET);
    CHECK(q.take(0.0, false, 0, &idx, &bytes, &excl) && idx == std::vector<uint32_t>({2}) && q.drained());
  }
  {  // deadlines: a waiting request past arrival + deadline is retired, an admitted one never
    const double arr[] = {0, 0, 1, 1};
    const uint64_t b[] = {100, 100, 100, 100};
    const double dl[] = {0.5, 0.5, 2.0, 0.0};
    ServeQueueParams p1 = p;
    p1.max_batch = 1;
    ServeQueue q(4, arr, b, dl, nullptr, p1);
    CHECK(q.take(0.1, false, 0, &idx, &bytes, &excl) && idx == std::vector<uint32_t>({0}));
    gone.clear();
    q.retire(0.5, &gone);  // request 1 is exactly at its deadline: still eligible
    CHECK(gone.empty());
    q.retire(0.6, &gone);
    CHECK(gone.size() == 1 && gone[0].first == 1 && gone[0].second == EO_ETIMEDOUT);
    CHECK(q.take(2.9, true, 100, &idx, &bytes, &excl) && idx == std::vector<uint32_t>({2}));
    gone.clear();
    q.retire(1000.0, &gone);  // no deadline: request 3 waits forever
    CHECK(gone.empty());
    CHECK(q.take(1000.0, true, 200, &idx, &bytes, &excl) && idx == std::vector<uint32_t>({3}) && q.drained());
  }
  {  // cancellation, also of a request behind the head and while the head waits
    const uint64_t b[] = {100, 100, 100, 100};
    uint32_t cancel[4] = {0, 1, 0, 0};
    ServeQueue q(4, nullptr, b, nullptr, cancel, p);
    CHECK(q.take(0.0, false, 0, &idx, &bytes, &excl) && idx == std::vector<uint32_t>({0}));  // stops at 1
    gone.clear();
    q.retire(0.0, &gone);
    CHECK(gone.size() == 1 && gone[0].first == 1 && gone[0].second == EO_ECANCELED);
    cancel[3] = 1;
    cancel[0] = 1;  // already admitted: no effect
    gone.clear();
    q.retire(0.0, &gone);
    CHECK(gone.size() == 1 && gone[0].first == 3);
    CHECK(q.take(0.0, false, 0, &idx, &bytes, &excl) && idx == std::vector<uint32_t>({2}) && q.drained());
  }
  {  // a request cancelled before it arrives leaves at once
    const double arr[] = {0, 10, 20};
    const uint64_t b[] = {100, 100, 100};
    uint32_t cancel[3] = {0, 0, 1};
    ServeQueue q(3, arr, b, nullptr, cancel, p);
    gone.clear();
    q.retire(1.0, &gone);
    CHECK(gone.size() == 1 && gone[0].first == 2 && gone[0].second == EO_ECANCELED);
    CHECK(q.take(1.0, false, 0, &idx, &bytes, &excl) && idx == std::vector<uint32_t>({0}));
    CHECK(!q.take(9.0, false, 0, &idx, &bytes, &excl) && q.next_arrival() == 10.0);
    CHECK(q.take(10.0, false, 0, &idx, &bytes, &excl) && idx == std::vector<uint32_t>({1}) && q.drained());
  }
  // randomized: every request leaves exactly once, admissions keep arrival
  // order, nothing admitted was cancelled or past its deadline at that time,
  // bytes in flight never pass the budget
  std::mt19937_64 g(99);
  for (int trial = 0; trial < 200; ++trial) {
    const uint32_t n = 1 + (uint32_t)(g() % 200);
    std::vector<double> arr(n), dl(n);
    std::vector<uint64_t> b(n);
    std::vector<uint32_t> cancel(n, 0);
    double t = 0;
    for (uint32_t i = 0; i < n; ++i) {
      t += (double)(g() % 100) / 100.0;
      arr[i] = t;
      b[i] = 1 + g() % 700;
      dl[i] = (g() % 4 == 0) ? (double)(g() % 300) / 100.0 : 0.0;
      cancel[i] = g() % 10 == 0;
    }
    ServeQueueParams pr;
    pr.budget = 1000;
    pr.batch_cap = 250 + g() % 500;
    pr.max_batch = 1 + (uint32_t)(g() % 8);
    ServeQueue q(n, arr.data(), b.data(), dl.data(), cancel.data(), pr);
    std::vector<int> seen(n, 0);
    std::vector<std::pair<double, uint64_t>> running;  // (finish time, bytes)
    double now = 0;
    int64_t last_admitted = -1;
    for (int step = 0; step < 100000 && !q.drained(); ++step) {
      now += 0.05;
      uint64_t in_flight = 0;
      std::vector<std::pair<double, uint64_t>> still;
      for (auto &r : running)
        if (r.first > now) {
          still.push_back(r);
          in_flight += r.second;
        }
      running.swap(still);
      gone.clear();
      q.retire(now, &gone);
      for (auto &x : gone) {
        CHECK(!seen[x.first]);
        seen[x.first] = 1;
        if (x.second == EO_ETIMEDOUT) CHECK(dl[x.first] > 0 && now > arr[x.first] + dl[x.first]);
        if (x.second == EO_ECANCELED) CHECK(cancel[x.first]);
        if (x.second != EO_ECANCELED) CHECK(arr[x.first] <= now);
        if (x.second == EO_EBUDGET) CHECK(b[x.first] > pr.budget);
      }
      while (q.take(now, !running.empty(), in_flight, &idx, &bytes, &excl)) {
        uint64_t sum = 0;
        for (uint32_t i : idx) {
          CHECK(!seen[i] && arr[i] <= now && !cancel[i]);
          CHECK(!(dl[i] > 0 && now > arr[i] + dl[i]));
          CHECK((int64_t)i > last_admitted);
          last_admitted = i;
          seen[i] = 1;
          sum += b[i];
        }
        CHECK(sum == bytes && idx.size() <= pr.max_batch);
        CHECK(excl ? (idx.size() == 1 && running.empty()) : sum <= pr.batch_cap);
        CHECK(in_flight + sum <= pr.budget);
        in_flight += sum;
        running.push_back({now + 0.01 * (double)(1 + g() % 50), sum});
      }
    }
    CHECK(q.drained());
    for (uint32_t i = 0; i < n; ++i) CHECK(seen[i]);
  }
}

int main() {
  check_oracle();
  check_ref_files();
  check_tune_table();
  check_planner();
  check_serve_queue();
  std::printf("host checks ok: %d checks\n", g_checks);
  return 0;
}
