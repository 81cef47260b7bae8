// The autotuner's table: size class (ceil log2 |R|, ceil log2 |S|) -> plan,
// its lookup by joins and batches, and its text file.  Host-only code (no HIP
// calls): the device-side timing that fills the table is in tuner.cpp.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "eojoin.h"
#include "plan.hpp"

namespace eo {

namespace {
int ceil_log2(uint64_t x) {
  int b = 0;
  while ((1ull << b) < x) ++b;
  return b;
}
}  // namespace

TuneKey tune_key(uint64_t nR, uint64_t nS) { return {ceil_log2(std::max<uint64_t>(nR, 1)), ceil_log2(std::max<uint64_t>(nS, 1))}; }

void apply_tuned(const eo_ctx *c, uint64_t nR, uint64_t nS, eo_join_opts *o) {
  if (o->radix_bits || o->passes || o->lds_tuples || o->s_tile) return;  // caller's explicit plan wins
  auto it = c->tuned.find(tune_key(nR, nS));
  if (it == c->tuned.end()) return;
  o->radix_bits = it->second.radix_bits ? it->second.radix_bits : -1;
  o->passes = it->second.passes;
  o->lds_tuples = it->second.lds_tuples;
  o->s_tile = it->second.s_tile;
}

}  // namespace eo

using namespace eo;

extern "C" {

eo_status eo_tune_lookup(eo_ctx *c, uint64_t nR, uint64_t nS, eo_tune_result *res) {
  if (!c || !res) return EO_EINVAL;
  std::memset(res, 0, sizeof(*res));
  auto it = c->tuned.find(tune_key(nR, nS));
  if (it == c->tuned.end()) return EO_EAGAIN;
  res->radix_bits = it->second.radix_bits;
  res->passes = it->second.passes;
  res->lds_tuples = it->second.lds_tuples;
  res->s_tile = it->second.s_tile;
  res->ms = it->second.ms;
  return EO_OK;
}

eo_status eo_tune_clear(eo_ctx *c) {
  if (!c) return EO_EINVAL;
  c->tuned.clear();
  return EO_OK;
}

// Text format, one size class per line:
//   eojoin-tune 1
//   <log2 |R|> <log2 |S|> <radix_bits> <passes> <s_tile> <ms>
eo_status eo_tune_save(eo_ctx *c, const char *path) {
  if (!c || !path) return EO_EINVAL;
  FILE *f = std::fopen(path, "w");
  if (!f) {
    set_last_error(std::string("cannot write ") + path);
    return EO_EINVAL;
  }
  std::fprintf(f, "eojoin-tune 2\n");  // rows: log2|R| log2|S| bits passes lds_tuples s_tile ms
  for (const auto &kv : c->tuned)
    std::fprintf(f, "%d %d %d %d %d %d %.6f\n", kv.first.first, kv.first.second, kv.second.radix_bits,
                 kv.second.passes, kv.second.lds_tuples, kv.second.s_tile, kv.second.ms);
  const bool ok = std::fclose(f) == 0;
  return ok ? EO_OK : EO_EINVAL;
}

eo_status eo_tune_load(eo_ctx *c, const char *path) {
  if (!c || !path) return EO_EINVAL;
  FILE *f = std::fopen(path, "r");
  if (!f) {
    set_last_error(std::string("cannot read ") + path);
    return EO_EINVAL;
  }
  int ver = 0;
  if (std::fscanf(f, "eojoin-tune %d", &ver) != 1 || (ver != 1 && ver != 2)) {
    std::fclose(f);
    set_last_error("not an eojoin tuning table");
    return EO_EINVAL;
  }
  std::vector<std::pair<TuneKey, TunedPlan>> rows;
  int kr, ks, b, p, lds = 0, st;
  double ms;
  int got;
  const int fields = ver == 1 ? 6 : 7;  // version 1 rows have no lds_tuples column
  while ((got = ver == 1 ? std::fscanf(f, "%d %d %d %d %d %lf", &kr, &ks, &b, &p, &st, &ms)
                         : std::fscanf(f, "%d %d %d %d %d %d %lf", &kr, &ks, &b, &p, &lds, &st, &ms)) == fields) {
    eo_join_opts o{};
    o.radix_bits = b ? b : -1;
    o.passes = p;
    o.lds_tuples = lds;
    o.s_tile = st;
    HashPlan hp;
    if (kr < 0 || kr > 31 || ks < 0 || ks > 31 || st < 0 ||
        plan_hash_join(1ull << kr, 1ull << ks, &o, &hp) != EO_OK) {
      std::fclose(f);
      set_last_error("invalid tuning table row");
      return EO_EINVAL;
    }
    rows.push_back({{kr, ks}, {b, p, lds, st, ms}});
  }
  const bool clean_end = got == EOF;
  std::fclose(f);
  if (!clean_end) {
    set_last_error("malformed tuning table");
    return EO_EINVAL;
  }
  for (auto &r : rows) c->tuned[r.first] = r.second;
  return EO_OK;
}

}  // extern "C"
