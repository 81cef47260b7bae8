// Resident build table: a relation partitioned once and kept in HBM.
#pragma once

#include <stdint.h>

#include "eojoin.h"
#include "plan.hpp"

struct eo_table {
  eo_ctx *ctx = nullptr;
  uint64_t n = 0;
  eo::HashPlan hp;
  eo_tuple *part = nullptr;   // partitioned tuples (R itself copied when bits == 0)
  uint32_t *start = nullptr;  // 2^bits + 1 partition starts
  void *mem = nullptr;
  uint64_t bytes = 0;
};
