// Admission policy of the request loop (eo_serve_run): which waiting requests
// form the next batch, and which leave the queue without running.  Host-only
// logic over plain arrays (no HIP calls), so it is tested on the CPU with
// synthetic clocks.
//
// The reference has no request loop: every join is a separate process run
// (joins/hashJoin_gpu/main_street/hj.cc:1189-1269) and its only failure
// handling is printf + exit(1) (hj.cc:212-227).
#pragma once

#include <stdint.h>

#include <utility>
#include <vector>

#include "eojoin.h"

namespace eo {

struct ServeQueueParams {
  uint64_t budget = 0;     // bytes admitted at once (a larger request is refused)
  uint64_t batch_cap = 0;  // bytes one batch may hold (a larger request runs alone on an idle device)
  uint32_t max_batch = 1;  // requests per batch
};

// Requests are kept in arrival order (arrival_ms non-decreasing).  A request
// has arrived at time `now` when arrival_ms[i] <= now; it then waits until it
// is admitted into a batch or retired:
//   - EO_ECANCELED: its cancel flag is non-zero (read at every decision; a
//     request cancelled before its arrival is retired at once);
//   - EO_ETIMEDOUT: deadline_ms[i] > 0 and now > arrival_ms[i] + deadline_ms[i];
//   - EO_EBUDGET: its bytes exceed the whole budget.
// Admission is FIFO: a batch takes the waiting requests in order while they
// fit (count, batch bytes, bytes in flight) and stops at the first that does
// not, so no request is overtaken by a later one.  A request above batch_cap
// runs alone, once the device is idle.  Admitted requests always run to
// completion (a kernel in flight is never interrupted).
class ServeQueue {
 public:
  // arrival_ms, deadline_ms, cancel: NULL means all 0 (all arrived at once,
  // no deadlines, nothing cancelled).  The arrays must outlive the queue.
  ServeQueue(uint32_t n, const double *arrival_ms, const uint64_t *bytes, const double *deadline_ms,
             const volatile uint32_t *cancel, const ServeQueueParams &p);

  // Retires every arrived, waiting request that is cancelled, past its
  // deadline or above the budget; appends (request, status) to *done.
  void retire(double now, std::vector<std::pair<uint32_t, eo_status>> *done);

  // Forms the next batch at time `now` (device_busy: a batch is in flight;
  // in_flight: bytes admitted and not yet completed).  Returns false when no
  // request is admitted now; otherwise *idx holds the batch in arrival order,
  // *bytes its admitted bytes, *exclusive whether it runs alone.
  bool take(double now, bool device_busy, uint64_t in_flight, std::vector<uint32_t> *idx, uint64_t *bytes,
            bool *exclusive);

  // True when every request has been admitted or retired.
  bool drained() const { return next_ >= n_; }
  // Arrival time of the first request not yet admitted or retired (the loop
  // sleeps until then when nothing else is pending); +inf when drained.
  double next_arrival() const;

  double arrival(uint32_t i) const { return arrival_ ? arrival_[i] : 0.0; }

 private:
  bool cancelled(uint32_t i) const;
  bool expired(uint32_t i, double now) const;
  void advance();  // moves next_ past handled requests

  uint32_t n_;
  const double *arrival_;
  const uint64_t *bytes_;
  const double *deadline_;
  const volatile uint32_t *cancel_;
  ServeQueueParams p_;
  uint32_t next_ = 0;             // first request neither admitted nor retired
  std::vector<uint8_t> handled_;  // admitted or retired
};

}  // namespace eo
