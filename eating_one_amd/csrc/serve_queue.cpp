#include "serve_queue.hpp"

#include <limits>

namespace eo {

ServeQueue::ServeQueue(uint32_t n, const double *arrival_ms, const uint64_t *bytes, const double *deadline_ms,
                       const volatile uint32_t *cancel, const ServeQueueParams &p)
    : n_(n), arrival_(arrival_ms), bytes_(bytes), deadline_(deadline_ms), cancel_(cancel), p_(p), handled_(n, 0) {
  if (p_.max_batch == 0) p_.max_batch = 1;
}

bool ServeQueue::cancelled(uint32_t i) const {
  // the flags may be set by another host thread while the loop runs
  return cancel_ && __atomic_load_n(const_cast<const uint32_t *>(&cancel_[i]), __ATOMIC_ACQUIRE) != 0;
}

bool ServeQueue::expired(uint32_t i, double now) const {
  return deadline_ && deadline_[i] > 0.0 && now > arrival(i) + deadline_[i];
}

void ServeQueue::advance() {
  while (next_ < n_ && handled_[next_]) ++next_;
}

double ServeQueue::next_arrival() const {
  return next_ < n_ ? arrival(next_) : std::numeric_limits<double>::infinity();
}

void ServeQueue::retire(double now, std::vector<std::pair<uint32_t, eo_status>> *done) {
  // a cancelled request leaves the queue even before it arrives
  for (uint32_t i = next_; i < n_ && (cancel_ || arrival(i) <= now); ++i) {
    if (handled_[i]) continue;
    eo_status st = EO_OK;
    if (cancelled(i)) st = EO_ECANCELED;
    else if (arrival(i) > now) continue;
    else if (expired(i, now)) st = EO_ETIMEDOUT;
    else if (bytes_[i] > p_.budget) st = EO_EBUDGET;
    if (st == EO_OK) continue;
    handled_[i] = 1;
    done->push_back({i, st});
  }
  advance();
}

bool ServeQueue::take(double now, bool device_busy, uint64_t in_flight, std::vector<uint32_t> *idx, uint64_t *bytes,
                      bool *exclusive) {
  idx->clear();
  *bytes = 0;
  *exclusive = false;
  advance();
  if (next_ >= n_ || arrival(next_) > now) return false;
  const uint32_t head = next_;
  // a request that must leave the queue is never admitted (retire() reports it)
  if (cancelled(head) || expired(head, now) || bytes_[head] > p_.budget) return false;
  if (bytes_[head] > p_.batch_cap) {
    if (device_busy) return false;  // runs alone: wait until the device is idle
    idx->push_back(head);
    *bytes = bytes_[head];
    *exclusive = true;
  } else {
    for (uint32_t i = head; i < n_ && arrival(i) <= now && idx->size() < p_.max_batch; ++i) {
      if (handled_[i]) continue;
      if (cancelled(i) || expired(i, now) || bytes_[i] > p_.budget) break;  // retired before it can run
      const uint64_t b = bytes_[i];
      if (b > p_.batch_cap || *bytes + b > p_.batch_cap || in_flight + *bytes + b > p_.budget) break;
      idx->push_back(i);
      *bytes += b;
    }
  }
  for (uint32_t i : *idx) handled_[i] = 1;
  advance();
  return !idx->empty();
}

}  // namespace eo
