// ba_ledger.hpp — bookkeeping of the event pairs that time one kernel's launches.
//
// A solve that profiles a kernel brackets each of its launches with a pair of events.  The pairs
// are created on demand and reused from solve to solve; `count()` of them hold launches that
// count.  PCG launches queued after the PCG state turned done return at once, so the solver keeps
// only as many of a batch's pairs as the device reports products applied (`keep`): that count
// comes from device memory, so it is clamped to the pairs that exist, and a pair is never indexed
// beyond the vector.  Ops supplies create(Event&) and destroy(Event); it is the HIP runtime in the
// library and a plain counter in the host test of this class.
#pragma once

#include <cstddef>
#include <utility>
#include <vector>

namespace ba {

template <class Event, class Ops>
class LaunchLedger {
   public:
    explicit LaunchLedger(Ops ops = Ops()) : ops_(ops) {}
    LaunchLedger(const LaunchLedger&) = delete;
    LaunchLedger& operator=(const LaunchLedger&) = delete;
    ~LaunchLedger() {
        for (auto& e : pairs_) {
            ops_.destroy(e.first);
            ops_.destroy(e.second);
        }
    }
    // The pair for the next launch (created when every existing pair is in use).
    std::pair<Event, Event>& next() {
        if (used_ >= pairs_.size()) {
            std::pair<Event, Event> e{};
            ops_.create(e.first);
            ops_.create(e.second);
            pairs_.push_back(e);
            used_ = pairs_.size() - 1;
        }
        return pairs_[used_++];
    }
    // Launches recorded so far.
    size_t count() const { return used_; }
    // Keep the first `n` recorded launches (clamped to the pairs that exist); later ones are reused.
    void keep(size_t n) { used_ = n < pairs_.size() ? n : pairs_.size(); }
    const std::pair<Event, Event>& operator[](size_t k) const { return pairs_[k]; }
    size_t capacity() const { return pairs_.size(); }

   private:
    Ops ops_;
    std::vector<std::pair<Event, Event>> pairs_;
    size_t used_ = 0;
};

}  // namespace ba
