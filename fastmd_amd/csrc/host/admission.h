// admission.h -- HBM-budget admission policy of the serving loop (pure C++).
// Jobs are admitted strictly in arrival order while their device footprint
// fits the bytes left in the budget and a resident slot is free; a job that
// does not fit waits (and blocks the jobs behind it) until running jobs finish.
#pragma once
#include <cstdint>
#include <deque>
#include <vector>

namespace fmd {

struct AdmitItem {
    int64_t job;
    uint64_t bytes;
};

class AdmissionQueue {
  public:
    AdmissionQueue(uint64_t budget, int max_resident) : budget_(budget), max_(max_resident) {}
    bool fits_ever(uint64_t bytes) const { return bytes <= budget_; }
    void push(const AdmitItem& it) { q_.push_back(it); }
    // A job handed back (it needs more than it was admitted with): it keeps its
    // place at the head of the queue.
    void push_front(const AdmitItem& it) { q_.push_front(it); }
    // An admitted job's charge changed by delta bytes (its real footprint).
    void adjust(int64_t delta) { used_ = (uint64_t)((int64_t)used_ + delta); }
    // Pops the jobs that can start now, in FIFO order.
    std::vector<AdmitItem> admit() {
        std::vector<AdmitItem> out;
        while (!q_.empty() && used_ + q_.front().bytes <= budget_ && resident_ < max_) {
            used_ += q_.front().bytes;
            resident_++;
            out.push_back(q_.front());
            q_.pop_front();
        }
        return out;
    }
    void release(uint64_t bytes) {
        used_ -= bytes;
        resident_--;
    }
    size_t queued() const { return q_.size(); }
    int resident() const { return resident_; }
    uint64_t used() const { return used_; }

  private:
    uint64_t budget_;
    int max_;
    uint64_t used_ = 0;
    int resident_ = 0;
    std::deque<AdmitItem> q_;
};

}  // namespace fmd
