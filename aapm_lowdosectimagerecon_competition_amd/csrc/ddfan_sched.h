// ddfan_sched.h -- admission policy of the reconstruction request loop.
//
// Pure host logic (no HIP): requests wait in arrival order and are admitted
// while a batch slot is free and their device footprint fits in what is left
// of the HBM budget.  A request larger than the whole budget can never run and
// is rejected at submission.  The device side (plans, batched launches) lives
// in ddfan_api.hip; this class only decides who runs.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <deque>
#include <vector>

namespace ddf {

class Admission {
public:
    Admission(size_t budget_bytes, int max_slots) : budget_(budget_bytes), slots_(max_slots) {}

    // false: the request can never fit (footprint > budget) and is not queued.
    bool submit(int64_t id, size_t bytes) {
        if (bytes > budget_) return false;
        queue_.push_back({id, bytes});
        return true;
    }

    // Admits waiting requests in arrival order while the head fits: a large
    // request at the head is not overtaken, so it cannot starve.
    std::vector<int64_t> admit() {
        std::vector<int64_t> out;
        while (!queue_.empty() && active_ < slots_ && used_ + queue_.front().bytes <= budget_) {
            used_ += queue_.front().bytes;
            ++active_;
            out.push_back(queue_.front().id);
            queue_.pop_front();
        }
        return out;
    }

    // Admits a new request at once, without queueing it, when no request is
    // waiting (so nothing is overtaken), a slot is free and it fits; false
    // leaves the state unchanged.
    bool admit_now(size_t bytes) {
        if (!queue_.empty() || active_ >= slots_ || bytes > budget_ || used_ + bytes > budget_)
            return false;
        used_ += bytes;
        ++active_;
        return true;
    }

    // An admitted request finished (or failed) and gave back its bytes.
    void release(size_t bytes) {
        used_ = bytes > used_ ? 0 : used_ - bytes;
        if (active_ > 0) --active_;
    }

    size_t used() const { return used_; }
    size_t budget() const { return budget_; }
    int active() const { return active_; }
    size_t waiting() const { return queue_.size(); }

private:
    struct Item {
        int64_t id;
        size_t bytes;
    };
    size_t budget_;
    int slots_;
    size_t used_ = 0;
    int active_ = 0;
    std::deque<Item> queue_;
};

}  // namespace ddf
