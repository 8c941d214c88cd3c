// ddfan_sched.h -- admission policy of the reconstruction request loop.
//
// Pure host logic (no HIP): requests wait and are admitted while a batch slot
// is free and their device footprint fits in what is left of the HBM budget,
// in arrival order (FIFO) or shortest first (SHORTEST: least work first, with
// an age limit).  A request larger than the whole budget can never run and is
// rejected at submission.  The device side (plans, batched launches) lives in
// ddfan_api.hip; this class only decides who runs.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <deque>
#include <vector>

namespace ddf {

class Admission {
public:
    enum Policy : int { FIFO = 0, SHORTEST = 1 };

    Admission(size_t budget_bytes, int max_slots) : budget_(budget_bytes), slots_(max_slots) {}

    // SHORTEST: among the waiting requests the one with the least work that
    // fits goes first (the earliest on ties), unless some request has waited
    // max_wait or longer: then the longest-waiting of those goes first and
    // nothing overtakes it until it fits, so none waits much beyond max_wait.
    void set_policy(int policy, double max_wait) {
        policy_ = policy == SHORTEST ? SHORTEST : FIFO;
        maxWait_ = max_wait;
    }
    int policy() const { return policy_; }

    // false: the request can never fit (footprint > budget) and is not queued.
    // work orders SHORTEST admission; now (any clock, seconds) ages it.
    bool submit(int64_t id, size_t bytes, double work = 0.0, double now = 0.0) {
        if (bytes > budget_) return false;
        queue_.push_back({id, bytes, work, now});
        return true;
    }

    // FIFO: admits waiting requests in arrival order while the head fits: a
    // large request at the head is not overtaken, so it cannot starve.
    // SHORTEST: as set_policy describes, at time now.
    std::vector<int64_t> admit(double now = 0.0) {
        std::vector<int64_t> out;
        auto take = [&](std::deque<Item>::iterator it) {
            used_ += it->bytes;
            ++active_;
            out.push_back(it->id);
            queue_.erase(it);
        };
        while (!queue_.empty() && active_ < slots_) {
            if (policy_ == FIFO) {
                if (used_ + queue_.front().bytes > budget_) break;
                take(queue_.begin());
                continue;
            }
            auto aged = queue_.end();  // the queue is in arrival order: the first is the oldest
            for (auto it = queue_.begin(); it != queue_.end(); ++it)
                if (now - it->t >= maxWait_) {
                    aged = it;
                    break;
                }
            if (aged != queue_.end()) {
                if (used_ + aged->bytes > budget_) break;
                take(aged);
                continue;
            }
            auto best = queue_.end();
            for (auto it = queue_.begin(); it != queue_.end(); ++it)
                if (used_ + it->bytes <= budget_ && (best == queue_.end() || it->work < best->work))
                    best = it;
            if (best == queue_.end()) break;
            take(best);
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
        double work, t;
    };
    size_t budget_;
    int slots_;
    int policy_ = FIFO;
    double maxWait_ = 0.0;
    size_t used_ = 0;
    int active_ = 0;
    std::deque<Item> queue_;
};

}  // namespace ddf
