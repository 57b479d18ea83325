// hss_sched.h -- admission and batch formation of the request server (host
// only, no device code; unit-tested on the CPU).
//
// A request is an interval of root positions [first, last] at one LIMIT.  Its
// per-position counters live in a device pool of `pool_positions` counter
// slots; a request is admitted (FIFO) when a request slot is free and the pool
// has a contiguous free range for all its positions.  Each launch takes work
// from every admitted request that still has positions left: first an equal
// share of the launch's work budget per request (at least one position each),
// then the rest of the budget to the requests with the least work left.
// Requests therefore join the launch stream as soon as they are admitted and
// short requests are not held behind long ones.  While a request could finish
// within a quarter of the budget, launches take half the budget, so that it
// finishes after a short launch (launches of long requests alone stay long:
// fewer, larger launches for throughput).  Positions of one request are
// handed out in order; launches can finish out of order (a launch that is run
// again), so the finished prefix [first, resume) is tracked separately from
// the count.
//
// No counterpart in the reference, which runs one interval per process
// (scripts/gpu_queue.py:65-98 starts one gpu_search_v42 per GPU).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <deque>
#include <map>
#include <vector>

namespace hss {

struct SchedConfig {
    uint32_t max_slots = 256;          // admitted requests at once
    uint64_t pool_positions = 1 << 20;  // counter slots in the device pool
    double launch_work = 3e10;         // work budget of one launch (planner cells)
    uint32_t max_launch_positions = 1 << 16;
};

struct Segment {
    uint64_t id;
    uint32_t begin, end;  // position indices [begin, end) of the request
};

class Sched {
public:
    explicit Sched(const SchedConfig& cfg);

    // Queue a request of npos positions.  False if it can never be admitted
    // (npos == 0 or npos > pool_positions).
    bool enqueue(uint64_t id, uint32_t npos);
    // Admit queued requests in FIFO order while slots and pool space allow.
    std::vector<uint64_t> admit();
    // Per-position work of an admitted request (required before it is scheduled).
    void set_costs(uint64_t id, std::vector<double> cost);
    // Work of the next launch, as segments of admitted requests (empty: none).
    std::vector<Segment> next_launch();
    // A launch's segments have finished; returns the requests now complete.
    std::vector<uint64_t> finished(const std::vector<Segment>& segs);
    // Release a complete request's slot and pool range.
    void retire(uint64_t id);
    // Stop scheduling a request: a queued one leaves the queue (returns 1); an
    // admitted one keeps only the positions already handed to launches
    // (returns 2 when they have all finished, so it completes now, else 3);
    // 0 for an unknown id.
    int cancel(uint64_t id);

    // first counter slot of an admitted request
    uint32_t pool_offset(uint64_t id) const;
    // positions of a request finished so far, and the length of the finished prefix
    uint32_t done(uint64_t id) const;
    uint32_t prefix(uint64_t id) const;
    size_t queued() const { return queue_.size(); }
    size_t active() const { return active_.size(); }
    bool has_work() const;  // an admitted request has unscheduled positions
    uint64_t pool_free() const;

private:
    struct Active {
        uint32_t npos = 0, pool_off = 0;
        uint32_t stop = 0;    // positions to search (npos, or fewer after a cancel)
        uint32_t cursor = 0;  // positions handed to launches
        uint32_t done = 0;    // positions in finished launches
        bool costed = false;
        double left = 0;  // work of the positions not handed out yet
        std::vector<double> cost;
        std::map<uint32_t, uint32_t> fin;  // finished ranges, merged: begin -> end
    };
    bool pool_alloc(uint32_t n, uint32_t& off);
    void pool_free_range(uint32_t off, uint32_t n);

    SchedConfig cfg_;
    std::deque<std::pair<uint64_t, uint32_t>> queue_;  // (id, npos)
    std::map<uint64_t, Active> active_;                // ordered by id = admission order
    std::map<uint32_t, uint32_t> free_;                // pool free ranges: offset -> length
};

}  // namespace hss
