// hss_sched.cpp -- admission and batch formation of the request server.
#include "hss_sched.h"

#include <algorithm>
#include <iterator>

namespace hss {

Sched::Sched(const SchedConfig& cfg) : cfg_(cfg) {
    if (cfg_.pool_positions > 0xFFFFFFFFull) cfg_.pool_positions = 0xFFFFFFFFull;
    if (cfg_.pool_positions) free_[0] = (uint32_t)cfg_.pool_positions;
    if (!cfg_.max_slots) cfg_.max_slots = 1;
    if (!cfg_.max_launch_positions) cfg_.max_launch_positions = 1;
}

bool Sched::enqueue(uint64_t id, uint32_t npos) {
    if (!npos || npos > cfg_.pool_positions) return false;
    queue_.emplace_back(id, npos);
    return true;
}

bool Sched::pool_alloc(uint32_t n, uint32_t& off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < n) continue;
        off = it->first;
        uint32_t rest = it->second - n;
        free_.erase(it);
        if (rest) free_[off + n] = rest;
        return true;
    }
    return false;
}

void Sched::pool_free_range(uint32_t off, uint32_t n) {
    auto it = free_.emplace(off, n).first;
    // merge with the following range
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        free_.erase(nx);
    }
    // merge with the preceding range
    if (it != free_.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_.erase(it);
        }
    }
}

uint64_t Sched::pool_free() const {
    uint64_t s = 0;
    for (const auto& kv : free_) s += kv.second;
    return s;
}

std::vector<uint64_t> Sched::admit() {
    std::vector<uint64_t> out;
    while (!queue_.empty() && active_.size() < cfg_.max_slots) {
        uint32_t off = 0;
        if (!pool_alloc(queue_.front().second, off)) break;  // FIFO: wait for memory
        Active a;
        a.npos = queue_.front().second;
        a.stop = a.npos;
        a.pool_off = off;
        active_.emplace(queue_.front().first, std::move(a));
        out.push_back(queue_.front().first);
        queue_.pop_front();
    }
    return out;
}

void Sched::set_costs(uint64_t id, std::vector<double> cost) {
    auto it = active_.find(id);
    if (it == active_.end()) return;
    cost.resize(it->second.npos, 1.0);
    it->second.left = 0;
    for (uint32_t i = it->second.cursor; i < it->second.stop; ++i) it->second.left += cost[i];
    it->second.cost = std::move(cost);
    it->second.costed = true;
}

bool Sched::has_work() const {
    for (const auto& kv : active_)
        if (kv.second.costed && kv.second.cursor < kv.second.stop) return true;
    return false;
}

int Sched::cancel(uint64_t id) {
    for (auto it = queue_.begin(); it != queue_.end(); ++it)
        if (it->first == id) {
            queue_.erase(it);
            return 1;
        }
    auto it = active_.find(id);
    if (it == active_.end()) return 0;
    Active& a = it->second;
    a.stop = a.cursor;
    a.left = 0;
    return a.done == a.stop ? 2 : 3;
}

std::vector<Segment> Sched::next_launch() {
    std::vector<std::pair<uint64_t, Active*>> live;
    for (auto& kv : active_)
        if (kv.second.costed && kv.second.cursor < kv.second.stop) live.emplace_back(kv.first, &kv.second);
    std::vector<Segment> segs;
    if (live.empty()) return segs;
    // a request that can finish within a quarter of the budget: a half-size launch
    bool short_left = false;
    for (const auto& [id, a] : live) short_left = short_left || a->left <= 0.25 * cfg_.launch_work;
    const double budget = short_left ? 0.5 * cfg_.launch_work : cfg_.launch_work;
    const double share = budget / (double)live.size();
    double total = 0;
    uint32_t positions = 0;
    std::vector<uint32_t> start(live.size());
    for (size_t i = 0; i < live.size(); ++i) start[i] = live[i].second->cursor;
    // pass 1: an equal share each, at least one position
    for (size_t i = 0; i < live.size() && positions < cfg_.max_launch_positions; ++i) {
        Active& a = *live[i].second;
        double acc = 0;
        while (a.cursor < a.stop && positions < cfg_.max_launch_positions) {
            const double w = a.cost[a.cursor];
            if (a.cursor > start[i] && acc + w > share) break;
            acc += w;
            a.left -= w;
            ++a.cursor;
            ++positions;
        }
        total += acc;
    }
    // pass 2: the rest of the budget, least work left first (admission order on ties)
    std::vector<size_t> order(live.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return live[x].second->left < live[y].second->left; });
    for (size_t i : order) {
        if (positions >= cfg_.max_launch_positions) break;
        Active& a = *live[i].second;
        while (a.cursor < a.stop && positions < cfg_.max_launch_positions) {
            const double w = a.cost[a.cursor];
            if (total + w > budget) break;
            total += w;
            a.left -= w;
            ++a.cursor;
            ++positions;
        }
    }
    for (size_t i = 0; i < live.size(); ++i)
        if (live[i].second->cursor > start[i]) segs.push_back(Segment{live[i].first, start[i], live[i].second->cursor});
    return segs;
}

std::vector<uint64_t> Sched::finished(const std::vector<Segment>& segs) {
    std::vector<uint64_t> out;
    for (const Segment& s : segs) {
        auto it = active_.find(s.id);
        if (it == active_.end()) continue;
        Active& a = it->second;
        a.done += s.end - s.begin;
        auto r = a.fin.emplace(s.begin, s.end).first;
        auto nx = std::next(r);
        if (nx != a.fin.end() && nx->first == r->second) {
            r->second = nx->second;
            a.fin.erase(nx);
        }
        if (r != a.fin.begin()) {
            auto pv = std::prev(r);
            if (pv->second == r->first) {
                pv->second = r->second;
                a.fin.erase(r);
            }
        }
        if (a.done == a.stop) out.push_back(s.id);
    }
    return out;
}

void Sched::retire(uint64_t id) {
    auto it = active_.find(id);
    if (it == active_.end()) return;
    pool_free_range(it->second.pool_off, it->second.npos);
    active_.erase(it);
}

uint32_t Sched::pool_offset(uint64_t id) const {
    auto it = active_.find(id);
    return it == active_.end() ? 0u : it->second.pool_off;
}

uint32_t Sched::prefix(uint64_t id) const {
    auto it = active_.find(id);
    if (it == active_.end() || it->second.fin.empty() || it->second.fin.begin()->first != 0) return 0;
    return it->second.fin.begin()->second;
}

uint32_t Sched::done(uint64_t id) const {
    auto it = active_.find(id);
    return it == active_.end() ? 0u : it->second.done;
}

}  // namespace hss
