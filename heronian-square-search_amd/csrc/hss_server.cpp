// hss_server.cpp -- the served form: a request loop over one context.
//
// Clients submit interval requests (LIMIT, FIRST_L, LAST_L) from any thread.
// One worker thread owns the context and its stream:
//   * admission (hss_sched.h): FIFO, bounded by request slots and by the
//     device pool of per-position counters (24 bytes per root position, the
//     per-request state that stays on the device until the request is done);
//   * continuous batching: every launch takes positions from all admitted
//     requests (an equal work share each, then the rest in arrival order), so
//     a request admitted while others run joins the next launch, and the
//     tiles of different requests (and LIMITs) interleave in the persistent
//     grid's work queue;
//   * two launch buffer sets: while one launch runs, the next is planned on
//     the host and queued behind it, so the device never waits for the host;
//     it is planned late, shortly before the running launch is predicted to
//     end (plan_delay), so that requests arriving meanwhile join it;
//   * results: counters accumulate in the pool; when a request's last
//     positions finish, its range is read back on a second stream (without
//     waiting for the launch that is running) and the request completes.
// Each position's counts come from the same kernels and arithmetic as a
// search of that position alone, so results do not depend on the mix.
//
// No counterpart in the reference: its supervisor (scripts/gpu_queue.py:65-98)
// starts one gpu_search_v42 process per interval and GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "hss_ctx.h"
#include "hss_sched.h"

using namespace hss;

namespace {

struct ReqRec {
    uint64_t limit = 0, first = 0, last = 0;
    uint32_t npos = 0;
    bool canceled = false;
    int32_t state = HSS_REQ_QUEUED;
    int32_t error = HSS_OK;
    uint32_t launches = 0;
    uint64_t resume = 0;
    double t_submit = -1, t_admit = -1, t_start = -1, t_done = -1;
    std::vector<hss_root_result> per;  // per position, filled as its launch finishes
    std::vector<HitOut> hits;          // accepted hits found so far
    uint64_t cand = 0, leaves = 0, pairs = 0;  // totals once finished
};

struct Launch {
    int buf = 0;
    std::vector<Segment> segs;
    std::vector<uint64_t> owner;  // plan position -> request id
    int attempts = 0;
    double work = 0;      // planner work estimate of the launch
    double t_start = -1;  // server clock when the device reached it (-1: not seen yet)
};

// Launch timing the worker learns as it goes: device milliseconds per unit of
// planner work, and host milliseconds to plan and enqueue a launch.
struct LaunchClock {
    double ms_per_work = 0;  // 0 until a launch has finished
    double plan_ms = 2.0;
    void finished(double work, double device_ms) {
        if (work <= 0 || device_ms <= 0) return;
        const double r = device_ms / work;
        ms_per_work = ms_per_work > 0 ? 0.7 * ms_per_work + 0.3 * r : r;
    }
    void planned(double host_ms) { plan_ms = 0.7 * plan_ms + 0.3 * host_ms; }
};

}  // namespace

struct hss_server {
    hss_ctx* ctx = nullptr;
    hipStream_t copy = nullptr;
    DevBuf pool;
    Counters cnt;
    uint64_t pool_positions = 0;
    LaunchBuffers bufs[2];
    HostBuf staging;  // pinned: counters of a finished launch
    Sched sched;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::map<uint64_t, ReqRec> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    int failed = HSS_OK;
    std::string fail_msg;
    std::thread worker;
    std::chrono::steady_clock::time_point t0;
    hss_server_stats st{};

    explicit hss_server(const SchedConfig& sc) : sched(sc) {}
    double now() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

// Mark every unfinished request failed (device error); the worker then exits.
void fail_all(hss_server* S, int code, const std::string& msg) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->failed = code;
    S->fail_msg = msg;
    const double t = S->now();
    for (auto& kv : S->reqs) {
        ReqRec& r = kv.second;
        if (r.state == HSS_REQ_DONE || r.state == HSS_REQ_FAILED) continue;
        r.state = HSS_REQ_FAILED;
        r.error = code;
        r.t_done = t;
        ++S->st.failed;
    }
    S->cv_done.notify_all();
}

// Read the counters of a finished launch's positions out of the pool (on the
// copy stream: the launch running meanwhile writes other positions).
int read_back(hss_server* S, const std::vector<Segment>& segs) {
    std::vector<std::pair<uint32_t, uint32_t>> src;  // (pool offset, count) per segment
    size_t total = 0;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        for (const Segment& sg : segs) {
            src.emplace_back(S->sched.pool_offset(sg.id) + sg.begin, sg.end - sg.begin);
            total += sg.end - sg.begin;
        }
    }
    HSS_HIP_TRY(S->staging.reserve(total * 24));
    unsigned long long* h = S->staging.as<unsigned long long>();
    size_t at = 0;
    for (const auto& [off, n] : src) {
        HSS_HIP_TRY(hipMemcpyAsync(h + at, S->cnt.cand + off, n * 8, hipMemcpyDeviceToHost, S->copy));
        HSS_HIP_TRY(hipMemcpyAsync(h + total + at, S->cnt.leaves + off, n * 8, hipMemcpyDeviceToHost, S->copy));
        HSS_HIP_TRY(hipMemcpyAsync(h + 2 * total + at, S->cnt.pairs + off, n * 8, hipMemcpyDeviceToHost, S->copy));
        at += n;
    }
    HSS_HIP_TRY(hipStreamSynchronize(S->copy));
    std::lock_guard<std::mutex> lk(S->mu);
    at = 0;
    for (const Segment& sg : segs) {
        ReqRec& r = S->reqs[sg.id];
        for (uint32_t i = sg.begin; i < sg.end; ++i, ++at)
            r.per[i] = hss_root_result{h[at], h[total + at], h[2 * total + at]};
    }
    return HSS_OK;
}

// A request whose last scheduled positions have finished: totals, sorted
// unique hits, release of its slot and pool range.  Caller holds S->mu.
void finalize_locked(hss_server* S, uint64_t id) {
    ReqRec& r = S->reqs[id];
    const uint32_t done = S->sched.done(id);
    r.cand = r.leaves = r.pairs = 0;
    for (uint32_t i = 0; i < r.npos; ++i) {
        r.cand += r.per[i].candidates;
        r.leaves += r.per[i].leaves;
        r.pairs += r.per[i].pairs;
    }
    std::set<std::tuple<uint64_t, uint64_t, uint64_t>> seen;
    std::vector<HitOut> uniq;
    for (const HitOut& x : r.hits)
        if (seen.insert({x.l, x.m, x.n}).second) uniq.push_back(x);
    std::sort(uniq.begin(), uniq.end(), [](const HitOut& a, const HitOut& b) {
        return std::tie(a.l, a.m, a.n) < std::tie(b.l, b.m, b.n);
    });
    r.hits.swap(uniq);
    r.resume = r.first + S->sched.prefix(id);
    r.state = r.canceled ? HSS_REQ_CANCELED : HSS_REQ_DONE;
    r.t_done = S->now();
    S->sched.retire(id);
    ++S->st.completed;
    S->st.positions += done;
    S->st.leaves += r.leaves;
    S->st.candidates += r.cand;
    S->cv_done.notify_all();
    S->cv_work.notify_all();
}

// Late planning.  With one launch on the device and the other buffer set
// free, the next launch is planned only shortly before the running one is
// predicted to end (its start, seen through its first event, plus the learned
// milliseconds per unit of work, minus a lead that covers planning and
// enqueueing), so requests that arrive meanwhile still join it: a request
// then waits for about the rest of the running launch and its own, not for a
// launch already queued.  Returns the milliseconds to wait before planning
// (<= 0: plan now).  A launch that has finished, or a clock that has no
// estimate yet, plans at once, so the device never waits for the host longer
// than one planning.
double plan_delay(hss_server* S, std::deque<Launch>& inflight, const LaunchClock& clk) {
    if (clk.ms_per_work <= 0 || inflight.size() != 1) return 0;
    Launch& L = inflight.front();
    const LaunchBuffers& B = S->bufs[L.buf];
    if (hipEventQuery(B.ev[5]) == hipSuccess) return 0;
    if (L.t_start < 0) {
        if (hipEventQuery(B.ev[0]) != hipSuccess) return 0.25;  // queued behind copies: look again shortly
        L.t_start = S->now();
    }
    const double lead = 2.0 * clk.plan_ms + 1.0;
    return L.t_start + clk.ms_per_work * L.work - lead - S->now();
}

void worker_main(hss_server* S) {
    hss_ctx* c = S->ctx;
    if (hipSetDevice(c->device) != hipSuccess) {
        fail_all(S, HSS_EDEVICE, "hipSetDevice failed in the server thread");
        return;
    }
    std::deque<Launch> inflight;
    LaunchClock clk;
    unsigned free_mask = 3;
    for (;;) {
        std::vector<uint64_t> admitted;
        std::vector<std::tuple<uint64_t, uint64_t, uint32_t>> todo;  // (limit, first, npos)
        std::vector<Segment> segs;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            for (;;) {
                admitted = S->sched.admit();
                if (!admitted.empty() || !inflight.empty() || (free_mask && S->sched.has_work())) break;
                if (S->stop && !S->sched.queued() && !S->sched.active()) return;
                S->cv_work.wait(lk);
            }
            const double t = S->now();
            for (uint64_t id : admitted) {
                ReqRec& r = S->reqs[id];
                r.state = HSS_REQ_ACTIVE;
                r.t_admit = t;
                r.per.assign(r.npos, hss_root_result{0, 0, 0});
                todo.emplace_back(r.limit, r.first, r.npos);
            }
            S->st.max_active = std::max<uint32_t>(S->st.max_active, (uint32_t)S->sched.active());
        }
        // per-position work estimates (host arithmetic, outside the lock)
        std::vector<std::vector<double>> costs(admitted.size());
        for (size_t a = 0; a < admitted.size(); ++a) {
            const auto& [limit, first, npos] = todo[a];
            costs[a].resize(npos);
            for (uint32_t i = 0; i < npos; ++i) {
                double e = estimate_position(first + i, limit, c->small_primes, c->h_off.data(), c->nmax);
                costs[a][i] = e > 0 ? e : 1.0;
            }
        }
        if (free_mask) {
            std::unique_lock<std::mutex> lk(S->mu);
            for (size_t a = 0; a < admitted.size(); ++a) S->sched.set_costs(admitted[a], std::move(costs[a]));
            if (S->sched.has_work() && !S->stop) {
                const double wait_ms = plan_delay(S, inflight, clk);
                if (wait_ms > 0) {
                    // wake for submissions; look at the running launch at least every millisecond
                    S->cv_work.wait_for(lk, std::chrono::duration<double, std::milli>(std::min(wait_ms, 1.0)));
                    continue;
                }
            }
            segs = S->sched.next_launch();
        } else if (!admitted.empty()) {
            std::lock_guard<std::mutex> lk(S->mu);
            for (size_t a = 0; a < admitted.size(); ++a) S->sched.set_costs(admitted[a], std::move(costs[a]));
        }
        if (!segs.empty()) {
            Launch L;
            L.buf = (free_mask & 1u) ? 0 : 1;
            L.segs = segs;
            std::vector<PosReq> pr;
            {
                std::lock_guard<std::mutex> lk(S->mu);
                const double t = S->now();
                for (const Segment& sg : segs) {
                    ReqRec& r = S->reqs[sg.id];
                    const uint32_t off = S->sched.pool_offset(sg.id);
                    for (uint32_t i = sg.begin; i < sg.end; ++i) {
                        pr.push_back(PosReq{r.first + i, r.limit, off + i});
                        L.owner.push_back(sg.id);
                    }
                    ++r.launches;
                    if (r.t_start < 0) r.t_start = t;
                }
                if (segs.size() > 1) ++S->st.shared_launches;
            }
            const double t_plan = S->now();
            int rc = launch_plan(c, S->bufs[L.buf], pr.data(), (uint32_t)pr.size());
            if (rc == HSS_OK) rc = launch_enqueue(c, S->bufs[L.buf], S->cnt, c->stream);
            L.work = S->bufs[L.buf].plan.est_work;
            clk.planned(S->now() - t_plan);
            if (rc != HSS_OK) {
                fail_all(S, rc, hss_last_error());
                (void)hipStreamSynchronize(c->stream);
                return;
            }
            free_mask &= ~(1u << L.buf);
            inflight.push_back(std::move(L));
            continue;  // a second launch may be queued behind this one
        }
        if (inflight.empty()) continue;
        // complete the oldest launch
        Launch L = std::move(inflight.front());
        inflight.pop_front();
        LaunchBuffers& B = S->bufs[L.buf];
        LaunchResult R;
        int rc = hipEventSynchronize(B.ev[5]) == hipSuccess ? HSS_OK : fail(HSS_EDEVICE, "launch failed");
        if (rc == HSS_OK) rc = launch_collect(c, B, S->copy, R);
        if (rc == HSS_OK && R.rerun) {
            if (++L.attempts > 3) {
                rc = fail(HSS_ENOMEM, "deferred-leaf buffer overflow");
            } else {
                {
                    std::lock_guard<std::mutex> lk(S->mu);
                    ++S->st.reruns;
                }
                rc = launch_enqueue(c, B, S->cnt, c->stream);
                if (rc == HSS_OK) {
                    L.t_start = -1;
                    inflight.push_back(std::move(L));
                    continue;
                }
            }
        }
        if (rc != HSS_OK) {
            fail_all(S, rc, hss_last_error());
            (void)hipStreamSynchronize(c->stream);
            return;
        }
        free_mask |= 1u << L.buf;
        clk.finished(L.work, (double)R.setup_ms + R.cells_ms + R.exact_ms);
        rc = read_back(S, L.segs);
        if (rc != HSS_OK) {
            fail_all(S, rc, hss_last_error());
            (void)hipStreamSynchronize(c->stream);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(S->mu);
            ++S->st.launches;
            S->st.gpu_ms += (double)R.setup_ms + R.cells_ms + R.exact_ms;
            for (const HitOut& h : R.hits)
                if (h.pos < L.owner.size()) S->reqs[L.owner[h.pos]].hits.push_back(h);
            const std::vector<uint64_t> done = S->sched.finished(L.segs);
            for (const Segment& sg : L.segs) {
                ReqRec& r = S->reqs[sg.id];
                r.resume = r.first + S->sched.prefix(sg.id);
            }
            for (uint64_t id : done) finalize_locked(S, id);
        }
    }
}

}  // namespace

extern "C" int hss_server_create(hss_ctx* ctx, const hss_server_config* cfg, hss_server** out) {
    if (!ctx || !out) return fail(HSS_EINVAL, "null argument");
    *out = nullptr;
    if (ctx->server) return fail(HSS_EINVAL, "context already has a server");
    hss_server_config k{};
    if (cfg) k = *cfg;
    SchedConfig sc;
    sc.max_slots = k.max_slots ? k.max_slots : 256;
    const uint64_t bytes = k.state_bytes ? k.state_bytes : (64ull << 20);
    sc.pool_positions = std::min<uint64_t>(bytes / 24, 0xFFFFFFFFull);
    if (!sc.pool_positions) return fail(HSS_EINVAL, "state_bytes holds no position (24 bytes each)");
    sc.launch_work = k.launch_work > 0 ? k.launch_work : 3e10;
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(HSS_EDEVICE, "hipSetDevice failed");
    hss_server* S = new hss_server(sc);
    S->ctx = ctx;
    S->pool_positions = sc.pool_positions;
    auto bail = [&](int code, const char* what, hipError_t e) {
        std::string msg = std::string(what) + ": " + hipGetErrorString(e);
        if (S->copy) (void)hipStreamDestroy(S->copy);
        S->pool.release();
        for (LaunchBuffers& b : S->bufs) b.release();
        S->staging.release();
        delete S;
        return fail(code, msg);
    };
    hipError_t e = hipStreamCreateWithFlags(&S->copy, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(HSS_EDEVICE, "copy stream", e);
    e = S->pool.reserve(sc.pool_positions * 24);
    if (e != hipSuccess) return bail(e == hipErrorOutOfMemory ? HSS_ENOMEM : HSS_EDEVICE, "state pool", e);
    S->cnt.cand = S->pool.as<unsigned long long>();
    S->cnt.leaves = S->cnt.cand + sc.pool_positions;
    S->cnt.pairs = S->cnt.cand + 2 * sc.pool_positions;
    for (LaunchBuffers& b : S->bufs) {
        e = b.create();
        if (e != hipSuccess) return bail(HSS_EDEVICE, "launch buffers", e);
    }
    S->t0 = std::chrono::steady_clock::now();
    S->st.pool_positions = (uint32_t)sc.pool_positions;
    ctx->server = S;
    S->worker = std::thread(worker_main, S);
    *out = S;
    return HSS_OK;
}

extern "C" int hss_server_submit(hss_server* S, uint64_t limit, uint64_t first, uint64_t last, uint64_t* id) {
    if (!S || !id) return fail(HSS_EINVAL, "null argument");
    *id = 0;
    if (limit < 1 || limit > S->ctx->nmax) return fail(HSS_EINVAL, "limit outside [1, n_max]");
    if (first < 1 || first > last || last > limit) return fail(HSS_EINVAL, "interval outside [1, limit]");
    if (last - first + 1 > S->pool_positions)
        return fail(HSS_EINVAL, "request has more positions than the state pool holds");
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->failed != HSS_OK) return fail(S->failed, "server failed: " + S->fail_msg);
    if (S->stop) return fail(HSS_EINVAL, "server is shutting down");
    const uint64_t rid = S->next_id++;
    ReqRec r;
    r.limit = limit;
    r.first = first;
    r.last = last;
    r.npos = (uint32_t)(last - first + 1);
    r.resume = first;
    r.t_submit = S->now();
    S->reqs.emplace(rid, std::move(r));
    S->sched.enqueue(rid, (uint32_t)(last - first + 1));
    ++S->st.submitted;
    S->cv_work.notify_all();
    *id = rid;
    return HSS_OK;
}

static void fill_info(uint64_t id, const ReqRec& r, hss_request_info* info) {
    if (!info) return;
    *info = hss_request_info{};
    info->id = id;
    info->limit = r.limit;
    info->first = r.first;
    info->last = r.last;
    info->resume = r.resume;
    const bool finished = r.state == HSS_REQ_DONE || r.state == HSS_REQ_CANCELED;
    if (finished) {
        info->candidates = r.cand;
        info->leaves = r.leaves;
        info->pairs = r.pairs;
    } else {
        const uint64_t prefix = r.per.empty() ? 0 : r.resume - r.first;
        for (uint64_t i = 0; i < prefix; ++i) {
            info->candidates += r.per[i].candidates;
            info->leaves += r.per[i].leaves;
            info->pairs += r.per[i].pairs;
        }
    }
    info->n_hits = (uint32_t)r.hits.size();
    info->launches = r.launches;
    info->state = r.state;
    info->error = r.error;
    info->submit_ms = r.t_submit;
    info->admit_ms = r.t_admit;
    info->start_ms = r.t_start;
    info->done_ms = r.t_done;
}

extern "C" int hss_server_wait(hss_server* S, uint64_t id, int64_t timeout_ms, hss_request_info* info) {
    if (!S) return fail(HSS_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(HSS_EINVAL, "unknown request id");
    auto finished = [&] {
        const int32_t st = it->second.state;
        return st == HSS_REQ_DONE || st == HSS_REQ_FAILED || st == HSS_REQ_CANCELED;
    };
    if (timeout_ms < 0)
        S->cv_done.wait(lk, finished);
    else
        S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), finished);
    fill_info(id, it->second, info);
    if (!finished()) return fail(HSS_EAGAIN, "request not finished yet");
    if (it->second.state == HSS_REQ_FAILED) return fail(it->second.error, "request failed: " + S->fail_msg);
    return HSS_OK;
}

extern "C" int hss_server_results(hss_server* S, uint64_t id, hss_root_result* per_root, uint64_t cap,
                                  hss_hit* hits, uint32_t hit_cap, uint32_t* n_hits) {
    if (!S) return fail(HSS_EINVAL, "null server");
    if ((cap && !per_root) || (hit_cap && !hits)) return fail(HSS_EINVAL, "null buffer");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(HSS_EINVAL, "unknown request id");
    const ReqRec& r = it->second;
    const uint64_t npos = r.last - r.first + 1;
    for (uint64_t i = 0; i < npos && i < cap; ++i) per_root[i] = i < r.per.size() ? r.per[i] : hss_root_result{0, 0, 0};
    // hits found so far, unique and sorted (a finished request's list already is)
    std::set<std::tuple<uint64_t, uint64_t, uint64_t>> seen;
    std::vector<HitOut> uniq;
    for (const HitOut& x : r.hits)
        if (seen.insert({x.l, x.m, x.n}).second) uniq.push_back(x);
    std::sort(uniq.begin(), uniq.end(), [](const HitOut& a, const HitOut& b) {
        return std::tie(a.l, a.m, a.n) < std::tie(b.l, b.m, b.n);
    });
    const uint32_t k = (uint32_t)uniq.size();
    for (uint32_t i = 0; i < k && i < hit_cap; ++i)
        hits[i] = hss_hit{uniq[i].l, uniq[i].m, uniq[i].n, uniq[i].area4_lo, uniq[i].area4_hi};
    if (n_hits) *n_hits = k;
    if (cap < npos || k > hit_cap) return fail(HSS_EOVERFLOW, "result buffers too small");
    return HSS_OK;
}

extern "C" int hss_server_release(hss_server* S, uint64_t id) {
    if (!S) return fail(HSS_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(HSS_EINVAL, "unknown request id");
    const int32_t st = it->second.state;
    if (st != HSS_REQ_DONE && st != HSS_REQ_FAILED && st != HSS_REQ_CANCELED)
        return fail(HSS_EAGAIN, "request not finished");
    S->reqs.erase(it);
    return HSS_OK;
}

extern "C" int hss_server_cancel(hss_server* S, uint64_t id) {
    if (!S) return fail(HSS_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(HSS_EINVAL, "unknown request id");
    ReqRec& r = it->second;
    if (r.state != HSS_REQ_QUEUED && r.state != HSS_REQ_ACTIVE) return HSS_OK;  // already finished
    r.canceled = true;
    const int how = S->sched.cancel(id);
    if (how == 1) {  // never admitted
        r.state = HSS_REQ_CANCELED;
        r.resume = r.first;
        r.t_done = S->now();
        ++S->st.completed;
        S->cv_done.notify_all();
    } else if (how == 2) {  // every scheduled position has finished
        finalize_locked(S, id);
    }
    // how == 3: finishes with the launches that hold its scheduled positions
    return HSS_OK;
}

extern "C" int hss_server_get_stats(hss_server* S, hss_server_stats* out) {
    if (!S || !out) return fail(HSS_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->st;
    out->uptime_ms = S->now();
    out->queued = (uint32_t)S->sched.queued();
    out->active = (uint32_t)S->sched.active();
    out->pool_free = (uint32_t)S->sched.pool_free();
    return HSS_OK;
}

extern "C" void hss_server_destroy(hss_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv_work.notify_all();
    }
    if (S->worker.joinable()) S->worker.join();
    (void)hipSetDevice(S->ctx->device);
    (void)hipStreamSynchronize(S->ctx->stream);
    if (S->copy) {
        (void)hipStreamSynchronize(S->copy);
        (void)hipStreamDestroy(S->copy);
    }
    for (LaunchBuffers& b : S->bufs) b.release();
    S->staging.release();
    S->pool.release();
    S->ctx->server = nullptr;
    delete S;
}

extern "C" int hss_split_by_work(const hss_ctx* c, uint64_t limit, uint64_t first, uint64_t last, uint32_t parts,
                                 uint64_t* bounds) {
    if (!c || !bounds) return fail(HSS_EINVAL, "null argument");
    if (!parts || parts > (1u << 20)) return fail(HSS_EINVAL, "parts outside [1, 2^20]");
    std::string err;
    if (!split_by_work(limit, first, last, parts, c->small_primes, c->h_off.data(), c->nmax, bounds, err))
        return fail(HSS_EINVAL, err);
    return HSS_OK;
}

extern "C" int hss_estimate(const hss_ctx* c, uint64_t limit, uint64_t first, uint64_t last, double* cells) {
    if (!c || !cells) return fail(HSS_EINVAL, "null argument");
    if (limit < 1 || limit > c->nmax) return fail(HSS_EINVAL, "limit outside [1, n_max]");
    if (first < 1 || first > last || last > limit) return fail(HSS_EINVAL, "interval outside [1, limit]");
    double s = 0;
    for (uint64_t l = first; l <= last; ++l) {
        double e = estimate_position(l, limit, c->small_primes, c->h_off.data(), c->nmax);
        if (e < 0) return fail(HSS_EINVAL, "position cannot be planned");
        s += e;
    }
    *cells = s;
    return HSS_OK;
}
