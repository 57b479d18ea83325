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
    int32_t state = HSS_REQ_QUEUED;
    int32_t error = HSS_OK;
    uint32_t launches = 0;
    uint64_t resume = 0;
    double t_submit = -1, t_admit = -1, t_start = -1, t_done = -1;
    std::vector<hss_root_result> per;
    std::vector<HitOut> hits;
    uint64_t cand = 0, leaves = 0, pairs = 0;
};

struct Launch {
    int buf = 0;
    std::vector<Segment> segs;
    std::vector<uint64_t> owner;  // plan position -> request id
    int attempts = 0;
};

}  // namespace

struct hss_server {
    hss_ctx* ctx = nullptr;
    hipStream_t copy = nullptr;
    DevBuf pool;
    Counters cnt;
    uint64_t pool_positions = 0;
    LaunchBuffers bufs[2];
    HostBuf readback;
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

// Copy a finished request's counters out of the pool and complete it.
int complete_request(hss_server* S, uint64_t id) {
    uint32_t off, npos;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        off = S->sched.pool_offset(id);
        npos = S->reqs[id].npos;
    }
    HSS_HIP_TRY(S->readback.reserve((size_t)npos * 24));
    unsigned long long* h = S->readback.as<unsigned long long>();
    HSS_HIP_TRY(hipMemcpyAsync(h, S->cnt.cand + off, (size_t)npos * 8, hipMemcpyDeviceToHost, S->copy));
    HSS_HIP_TRY(hipMemcpyAsync(h + npos, S->cnt.leaves + off, (size_t)npos * 8, hipMemcpyDeviceToHost, S->copy));
    HSS_HIP_TRY(hipMemcpyAsync(h + 2 * (size_t)npos, S->cnt.pairs + off, (size_t)npos * 8, hipMemcpyDeviceToHost,
                               S->copy));
    HSS_HIP_TRY(hipStreamSynchronize(S->copy));
    std::lock_guard<std::mutex> lk(S->mu);
    ReqRec& r = S->reqs[id];
    r.per.resize(npos);
    for (uint32_t i = 0; i < npos; ++i) {
        r.per[i] = hss_root_result{h[i], h[npos + i], h[2 * (size_t)npos + i]};
        r.cand += h[i];
        r.leaves += h[npos + i];
        r.pairs += h[2 * (size_t)npos + i];
    }
    std::set<std::tuple<uint64_t, uint64_t, uint64_t>> seen;
    std::vector<HitOut> uniq;
    for (const HitOut& x : r.hits)
        if (seen.insert({x.l, x.m, x.n}).second) uniq.push_back(x);
    std::sort(uniq.begin(), uniq.end(), [](const HitOut& a, const HitOut& b) {
        return std::tie(a.l, a.m, a.n) < std::tie(b.l, b.m, b.n);
    });
    r.hits.swap(uniq);
    r.resume = r.last + 1;
    r.state = HSS_REQ_DONE;
    r.t_done = S->now();
    S->sched.retire(id);
    ++S->st.completed;
    S->st.positions += npos;
    S->st.leaves += r.leaves;
    S->st.candidates += r.cand;
    S->cv_done.notify_all();
    return HSS_OK;
}

void worker_main(hss_server* S) {
    hss_ctx* c = S->ctx;
    if (hipSetDevice(c->device) != hipSuccess) {
        fail_all(S, HSS_EDEVICE, "hipSetDevice failed in the server thread");
        return;
    }
    std::deque<Launch> inflight;
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
            std::lock_guard<std::mutex> lk(S->mu);
            for (size_t a = 0; a < admitted.size(); ++a) S->sched.set_costs(admitted[a], std::move(costs[a]));
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
            int rc = launch_plan(c, S->bufs[L.buf], pr.data(), (uint32_t)pr.size());
            if (rc == HSS_OK) rc = launch_enqueue(c, S->bufs[L.buf], S->cnt, c->stream);
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
        std::vector<uint64_t> done;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            ++S->st.launches;
            S->st.gpu_ms += (double)R.setup_ms + R.cells_ms + R.exact_ms;
            for (const HitOut& h : R.hits)
                if (h.pos < L.owner.size()) S->reqs[L.owner[h.pos]].hits.push_back(h);
            done = S->sched.finished(L.segs);
            for (const Segment& sg : L.segs) {
                ReqRec& r = S->reqs[sg.id];
                r.resume = r.first + S->sched.prefix(sg.id);
            }
        }
        for (uint64_t id : done) {
            rc = complete_request(S, id);
            if (rc != HSS_OK) {
                fail_all(S, rc, hss_last_error());
                (void)hipStreamSynchronize(c->stream);
                return;
            }
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
        S->readback.release();
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
    info->candidates = r.cand;
    info->leaves = r.leaves;
    info->pairs = r.pairs;
    info->n_hits = r.state == HSS_REQ_DONE ? (uint32_t)r.hits.size() : 0u;
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
    auto finished = [&] { return it->second.state == HSS_REQ_DONE || it->second.state == HSS_REQ_FAILED; };
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
    if (r.state != HSS_REQ_DONE) return fail(HSS_EAGAIN, "request not finished");
    for (uint64_t i = 0; i < r.per.size() && i < cap; ++i) per_root[i] = r.per[i];
    const uint32_t k = (uint32_t)r.hits.size();
    for (uint32_t i = 0; i < k && i < hit_cap; ++i)
        hits[i] = hss_hit{r.hits[i].l, r.hits[i].m, r.hits[i].n, r.hits[i].area4_lo, r.hits[i].area4_hi};
    if (n_hits) *n_hits = k;
    if (cap < r.per.size() || k > hit_cap) return fail(HSS_EOVERFLOW, "result buffers too small");
    return HSS_OK;
}

extern "C" int hss_server_release(hss_server* S, uint64_t id) {
    if (!S) return fail(HSS_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(HSS_EINVAL, "unknown request id");
    if (it->second.state != HSS_REQ_DONE && it->second.state != HSS_REQ_FAILED)
        return fail(HSS_EAGAIN, "request not finished");
    S->reqs.erase(it);
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
    S->readback.release();
    S->pool.release();
    S->ctx->server = nullptr;
    delete S;
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
