// hss_runtime.cpp -- host runtime behind the C ABI of include/hss.h:
// context creation (resident number-theory tables), the launch sequence of one
// batch of root positions, and the synchronous search entry points.
//
// One context = one HIP device, one stream (default priority), the resident
// tables, and grow-only buffer sets for launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "hss_ctx.h"

using namespace hss;

namespace hss {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

hipError_t DevBuf::reserve(size_t want) {
    if (want <= bytes) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    size_t n = std::max<size_t>(want + want / 4, 4096);
    hipError_t e = hipMalloc(&p, n);
    if (e == hipSuccess) bytes = n;
    return e;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

hipError_t HostBuf::reserve(size_t want) {
    if (want <= bytes) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    size_t n = std::max<size_t>(want + want / 4, 4096);
    hipError_t e = hipHostMalloc(&p, n, hipHostMallocDefault);
    if (e == hipSuccess) bytes = n;
    return e;
}

void HostBuf::release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
}

hipError_t LaunchBuffers::create() {
    for (hipEvent_t& e : ev) {
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) return r;
    }
    return h_misc.reserve(64);
}

void LaunchBuffers::drop_graph() {
    for (int i = 0; i < STAGES; ++i) {
        if (exec[i]) (void)hipGraphExecDestroy(exec[i]);
        if (graph[i]) (void)hipGraphDestroy(graph[i]);
        exec[i] = nullptr;
        graph[i] = nullptr;
    }
    layout.clear();
}

void LaunchBuffers::release() {
    drop_graph();
    for (DevBuf* b : {&header, &pos, &lefts, &builds, &tiles, &wtiles, &sorted, &perm, &btab, &lfast, &lre, &lim, &misc,
                      &def, &raw, &hits})
        b->release();
    upload.release();
    upload2.release();
    h_misc.release();
    for (hipEvent_t& e : ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
}

// Forget device buffers that are views into memory B does not own (the
// no-allocation search's scratch array), so that B.release() frees none of it.
void clear_views(LaunchBuffers& B) {
    for (DevBuf* d : {&B.header, &B.pos, &B.lefts, &B.builds, &B.tiles, &B.wtiles, &B.sorted, &B.perm, &B.btab,
                      &B.lfast, &B.lre, &B.lim, &B.misc, &B.def, &B.raw, &B.hits})
        *d = DevBuf();
}

}  // namespace hss

namespace {

#define HIP_TRY HSS_HIP_TRY

// HSS_VERBOSE=1: synchronize and report after each phase (diagnostics only)
bool verbose() {
    static const bool v = std::getenv("HSS_VERBOSE") != nullptr;
    return v;
}

void vlog(hipStream_t st, const char* what) {
    if (!verbose()) return;
    static const auto t0 = std::chrono::steady_clock::now();
    hipError_t e = st ? hipStreamSynchronize(st) : hipSuccess;
    std::fprintf(stderr, "[hss] %-28s t=%8.3f s  %s\n", what,
                 std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(),
                 e == hipSuccess ? "" : hipGetErrorString(e));
    std::fflush(stderr);
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" const char* hss_last_error(void) { return hss::g_error.c_str(); }

extern "C" uint64_t hss_ctx_nmax(const hss_ctx* c) { return c ? c->nmax : 0; }

extern "C" uint64_t hss_ctx_device_bytes(const hss_ctx* c) { return c ? c->device_bytes : 0; }

extern "C" void hss_ctx_destroy(hss_ctx* c) {
    if (!c) return;
    if (c->server) hss_server_destroy(c->server);  // finishes its requests first
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    (void)hipFree(c->d_off);
    (void)hipFree(c->d_cre);
    (void)hipFree(c->d_cim);
    (void)hipFree(c->d_cres);
    (void)hipFree(c->d_vbase);
    (void)hipFree(c->d_vre);
    (void)hipFree(c->d_vim);
    c->sync.release();
    clear_views(c->noalloc);  // views into the scratch array: nothing to free
    c->noalloc.release();
    c->counters.release();
    if (c->capture) (void)hipStreamDestroy(c->capture);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

static int build_tables(hss_ctx* c) {
    const uint64_t n = c->nmax;
    hipStream_t st = c->stream;
    uint32_t* d_spf = nullptr;
    uint32_t* d_H = nullptr;
    uint32_t* d_sums = nullptr;
    uint32_t* d_small = nullptr;
    uint32_t *d_plist = nullptr, *d_psize = nullptr, *d_vstart = nullptr, *d_cnt = nullptr, *d_psums = nullptr;
    int64_t *d_pa = nullptr, *d_pb = nullptr;
    struct Cleanup {
        std::vector<void*> ptrs;
        ~Cleanup() {
            for (void* p : ptrs)
                if (p) (void)hipFree(p);
        }
    } cleanup;
    const uint32_t root = (uint32_t)std::sqrt((double)n) + 1;
    c->small_primes = primes_upto(std::max<uint32_t>(root, 2u));
    std::vector<uint32_t> sieve_primes;
    for (uint32_t p : c->small_primes)
        if ((uint64_t)p * p <= n) sieve_primes.push_back(p);
    HIP_TRY(hipMalloc(&d_spf, (n + 1) * 4));
    cleanup.ptrs.push_back(d_spf);
    HIP_TRY(hipMalloc(&d_H, (n + 2) * 4));
    cleanup.ptrs.push_back(d_H);
    HIP_TRY(hipMalloc(&d_sums, ((n + 2 + 4095) / 4096 + 1) * 4));
    cleanup.ptrs.push_back(d_sums);
    HIP_TRY(hipMalloc(&d_small, std::max<size_t>(sieve_primes.size(), 1) * 4));
    cleanup.ptrs.push_back(d_small);
    if (!sieve_primes.empty())
        HIP_TRY(hipMemcpyAsync(d_small, sieve_primes.data(), sieve_primes.size() * 4, hipMemcpyHostToDevice, st));
    BuildBuffers B{n, d_spf, d_H, d_small, (uint32_t)sieve_primes.size()};
    HIP_TRY(launch_build_tables(B, st));
    vlog(st, "sieve + half counts");
    HIP_TRY(hipMalloc(&c->d_off, (n + 2) * 4));
    HIP_TRY(launch_scan(d_H, c->d_off, d_sums, n + 2, st));
    vlog(st, "offset scan");
    c->h_off.resize(n + 2);
    HIP_TRY(hipMemcpyAsync(c->h_off.data(), c->d_off, (n + 2) * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    uint64_t total = c->h_off[n + 1];
    if (total >= 0xFFFFFFFFull) return fail(HSS_EINVAL, "cache larger than 2^32 entries");
    for (uint64_t k = 0; k <= n; ++k)
        if (c->h_off[k + 1] - c->h_off[k] > MAX_HALF_SET) return fail(HSS_EINVAL, "half-set larger than 2^15 - 1");
    c->cache_total = (uint32_t)total;
    // split primes and their Gaussian value blocks
    uint64_t cap = n >= 5 ? (n - 5) / 4 + 1 : 1;
    HIP_TRY(hipMalloc(&d_plist, cap * 4));
    cleanup.ptrs.push_back(d_plist);
    HIP_TRY(hipMalloc(&d_psize, cap * 4));
    cleanup.ptrs.push_back(d_psize);
    HIP_TRY(hipMalloc(&d_vstart, cap * 4));
    cleanup.ptrs.push_back(d_vstart);
    HIP_TRY(hipMalloc(&d_cnt, 4));
    cleanup.ptrs.push_back(d_cnt);
    HIP_TRY(hipMemsetAsync(d_cnt, 0, 4, st));
    HIP_TRY(launch_collect_split(d_spf, n, d_plist, d_psize, d_cnt, st));
    vlog(st, "split primes collected");
    uint32_t np = 0;
    HIP_TRY(hipMemcpyAsync(&np, d_cnt, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    uint32_t nvals = 0;
    HIP_TRY(hipMalloc(&c->d_vbase, (n + 1) * 4));
    if (np) {
        HIP_TRY(hipMalloc(&d_psums, ((np + 4095) / 4096 + 1) * 4));
        cleanup.ptrs.push_back(d_psums);
        HIP_TRY(launch_scan(d_psize, d_vstart, d_psums, np, st));
        uint32_t last_start = 0, last_size = 0;
        HIP_TRY(hipMemcpyAsync(&last_start, d_vstart + np - 1, 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(&last_size, d_psize + np - 1, 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        nvals = last_start + last_size;
        HIP_TRY(hipMalloc(&d_pa, np * 8));
        cleanup.ptrs.push_back(d_pa);
        HIP_TRY(hipMalloc(&d_pb, np * 8));
        cleanup.ptrs.push_back(d_pb);
        HIP_TRY(hipMalloc(&c->d_vre, (size_t)nvals * sizeof(i128)));
        HIP_TRY(hipMalloc(&c->d_vim, (size_t)nvals * sizeof(i128)));
        HIP_TRY(launch_prime_values(d_plist, np, d_vstart, nvals, c->d_vbase, d_pa, d_pb, c->d_vre, c->d_vim, st));
        vlog(st, "prime values");
    }
    HIP_TRY(hipMalloc(&c->d_cre, std::max<size_t>(total, 1) * 8));
    HIP_TRY(hipMalloc(&c->d_cim, std::max<size_t>(total, 1) * 8));
    HIP_TRY(hipMalloc(&c->d_cres, std::max<size_t>(total, 1) * 4));
    HIP_TRY(launch_fill_cache(d_spf, c->d_off, n, (uint32_t)total, c->d_vbase, c->d_vre, c->d_vim, c->d_cre,
                              c->d_cim, c->d_cres, st));
    vlog(st, "cache filled");
    HIP_TRY(hipStreamSynchronize(st));
    c->device_bytes = (n + 2) * 4 + (n + 1) * 4 + (uint64_t)nvals * 32 + total * 20;
    return HSS_OK;
}

extern "C" int hss_ctx_create(uint64_t n_max, int device, hss_ctx** out) {
    if (!out) return fail(HSS_EINVAL, "out is null");
    *out = nullptr;
    if (n_max < 1 || n_max > (1ull << 26)) return fail(HSS_EINVAL, "n_max must be in [1, 2^26]");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(HSS_EDEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return fail(HSS_EINVAL, "device index out of range");
    hss_ctx* c = new hss_ctx();
    c->device = device;
    c->nmax = n_max;
    c->qr = make_qr4();
    auto bail = [&](int code) {
        std::string keep = g_error;
        hss_ctx_destroy(c);
        g_error = keep;
        return code;
    };
    vlog(nullptr, "device query");
    if (hipSetDevice(device) != hipSuccess) return bail(fail(HSS_EDEVICE, "hipSetDevice failed"));
    vlog(nullptr, "device set");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return bail(fail(HSS_EDEVICE, "device properties"));
    c->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->capture, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(HSS_EDEVICE, "stream create failed"));
    c->use_graphs = std::getenv("HSS_NO_GRAPH") == nullptr && !verbose();
    if (c->sync.create() != hipSuccess || c->noalloc.create() != hipSuccess)
        return bail(fail(HSS_EDEVICE, "launch buffers: event or pinned memory"));
    if (noalloc_scratch(device, &c->scratch) != hipSuccess)
        return bail(fail(HSS_EDEVICE, "address of the no-allocation scratch array"));
    int bpc = 0;
    if (cell_occupancy(&bpc) != hipSuccess || bpc < 1) bpc = 1;
    c->blocks_per_cu = bpc;
    bpc = 0;
    if (cell_win_occupancy(&bpc) != hipSuccess || bpc < 1) bpc = 1;
    c->win_blocks_per_cu = bpc;
    vlog(nullptr, "stream + occupancy");
    int rc = build_tables(c);
    vlog(c->stream, "tables built");
    if (rc != HSS_OK) return bail(rc);
    *out = c;
    return HSS_OK;
}

extern "C" int hss_ctx_use_graphs(hss_ctx* c, int enable) {
    if (!c) return fail(HSS_EINVAL, "null context");
    const int old = c->use_graphs ? 1 : 0;
    c->use_graphs = enable != 0;
    if (!enable) c->sync.drop_graph();
    return old;
}

extern "C" int hss_ctx_get_tuning(const hss_ctx* c, hss_tuning* out) {
    if (!c || !out) return fail(HSS_EINVAL, "null argument");
    *out = hss_tuning{};
    out->tiles_per_wave = c->tune.tiles_per_wave;
    for (int i = 0; i < HSS_WEIGHT_CLASSES; ++i) out->win_max_frac[i] = c->tune.win_max_frac[i];
    out->win_item_cost = c->tune.win_item_cost;
    out->win_partner_cost = c->tune.win_partner_cost;
    out->plain_partner_cost = c->tune.plain_partner_cost;
    return HSS_OK;
}

extern "C" int hss_ctx_set_tuning(hss_ctx* c, const hss_tuning* in) {
    if (!c || !in) return fail(HSS_EINVAL, "null argument");
    if (c->server) return fail(HSS_EINVAL, "context is attached to a server");
    if (!in->tiles_per_wave || in->tiles_per_wave > 1024) return fail(HSS_EINVAL, "tiles_per_wave outside [1, 1024]");
    for (int i = 0; i < HSS_WEIGHT_CLASSES; ++i)
        if (!(in->win_max_frac[i] >= 0.0 && in->win_max_frac[i] <= 1.0))
            return fail(HSS_EINVAL, "win_max_frac outside [0, 1]");
    if (!(in->win_item_cost > 0.0) || !(in->win_partner_cost >= 0.0) || !(in->plain_partner_cost >= 0.0))
        return fail(HSS_EINVAL, "bad cost weights");
    c->tune.tiles_per_wave = in->tiles_per_wave;
    for (int i = 0; i < HSS_WEIGHT_CLASSES; ++i) c->tune.win_max_frac[i] = in->win_max_frac[i];
    c->tune.win_item_cost = in->win_item_cost;
    c->tune.win_partner_cost = in->win_partner_cost;
    c->tune.plain_partner_cost = in->plain_partner_cost;
    return HSS_OK;
}

extern "C" int hss_autotune(hss_ctx* c, uint64_t limit, const uint64_t* roots, uint32_t n_roots, uint32_t repeats,
                            hss_tuning* out, double* best_ms) {
    if (!c || (!roots && n_roots)) return fail(HSS_EINVAL, "null argument");
    if (!n_roots) return fail(HSS_EINVAL, "no calibration roots");
    if (c->server) return fail(HSS_EINVAL, "context is attached to a server");
    if (!repeats) repeats = 1;
    const Tuning start = c->tune;
    uint64_t ref_total = 0;
    bool have_ref = false;
    // device time (setup + cells + exact) of the calibration batch under t, best of `repeats`
    auto measure = [&](const Tuning& t, double& ms) -> int {
        c->tune = t;
        ms = 1e300;
        for (uint32_t r = 0; r < repeats; ++r) {
            uint64_t total = 0;
            int rc = hss_search(c, limit, roots, n_roots, nullptr, &total, nullptr, 0, nullptr);
            if (rc != HSS_OK && rc != HSS_EOVERFLOW) return rc;
            if (!have_ref) {
                ref_total = total;
                have_ref = true;
            } else if (total != ref_total) {
                return fail(HSS_EDEVICE, "autotune: a tuning changed the candidate count");
            }
            ms = std::min(ms, c->stats.setup_ms + c->stats.cells_ms + c->stats.exact_ms);
        }
        return HSS_OK;
    };
    // weight classes in order of their share of the batch's cells
    double share[WEIGHT_CLASSES] = {0};
    {
        std::vector<PosReq> reqs(n_roots);
        for (uint32_t s = 0; s < n_roots; ++s) reqs[s] = PosReq{roots[s], limit, s};
        Plan P;
        std::string err;
        if (!plan_batch(reqs.data(), n_roots, c->small_primes, c->h_off.data(), c->nmax, c->waves(), start, P, err))
            return fail(HSS_EINVAL, err);
        for (size_t li = 0; li < P.lefts.size(); ++li) {
            const int wc = weight_class(P.lefts[li].H);
            if (wc >= 0)
                share[wc] += (double)P.lefts[li].H * ((double)c->h_off[P.kmax[li] + 1] - c->h_off[P.kmin[li]]);
        }
    }
    std::vector<int> order;
    for (int i = 0; i < WEIGHT_CLASSES; ++i)
        if (share[i] > 0) order.push_back(i);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return share[a] > share[b]; });
    Tuning best = start;
    double best_t = 0;
    int rc = measure(best, best_t);
    if (rc != HSS_OK) {
        c->tune = start;
        return rc;
    }
    static const double grid[] = {0.0, 0.05, 0.1, 0.2, 0.3, 0.45, 0.6, 0.8, 1.0};
    for (int wc : order)
        for (double g : grid) {
            if (g == best.win_max_frac[wc]) continue;
            Tuning t = best;
            t.win_max_frac[wc] = g;
            double ms;
            rc = measure(t, ms);
            if (rc != HSS_OK) {
                c->tune = start;
                return rc;
            }
            if (ms < best_t) {
                best_t = ms;
                best = t;
            }
        }
    for (uint32_t tpw : {8u, 16u, 32u, 48u, 64u}) {
        if (tpw == best.tiles_per_wave) continue;
        Tuning t = best;
        t.tiles_per_wave = tpw;
        double ms;
        rc = measure(t, ms);
        if (rc != HSS_OK) {
            c->tune = start;
            return rc;
        }
        if (ms < best_t) {
            best_t = ms;
            best = t;
        }
    }
    c->tune = best;
    if (out) (void)hss_ctx_get_tuning(c, out);
    if (best_ms) *best_ms = best_t;
    return HSS_OK;
}

extern "C" int hss_get_stats(const hss_ctx* c, hss_stats* out) {
    if (!c || !out) return fail(HSS_EINVAL, "null argument");
    *out = c->stats;
    return HSS_OK;
}

namespace hss {

int launch_plan(hss_ctx* c, LaunchBuffers& B, const PosReq* reqs, uint32_t n) {
    std::string err;
    if (!plan_batch(reqs, n, c->small_primes, c->h_off.data(), c->nmax, c->waves(), c->tune, B.plan, err))
        return fail(HSS_EINVAL, err);
    B.npos = n;
    return HSS_OK;
}

int launch_plan_lefts(hss_ctx* c, LaunchBuffers& B, const PosReq* reqs, uint32_t n) {
    std::string err;
    if (!plan_lefts(reqs, n, c->small_primes, c->h_off.data(), c->nmax, c->tune, B.plan, err))
        return fail(HSS_EINVAL, err);
    B.npos = n;
    return HSS_OK;
}

int launch_plan_tiles(hss_ctx* c, LaunchBuffers& B) {
    if (!plan_tiles(c->h_off.data(), c->waves(), c->tune, B.plan)) return fail(HSS_EINVAL, "tile planning failed");
    return HSS_OK;
}

// The kernel chain of one launch in four stages on stream st (timing events
// go between the stages): 0 = left-set sort and build and the counter reset,
// 1 = window-mode cell kernel, 2 = plain cell kernel, 3 = exact and
// acceptance kernels.  Every size comes from the device header and every grid
// is fixed, so one capture of a stage serves every batch.
static int enqueue_stage(hss_ctx* c, LaunchBuffers& B, const CellParams& K, int stage, hipStream_t st) {
    uint32_t* misc = B.misc.as<uint32_t>();
    if (stage == 0) {
        HIP_TRY(launch_build_left(B.header.as<BatchHeader>(), B.builds.as<LeftBuild>(), B.lefts.as<LeftDesc>(),
                                  c->d_vbase, c->d_vre, c->d_vim, B.sorted.as<uint32_t>(), B.perm.as<uint32_t>(),
                                  B.btab.as<uint32_t>(), B.lre.as<i128>(), B.lim.as<i128>(), B.lfast.as<LeftFast>(),
                                  st));
        HIP_TRY(launch_batch_reset(B.header.as<BatchHeader>(), K.pos, K.cand, K.leaves, K.pairs, misc, st));
    } else if (stage == 1) {
        HIP_TRY(launch_cells_win(K, c->win_grid_blocks(), st));
    } else if (stage == 2) {
        HIP_TRY(launch_cells(K, c->grid_blocks(), st));
    } else {
        HIP_TRY(launch_exact(K, (uint32_t)c->cus * 4, st));
        HIP_TRY(launch_accept(K.pos, K.raw, K.nraw, LaunchBuffers::RAW_CAP, B.hits.as<HitOut>(), misc + 3,
                              LaunchBuffers::HIT_CAP, st));
    }
    return HSS_OK;
}

// Capture the stages (once per buffer layout).  false: capture is not
// available here (the reason goes to c->graph_error).
static bool capture_stages(hss_ctx* c, LaunchBuffers& B, const CellParams& K) {
    B.drop_graph();
    std::string why;
    for (int stage = 0; stage < LaunchBuffers::STAGES; ++stage) {
        hipError_t e = hipStreamBeginCapture(c->capture, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            why = hipGetErrorString(e);
            break;
        }
        const int rc = enqueue_stage(c, B, K, stage, c->capture);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(c->capture, &g);
        if (rc != HSS_OK || e != hipSuccess || !g) {
            why = rc != HSS_OK ? std::string(hss_last_error()) : std::string(hipGetErrorString(e));
            if (g) (void)hipGraphDestroy(g);
            break;
        }
        e = hipGraphInstantiate(&B.exec[stage], g, nullptr, nullptr, 0);
        B.graph[stage] = g;
        if (e != hipSuccess) {
            B.exec[stage] = nullptr;
            why = hipGetErrorString(e);
            break;
        }
    }
    if (!why.empty()) {
        B.drop_graph();
        c->graph_error = why;
        (void)hipGetLastError();
        return false;
    }
    return true;
}

// Device buffers of the positions and left sets (setup) and of the tiles.
static int reserve_setup(LaunchBuffers& B) {
    const Plan& P = B.plan;
    const uint32_t nleft = (uint32_t)P.lefts.size(), nsorted = (uint32_t)P.sorted.size();
    const size_t entries = std::max<uint64_t>(P.left_entries, 1);
    HIP_TRY(B.header.reserve(sizeof(BatchHeader)));
    HIP_TRY(B.pos.reserve(std::max<size_t>(B.npos, 1) * sizeof(PosDesc)));
    HIP_TRY(B.lefts.reserve(std::max<size_t>(nleft, 1) * sizeof(LeftDesc)));
    HIP_TRY(B.builds.reserve(std::max<size_t>(nleft, 1) * sizeof(LeftBuild)));
    HIP_TRY(B.sorted.reserve(std::max<size_t>(nsorted, 1) * 4));
    HIP_TRY(B.btab.reserve(std::max<size_t>(nsorted, 1) * (WIN_NB + 1) * 4));
    HIP_TRY(B.perm.reserve(entries * 4));
    HIP_TRY(B.lfast.reserve((entries + WIN_UNIT) * sizeof(LeftFast)));  // padding read by window units
    HIP_TRY(B.lre.reserve(entries * sizeof(i128)));
    HIP_TRY(B.lim.reserve(entries * sizeof(i128)));
    HIP_TRY(B.misc.reserve(64));
    HIP_TRY(B.raw.reserve(LaunchBuffers::RAW_CAP * sizeof(HitOut)));
    HIP_TRY(B.hits.reserve(LaunchBuffers::HIT_CAP * sizeof(HitOut)));
    if (!B.def_cap) {
        B.def_cap = 1u << 22;
        HIP_TRY(B.def.reserve((size_t)B.def_cap * sizeof(DefItem)));
    }
    return HSS_OK;
}

static int reserve_cells(LaunchBuffers& B) {
    HIP_TRY(B.tiles.reserve(std::max<size_t>(B.plan.tiles.size(), 1) * sizeof(TileDesc)));
    HIP_TRY(B.wtiles.reserve(std::max<size_t>(B.plan.wtiles.size(), 1) * sizeof(TileDesc)));
    return HSS_OK;
}

static BatchHeader batch_header(const LaunchBuffers& B, bool with_tiles) {
    BatchHeader hdr{};
    hdr.npos = B.npos;
    hdr.nleft = (uint32_t)B.plan.lefts.size();
    hdr.nsorted = (uint32_t)B.plan.sorted.size();
    hdr.ntiles = with_tiles ? (uint32_t)B.plan.tiles.size() : 0u;
    hdr.nwtiles = with_tiles ? (uint32_t)B.plan.wtiles.size() : 0u;
    return hdr;
}

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// Stage the header (with or without tile counts), positions, left sets and
// sorted list in `up` and copy them to the device.
static int upload_setup(LaunchBuffers& B, HostBuf& up, bool with_tiles, hipStream_t st) {
    const Plan& P = B.plan;
    const uint32_t npos = B.npos, nleft = (uint32_t)P.lefts.size(), nsorted = (uint32_t)P.sorted.size();
    const size_t b_hdr = al256(sizeof(BatchHeader)), b_pos = al256(npos * sizeof(PosDesc)),
                 b_left = al256(nleft * sizeof(LeftDesc)), b_build = al256(nleft * sizeof(LeftBuild)),
                 b_sorted = al256(nsorted * 4);
    HIP_TRY(up.reserve(b_hdr + b_pos + b_left + b_build + b_sorted));
    char* const up_hdr = up.as<char>();
    char* const up_pos = up_hdr + b_hdr;
    char* const up_left = up_pos + b_pos;
    char* const up_build = up_left + b_left;
    char* const up_sorted = up_build + b_build;
    const BatchHeader hdr = batch_header(B, with_tiles);
    std::memcpy(up_hdr, &hdr, sizeof hdr);
    std::memcpy(up_pos, P.pos.data(), npos * sizeof(PosDesc));
    std::memcpy(up_left, P.lefts.data(), nleft * sizeof(LeftDesc));
    std::memcpy(up_build, P.builds.data(), nleft * sizeof(LeftBuild));
    std::memcpy(up_sorted, P.sorted.data(), nsorted * 4);
    HIP_TRY(hipMemcpyAsync(B.header.p, up_hdr, sizeof hdr, hipMemcpyHostToDevice, st));
    if (npos) HIP_TRY(hipMemcpyAsync(B.pos.p, up_pos, npos * sizeof(PosDesc), hipMemcpyHostToDevice, st));
    if (nleft) {
        HIP_TRY(hipMemcpyAsync(B.lefts.p, up_left, nleft * sizeof(LeftDesc), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(B.builds.p, up_build, nleft * sizeof(LeftBuild), hipMemcpyHostToDevice, st));
    }
    if (nsorted) HIP_TRY(hipMemcpyAsync(B.sorted.p, up_sorted, nsorted * 4, hipMemcpyHostToDevice, st));
    return HSS_OK;
}

// Stage the tiles (and, with_header, the full header) in `up` and copy them.
static int upload_cells(LaunchBuffers& B, HostBuf& up, bool with_header, hipStream_t st) {
    const Plan& P = B.plan;
    const uint32_t ntiles = (uint32_t)P.tiles.size(), nwtiles = (uint32_t)P.wtiles.size();
    const size_t b_hdr = al256(sizeof(BatchHeader)), b_tile = al256(ntiles * sizeof(TileDesc)),
                 b_wtile = al256(nwtiles * sizeof(TileDesc));
    HIP_TRY(up.reserve(b_hdr + b_tile + b_wtile));
    char* const up_hdr = up.as<char>();
    char* const up_tile = up_hdr + b_hdr;
    char* const up_wtile = up_tile + b_tile;
    std::memcpy(up_tile, P.tiles.data(), ntiles * sizeof(TileDesc));
    std::memcpy(up_wtile, P.wtiles.data(), nwtiles * sizeof(TileDesc));
    if (with_header) {
        const BatchHeader hdr = batch_header(B, true);
        std::memcpy(up_hdr, &hdr, sizeof hdr);
        HIP_TRY(hipMemcpyAsync(B.header.p, up_hdr, sizeof hdr, hipMemcpyHostToDevice, st));
    }
    if (ntiles) HIP_TRY(hipMemcpyAsync(B.tiles.p, up_tile, ntiles * sizeof(TileDesc), hipMemcpyHostToDevice, st));
    if (nwtiles)
        HIP_TRY(hipMemcpyAsync(B.wtiles.p, up_wtile, nwtiles * sizeof(TileDesc), hipMemcpyHostToDevice, st));
    return HSS_OK;
}

static CellParams cell_params(hss_ctx* c, LaunchBuffers& B, const Counters& cnt) {
    // misc: [0] next plain tile, [1] deferred, [2] raw, [3] hits, [4] next window tile,
    // [6..9] evaluated (2 x u64), [10..13] candidates per kernel (2 x u64)
    uint32_t* misc = B.misc.as<uint32_t>();
    CellParams K{};
    K.off = c->d_off;
    K.cre = c->d_cre;
    K.cim = c->d_cim;
    K.cres = c->d_cres;
    K.pos = B.pos.as<PosDesc>();
    K.lefts = B.lefts.as<LeftDesc>();
    K.hdr = B.header.as<BatchHeader>();
    K.tiles = B.tiles.as<TileDesc>();
    K.wtiles = B.wtiles.as<TileDesc>();
    K.btab = B.btab.as<uint32_t>();
    K.lfast = B.lfast.as<LeftFast>();
    K.lre = B.lre.as<i128>();
    K.lim = B.lim.as<i128>();
    K.next_tile = misc;
    K.next_wtile = misc + 4;
    K.cand = cnt.cand;
    K.leaves = cnt.leaves;
    K.pairs = cnt.pairs;
    K.evaluated = reinterpret_cast<unsigned long long*>(misc + 6);
    K.def = B.def.as<DefItem>();
    K.ndef = misc + 1;
    K.def_cap = B.def_cap;
    K.raw = B.raw.as<HitOut>();
    K.nraw = misc + 2;
    K.raw_cap = LaunchBuffers::RAW_CAP;
    K.qr = c->qr;
    return K;
}

// Stages first .. STAGES - 1, replayed from graphs when they are available
// (captured once per buffer layout), then the read-back of misc.
static int run_stages(hss_ctx* c, LaunchBuffers& B, const CellParams& K, const Counters& cnt, int first,
                      hipStream_t st) {
    const uint32_t ntiles = (uint32_t)B.plan.tiles.size(), nwtiles = (uint32_t)B.plan.wtiles.size();
    bool replay = false;
    if (c->use_graphs) {
        std::vector<uintptr_t> layout;
        for (const DevBuf* d : {&B.header, &B.pos, &B.lefts, &B.builds, &B.tiles, &B.wtiles, &B.sorted, &B.perm,
                                &B.btab, &B.lfast, &B.lre, &B.lim, &B.misc, &B.def, &B.raw, &B.hits})
            layout.push_back((uintptr_t)d->p);
        for (const void* p : {(const void*)cnt.cand, (const void*)cnt.leaves, (const void*)cnt.pairs})
            layout.push_back((uintptr_t)p);
        layout.push_back(B.def_cap);
        if (!B.exec[0] || layout != B.layout) {
            if (capture_stages(c, B, K))
                B.layout = layout;
            else
                c->use_graphs = false;  // launch the stages directly from now on
        }
        replay = B.exec[0] != nullptr;
    }
    for (int stage = first; stage < LaunchBuffers::STAGES; ++stage) {
        const bool empty = (stage == 1 && !nwtiles) || (stage == 2 && !ntiles);  // nothing to run
        if (!empty && replay) HIP_TRY(hipGraphLaunch(B.exec[stage], st));
        if (!empty && !replay) {
            const int rc = enqueue_stage(c, B, K, stage, st);
            if (rc != HSS_OK) return rc;
        }
        HIP_TRY(hipEventRecord(B.ev[1 + stage], st));
        if (stage == 0) HIP_TRY(hipEventRecord(B.ev[6], st));
    }
    B.replayed = replay;
    vlog(st, "kernel chain");
    HIP_TRY(hipMemcpyAsync(B.h_misc.p, B.misc.p, 64, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipEventRecord(B.ev[5], st));
    return HSS_OK;
}

int launch_enqueue(hss_ctx* c, LaunchBuffers& B, const Counters& cnt, hipStream_t st) {
    int rc = reserve_setup(B);
    if (rc == HSS_OK) rc = reserve_cells(B);
    if (rc != HSS_OK) return rc;
    HIP_TRY(hipEventRecord(B.ev[0], st));
    rc = upload_setup(B, B.upload, true, st);
    if (rc == HSS_OK) rc = upload_cells(B, B.upload2, false, st);
    if (rc != HSS_OK) return rc;
    return run_stages(c, B, cell_params(c, B, cnt), cnt, 0, st);
}

int launch_enqueue_setup(hss_ctx* c, LaunchBuffers& B, const Counters& cnt, hipStream_t st) {
    int rc = reserve_setup(B);
    if (rc != HSS_OK) return rc;
    HIP_TRY(hipEventRecord(B.ev[0], st));
    rc = upload_setup(B, B.upload, false, st);  // no tiles yet: their counts follow with them
    if (rc != HSS_OK) return rc;
    rc = enqueue_stage(c, B, cell_params(c, B, cnt), 0, st);
    if (rc != HSS_OK) return rc;
    HIP_TRY(hipEventRecord(B.ev[1], st));
    return HSS_OK;
}

int launch_enqueue_cells(hss_ctx* c, LaunchBuffers& B, const Counters& cnt, hipStream_t st) {
    int rc = reserve_cells(B);
    if (rc == HSS_OK) rc = upload_cells(B, B.upload2, true, st);
    if (rc != HSS_OK) return rc;
    HIP_TRY(hipEventRecord(B.ev[6], st));  // the window kernel's time starts here, not at the end of the setup
    return run_stages(c, B, cell_params(c, B, cnt), cnt, 1, st);
}

void launch_read_misc(const LaunchBuffers& B, LaunchResult& out) {
    out = LaunchResult();
    const uint32_t* m4 = B.h_misc.as<uint32_t>();
    out.deferred = m4[1];
    out.raw = m4[2];
    std::memcpy(&out.win_evaluated, m4 + 6, 8);
    std::memcpy(&out.plain_evaluated, m4 + 8, 8);
    std::memcpy(&out.win_cand, m4 + 10, 8);
    std::memcpy(&out.plain_cand, m4 + 12, 8);
    out.evaluated = out.win_evaluated + out.plain_evaluated;
    out.graph = B.replayed;
    (void)hipEventElapsedTime(&out.setup_ms, B.ev[0], B.ev[1]);
    (void)hipEventElapsedTime(&out.win_ms, B.ev[6], B.ev[2]);
    (void)hipEventElapsedTime(&out.plain_ms, B.ev[2], B.ev[3]);
    (void)hipEventElapsedTime(&out.exact_ms, B.ev[3], B.ev[4]);
    out.cells_ms = out.win_ms + out.plain_ms;
}

int launch_collect(hss_ctx* c, LaunchBuffers& B, hipStream_t copy, LaunchResult& out) {
    launch_read_misc(B, out);
    const uint32_t* m4 = B.h_misc.as<uint32_t>();
    if (m4[1] > B.def_cap) {
        // deferred buffer overflowed: the counts are incomplete; grow for a rerun
        if (m4[1] > (1u << 30)) return fail(HSS_ENOMEM, "deferred-leaf buffer overflow");
        B.def_cap = m4[1] + m4[1] / 2;
        HIP_TRY(B.def.reserve((size_t)B.def_cap * sizeof(DefItem)));
        out.rerun = true;
        return HSS_OK;
    }
    if (m4[2] > LaunchBuffers::RAW_CAP) return fail(HSS_EOVERFLOW, "raw fourth-power buffer overflow");
    if (m4[3] > LaunchBuffers::HIT_CAP) return fail(HSS_EOVERFLOW, "accepted-hit buffer overflow");
    out.hits.resize(m4[3]);
    if (m4[3]) {
        HIP_TRY(hipMemcpyAsync(out.hits.data(), B.hits.p, m4[3] * sizeof(HitOut), hipMemcpyDeviceToHost, copy));
        HIP_TRY(hipStreamSynchronize(copy));
    }
    (void)c;
    return HSS_OK;
}

}  // namespace hss

static int finish_search(hss_ctx* c, const LaunchBuffers& B, const LaunchResult& R,
                         const std::vector<unsigned long long>& h, uint32_t n_roots, hss_root_result* per_root,
                         uint64_t* candidates_total, hss_hit* hits, uint32_t hit_cap, uint32_t* n_hits,
                         std::chrono::steady_clock::time_point t0);

extern "C" int hss_search(hss_ctx* c, uint64_t limit, const uint64_t* roots, uint32_t n_roots,
                          hss_root_result* per_root, uint64_t* candidates_total, hss_hit* hits, uint32_t hit_cap,
                          uint32_t* n_hits) {
    if (!c) return fail(HSS_EINVAL, "null context");
    if (c->server) return fail(HSS_EINVAL, "context is attached to a server");
    if (n_roots && !roots) return fail(HSS_EINVAL, "roots is null");
    if (hit_cap && !hits) return fail(HSS_EINVAL, "hits is null");
    if (hipSetDevice(c->device) != hipSuccess) return fail(HSS_EDEVICE, "hipSetDevice failed");
    c->stats = hss_stats{};
    if (n_hits) *n_hits = 0;
    if (candidates_total) *candidates_total = 0;
    if (!n_roots) return HSS_OK;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<PosReq> reqs(n_roots);
    for (uint32_t s = 0; s < n_roots; ++s) reqs[s] = PosReq{roots[s], limit, s};
    LaunchBuffers& B = c->sync;
    HIP_TRY(c->counters.reserve((size_t)n_roots * 3 * 8));
    Counters cnt;
    cnt.cand = c->counters.as<unsigned long long>();
    cnt.leaves = cnt.cand + n_roots;
    cnt.pairs = cnt.cand + 2 * (size_t)n_roots;
    // left sets first: their sort and build run on the device while the host
    // plans the tiles
    int rc = launch_plan_lefts(c, B, reqs.data(), n_roots);
    if (rc == HSS_OK) rc = launch_enqueue_setup(c, B, cnt, c->stream);
    if (rc == HSS_OK) rc = launch_plan_tiles(c, B);
    if (rc != HSS_OK) return rc;
    c->stats.plan_ms = ms_since(t0);
    LaunchResult R;
    for (int attempt = 0;; ++attempt) {
        // a second attempt (deferred buffer grown) replays the whole chain
        rc = attempt == 0 ? launch_enqueue_cells(c, B, cnt, c->stream) : launch_enqueue(c, B, cnt, c->stream);
        if (rc != HSS_OK) return rc;
        HIP_TRY(hipEventSynchronize(B.ev[5]));
        rc = launch_collect(c, B, c->stream, R);
        if (rc != HSS_OK) return rc;
        if (!R.rerun) break;
        if (attempt == 3) return fail(HSS_ENOMEM, "deferred-leaf buffer overflow");
    }
    std::vector<unsigned long long> h((size_t)n_roots * 3);
    HIP_TRY(hipMemcpyAsync(h.data(), c->counters.p, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return finish_search(c, B, R, h, n_roots, per_root, candidates_total, hits, hit_cap, n_hits, t0);
}

// Per-root counters, statistics and the de-duplicated, sorted hit list of a
// completed search launch (shared by hss_search and hss_search_noalloc).
static int finish_search(hss_ctx* c, const LaunchBuffers& B, const LaunchResult& R,
                         const std::vector<unsigned long long>& h, uint32_t n_roots, hss_root_result* per_root,
                         uint64_t* candidates_total, hss_hit* hits, uint32_t hit_cap, uint32_t* n_hits,
                         std::chrono::steady_clock::time_point t0) {
    uint64_t total = 0, leaves = 0;
    for (uint32_t s = 0; s < n_roots; ++s) {
        total += h[s];
        leaves += h[n_roots + s];
        if (per_root) per_root[s] = hss_root_result{h[s], h[n_roots + s], h[2 * (size_t)n_roots + s]};
    }
    c->stats.setup_ms = R.setup_ms;
    c->stats.cells_ms = R.cells_ms;
    c->stats.exact_ms = R.exact_ms;
    c->stats.tiles = B.plan.tiles.size() + B.plan.wtiles.size();
    c->stats.window_tiles = B.plan.wtiles.size();
    c->stats.evaluated = R.evaluated;
    c->stats.graph_launch = R.graph ? 1 : 0;
    c->stats.window_evaluated = R.win_evaluated;
    c->stats.window_candidates = R.win_cand;
    c->stats.plain_candidates = R.plain_cand;
    c->stats.window_ms = R.win_ms;
    c->stats.plain_ms = R.plain_ms;
    c->stats.sorted_lefts = B.plan.sorted.size();
    c->stats.left_entries = B.plan.left_entries;
    c->stats.deferred = R.deferred;
    c->stats.raw_roots = R.raw;
    c->stats.grid_blocks = c->grid_blocks();
    c->stats.block_threads = 256;
    c->stats.candidates = total;
    c->stats.leaves = leaves;
    if (candidates_total) *candidates_total = total;
    std::set<std::tuple<uint64_t, uint64_t, uint64_t>> seen;
    std::vector<HitOut> uniq;
    for (const HitOut& x : R.hits)
        if (seen.insert({x.l, x.m, x.n}).second) uniq.push_back(x);
    std::sort(uniq.begin(), uniq.end(), [](const HitOut& a, const HitOut& b) {
        return std::tie(a.l, a.m, a.n) < std::tie(b.l, b.m, b.n);
    });
    uint32_t k = (uint32_t)uniq.size();
    for (uint32_t i = 0; i < k && i < hit_cap; ++i)
        hits[i] = hss_hit{uniq[i].l, uniq[i].m, uniq[i].n, uniq[i].area4_lo, uniq[i].area4_hi};
    if (n_hits) *n_hits = k;
    c->stats.total_ms = ms_since(t0);
    if (k > hit_cap) return fail(HSS_EOVERFLOW, "more hits than hit_cap");
    return HSS_OK;
}

extern "C" int hss_search_interval(hss_ctx* c, uint64_t limit, uint64_t first, uint64_t last,
                                   uint64_t* candidates_total, hss_hit* hits, uint32_t hit_cap, uint32_t* n_hits) {
    if (first < 1 || first > last || last - first >= (1ull << 31)) return fail(HSS_EINVAL, "bad interval");
    std::vector<uint64_t> roots(last - first + 1);
    for (uint64_t i = 0; i < roots.size(); ++i) roots[i] = first + i;
    return hss_search(c, limit, roots.data(), (uint32_t)roots.size(), nullptr, candidates_total, hits, hit_cap,
                      n_hits);
}

namespace {

// Point every device buffer of B into the scratch array, in order, each on a
// 256-byte boundary; false when the batch does not fit.  The views are never
// released (clear_views drops them before B.release()).
bool carve_views(hss_ctx* c, LaunchBuffers& B, uint32_t n_roots, Counters& cnt, size_t& need) {
    const Plan& P = B.plan;
    char* const base = static_cast<char*>(c->scratch);
    size_t used = 0;
    bool fits = true;
    auto take = [&](size_t bytes) -> char* {
        const size_t at = (used + 255) & ~(size_t)255;
        used = at + std::max<size_t>(bytes, 1);
        if (used > NOALLOC_SCRATCH_BYTES) fits = false;
        return fits ? base + at : nullptr;
    };
    auto view = [&](DevBuf& d, size_t bytes) {
        d.p = take(bytes);
        d.bytes = d.p ? bytes : 0;
    };
    const size_t entries = std::max<uint64_t>(P.left_entries, 1);
    view(B.header, sizeof(BatchHeader));
    view(B.misc, 64);
    view(B.pos, (size_t)n_roots * sizeof(PosDesc));
    view(B.lefts, P.lefts.size() * sizeof(LeftDesc));
    view(B.builds, P.lefts.size() * sizeof(LeftBuild));
    view(B.sorted, P.sorted.size() * 4);
    view(B.btab, P.sorted.size() * (WIN_NB + 1) * 4);
    view(B.perm, entries * 4);
    view(B.lfast, (entries + WIN_UNIT) * sizeof(LeftFast));  // padding read by window units
    view(B.lre, entries * sizeof(i128));
    view(B.lim, entries * sizeof(i128));
    view(B.tiles, P.tiles.size() * sizeof(TileDesc));
    view(B.wtiles, P.wtiles.size() * sizeof(TileDesc));
    view(B.raw, LaunchBuffers::RAW_CAP * sizeof(HitOut));
    view(B.hits, LaunchBuffers::HIT_CAP * sizeof(HitOut));
    B.def_cap = NOALLOC_DEF_CAP;
    view(B.def, (size_t)B.def_cap * sizeof(DefItem));
    cnt.cand = reinterpret_cast<unsigned long long*>(take((size_t)n_roots * 3 * 8));
    cnt.leaves = cnt.cand ? cnt.cand + n_roots : nullptr;
    cnt.pairs = cnt.cand ? cnt.cand + 2 * (size_t)n_roots : nullptr;
    need = used;
    return fits;
}

}  // namespace

// The search of hss_search with no device or pinned-host allocation: the
// plan is copied straight from host memory into views of the scratch array,
// the kernel chain is launched directly (no graph capture) and the deferred
// buffer has a fixed capacity.
extern "C" int hss_search_noalloc(hss_ctx* c, uint64_t limit, const uint64_t* roots, uint32_t n_roots,
                                  hss_root_result* per_root, uint64_t* candidates_total, hss_hit* hits,
                                  uint32_t hit_cap, uint32_t* n_hits) {
    if (!c) return fail(HSS_EINVAL, "null context");
    if (c->server) return fail(HSS_EINVAL, "context is attached to a server");
    if (n_roots && !roots) return fail(HSS_EINVAL, "roots is null");
    if (hit_cap && !hits) return fail(HSS_EINVAL, "hits is null");
    if (!c->scratch) return fail(HSS_EDEVICE, "no scratch array on this device");
    if (hipSetDevice(c->device) != hipSuccess) return fail(HSS_EDEVICE, "hipSetDevice failed");
    c->stats = hss_stats{};
    if (n_hits) *n_hits = 0;
    if (candidates_total) *candidates_total = 0;
    if (!n_roots) return HSS_OK;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<PosReq> reqs(n_roots);
    for (uint32_t s = 0; s < n_roots; ++s) reqs[s] = PosReq{roots[s], limit, s};
    LaunchBuffers& B = c->noalloc;
    int rc = launch_plan(c, B, reqs.data(), n_roots);
    if (rc != HSS_OK) return rc;
    c->stats.plan_ms = ms_since(t0);
    Counters cnt;
    size_t need = 0;
    if (!carve_views(c, B, n_roots, cnt, need)) {
        clear_views(B);
        return fail(HSS_ENOMEM, "batch needs " + std::to_string(need) + " bytes of scratch; the no-allocation "
                                "search has " + std::to_string(NOALLOC_SCRATCH_BYTES) + ": split the batch");
    }
    struct Drop {
        LaunchBuffers& b;
        ~Drop() { clear_views(b); }
    } drop{B};
    const Plan& P = B.plan;
    const hipStream_t st = c->stream;
    const BatchHeader hdr = batch_header(B, true);
    HIP_TRY(hipEventRecord(B.ev[0], st));
    HIP_TRY(hipMemcpyAsync(B.header.p, &hdr, sizeof hdr, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(B.pos.p, P.pos.data(), n_roots * sizeof(PosDesc), hipMemcpyHostToDevice, st));
    if (!P.lefts.empty()) {
        HIP_TRY(hipMemcpyAsync(B.lefts.p, P.lefts.data(), P.lefts.size() * sizeof(LeftDesc), hipMemcpyHostToDevice,
                               st));
        HIP_TRY(hipMemcpyAsync(B.builds.p, P.builds.data(), P.builds.size() * sizeof(LeftBuild),
                               hipMemcpyHostToDevice, st));
    }
    if (!P.sorted.empty())
        HIP_TRY(hipMemcpyAsync(B.sorted.p, P.sorted.data(), P.sorted.size() * 4, hipMemcpyHostToDevice, st));
    if (!P.tiles.empty())
        HIP_TRY(hipMemcpyAsync(B.tiles.p, P.tiles.data(), P.tiles.size() * sizeof(TileDesc), hipMemcpyHostToDevice,
                               st));
    if (!P.wtiles.empty())
        HIP_TRY(hipMemcpyAsync(B.wtiles.p, P.wtiles.data(), P.wtiles.size() * sizeof(TileDesc),
                               hipMemcpyHostToDevice, st));
    const CellParams K = cell_params(c, B, cnt);
    for (int stage = 0; stage < LaunchBuffers::STAGES; ++stage) {
        const bool empty = (stage == 1 && P.wtiles.empty()) || (stage == 2 && P.tiles.empty());
        if (!empty) {
            rc = enqueue_stage(c, B, K, stage, st);
            if (rc != HSS_OK) return rc;
        }
        HIP_TRY(hipEventRecord(B.ev[1 + stage], st));
        if (stage == 0) HIP_TRY(hipEventRecord(B.ev[6], st));
    }
    B.replayed = false;
    HIP_TRY(hipMemcpyAsync(B.h_misc.p, B.misc.p, 64, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipEventRecord(B.ev[5], st));
    HIP_TRY(hipEventSynchronize(B.ev[5]));
    LaunchResult R;
    launch_read_misc(B, R);
    const uint32_t* m4 = B.h_misc.as<uint32_t>();
    if (m4[1] > B.def_cap)
        return fail(HSS_ENOMEM, "deferred-leaf buffer of the no-allocation search overflowed (" +
                                    std::to_string(m4[1]) + " > " + std::to_string(B.def_cap) + "): split the batch");
    if (m4[2] > LaunchBuffers::RAW_CAP) return fail(HSS_EOVERFLOW, "raw fourth-power buffer overflow");
    if (m4[3] > LaunchBuffers::HIT_CAP) return fail(HSS_EOVERFLOW, "accepted-hit buffer overflow");
    R.hits.resize(m4[3]);
    if (m4[3]) HIP_TRY(hipMemcpyAsync(R.hits.data(), B.hits.p, m4[3] * sizeof(HitOut), hipMemcpyDeviceToHost, st));
    std::vector<unsigned long long> h((size_t)n_roots * 3);
    HIP_TRY(hipMemcpyAsync(h.data(), cnt.cand, h.size() * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return finish_search(c, B, R, h, n_roots, per_root, candidates_total, hits, hit_cap, n_hits, t0);
}

extern "C" uint32_t hss_gaussian_weight(uint64_t l) {
    static const std::vector<uint32_t> primes = primes_upto(1u << 13);  // covers l <= 2^26
    if (!l || l > (1ull << 26)) return 0;
    return gaussian_weight(l, primes);
}

extern "C" int hss_verify_triple(uint64_t l, uint64_t m, uint64_t n, uint64_t area_lo, uint64_t area_hi) {
    u128 area = ((u128)area_hi << 64) | area_lo;
    if (area >> 124) return 0;
    if (!l || !m || !n || gcd_u64(gcd_u64(l, m), n) != 1) return 0;
    return heron_matches(l, m, n, area * 4) ? 1 : 0;
}
