// hss_runtime.cpp -- host runtime behind the C ABI of include/hss.h.
//
// One context = one HIP device, one stream (default priority), the resident
// number-theory tables, and grow-only scratch buffers for batches.
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

#include "../../include/hss.h"
#include "hss_kernels.h"
#include "hss_plan.h"

using namespace hss;

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

#define HIP_TRY(x)                                                                         \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return fail(e_ == hipErrorOutOfMemory ? HSS_ENOMEM : HSS_EDEVICE,              \
                        std::string(#x) + ": " + hipGetErrorString(e_));                   \
    } while (0)

// grow-only device buffer
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t want) {
        if (want <= bytes) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        size_t n = std::max<size_t>(want + want / 4, 4096);
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

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

struct hss_ctx {
    int device = 0;
    uint64_t nmax = 0;
    hipStream_t stream = nullptr;
    int cus = 0;
    int blocks_per_cu = 1;
    uint64_t device_bytes = 0;
    std::vector<uint32_t> small_primes;
    std::vector<uint32_t> h_off;  // half-set offsets, nmax + 2 entries
    uint32_t cache_total = 0;
    QR4 qr{};
    // resident tables
    uint32_t* d_off = nullptr;
    double* d_cre = nullptr;
    double* d_cim = nullptr;
    uint32_t* d_cres = nullptr;
    uint32_t* d_vbase = nullptr;
    i128* d_vre = nullptr;
    i128* d_vim = nullptr;
    // batch scratch
    DevBuf pos, lefts, builds, tiles, lfast, lre, lim, counters, misc, def, raw, hits;
    uint32_t def_cap = 0;
    hipEvent_t ev[6] = {};
    hss_stats stats{};
};

extern "C" const char* hss_last_error(void) { return g_error.c_str(); }

extern "C" uint64_t hss_ctx_nmax(const hss_ctx* c) { return c ? c->nmax : 0; }

extern "C" uint64_t hss_ctx_device_bytes(const hss_ctx* c) { return c ? c->device_bytes : 0; }

extern "C" void hss_ctx_destroy(hss_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    (void)hipFree(c->d_off);
    (void)hipFree(c->d_cre);
    (void)hipFree(c->d_cim);
    (void)hipFree(c->d_cres);
    (void)hipFree(c->d_vbase);
    (void)hipFree(c->d_vre);
    (void)hipFree(c->d_vim);
    for (DevBuf* b : {&c->pos, &c->lefts, &c->builds, &c->tiles, &c->lfast, &c->lre, &c->lim, &c->counters,
                      &c->misc, &c->def, &c->raw, &c->hits})
        b->release();
    for (hipEvent_t& e : c->ev)
        if (e) (void)hipEventDestroy(e);
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
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(HSS_EDEVICE, "stream create failed"));
    for (hipEvent_t& e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(fail(HSS_EDEVICE, "event create failed"));
    int bpc = 0;
    if (cell_occupancy(&bpc) != hipSuccess || bpc < 1) bpc = 1;
    c->blocks_per_cu = bpc;
    vlog(nullptr, "stream + occupancy");
    int rc = build_tables(c);
    vlog(c->stream, "tables built");
    if (rc != HSS_OK) return bail(rc);
    *out = c;
    return HSS_OK;
}

extern "C" int hss_get_stats(const hss_ctx* c, hss_stats* out) {
    if (!c || !out) return fail(HSS_EINVAL, "null argument");
    *out = c->stats;
    return HSS_OK;
}

static int run_batch(hss_ctx* c, uint64_t limit, const uint64_t* roots, uint32_t n_roots,
                     std::vector<unsigned long long>& cnt, std::vector<HitOut>& hits_out, uint32_t& raw_count) {
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = c->stream;
    Plan P;
    std::string err;
    const uint64_t waves = (uint64_t)c->cus * c->blocks_per_cu * 4;
    if (!plan_batch(limit, roots, n_roots, c->small_primes, c->h_off.data(), c->nmax, waves, P, err))
        return fail(HSS_EINVAL, err);
    c->stats.plan_ms = ms_since(t0);
    const uint32_t npos = n_roots, nleft = (uint32_t)P.lefts.size(), ntiles = (uint32_t)P.tiles.size();
    HIP_TRY(c->pos.reserve(npos * sizeof(PosDesc)));
    HIP_TRY(c->lefts.reserve(std::max<size_t>(nleft, 1) * sizeof(LeftDesc)));
    HIP_TRY(c->builds.reserve(std::max<size_t>(nleft, 1) * sizeof(LeftBuild)));
    HIP_TRY(c->tiles.reserve(std::max<size_t>(ntiles, 1) * sizeof(TileDesc)));
    HIP_TRY(c->lfast.reserve(std::max<uint64_t>(P.left_entries, 1) * sizeof(LeftFast)));
    HIP_TRY(c->lre.reserve(std::max<uint64_t>(P.left_entries, 1) * sizeof(i128)));
    HIP_TRY(c->lim.reserve(std::max<uint64_t>(P.left_entries, 1) * sizeof(i128)));
    HIP_TRY(c->counters.reserve((size_t)npos * 3 * 8));
    HIP_TRY(c->misc.reserve(64));
    const uint32_t raw_cap = 1u << 16, hit_cap = 1u << 16;
    HIP_TRY(c->raw.reserve(raw_cap * sizeof(HitOut)));
    HIP_TRY(c->hits.reserve(hit_cap * sizeof(HitOut)));
    if (!c->def_cap) {
        c->def_cap = 1u << 22;
        HIP_TRY(c->def.reserve((size_t)c->def_cap * sizeof(DefItem)));
    }
    HIP_TRY(hipEventRecord(c->ev[0], st));
    HIP_TRY(hipMemcpyAsync(c->pos.p, P.pos.data(), npos * sizeof(PosDesc), hipMemcpyHostToDevice, st));
    if (nleft) {
        HIP_TRY(hipMemcpyAsync(c->lefts.p, P.lefts.data(), nleft * sizeof(LeftDesc), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(c->builds.p, P.builds.data(), nleft * sizeof(LeftBuild), hipMemcpyHostToDevice, st));
    }
    if (ntiles)
        HIP_TRY(hipMemcpyAsync(c->tiles.p, P.tiles.data(), ntiles * sizeof(TileDesc), hipMemcpyHostToDevice, st));
    HIP_TRY(launch_build_left(c->builds.as<LeftBuild>(), c->lefts.as<LeftDesc>(), nleft, c->d_vbase, c->d_vre,
                              c->d_vim, c->lre.as<i128>(), c->lim.as<i128>(), c->lfast.as<LeftFast>(), st));
    CellParams K{};
    K.off = c->d_off;
    K.cre = c->d_cre;
    K.cim = c->d_cim;
    K.cres = c->d_cres;
    K.pos = c->pos.as<PosDesc>();
    K.lefts = c->lefts.as<LeftDesc>();
    K.tiles = c->tiles.as<TileDesc>();
    K.ntiles = ntiles;
    K.lfast = c->lfast.as<LeftFast>();
    K.lre = c->lre.as<i128>();
    K.lim = c->lim.as<i128>();
    uint32_t* misc = c->misc.as<uint32_t>();  // [0] next tile, [1] deferred, [2] raw, [3] hits
    K.next_tile = misc;
    K.cand = c->counters.as<unsigned long long>();
    K.leaves = K.cand + npos;
    K.pairs = K.cand + 2 * npos;
    K.ndef = misc + 1;
    K.raw = c->raw.as<HitOut>();
    K.nraw = misc + 2;
    K.raw_cap = raw_cap;
    K.limit = limit;
    double ld = (double)(limit * limit);
    K.n4_d = ld * ld;
    K.qr = c->qr;
    const uint32_t blocks = (uint32_t)(c->cus * c->blocks_per_cu);
    c->stats.grid_blocks = blocks;
    c->stats.block_threads = 256;
    uint32_t ndef = 0;
    for (int attempt = 0; attempt < 4; ++attempt) {
        K.def = c->def.as<DefItem>();
        K.def_cap = c->def_cap;
        HIP_TRY(hipMemsetAsync(c->counters.p, 0, (size_t)npos * 3 * 8, st));
        HIP_TRY(hipMemsetAsync(misc, 0, 16, st));
        HIP_TRY(hipEventRecord(c->ev[1], st));
        vlog(st, "left sets built");
        if (ntiles) HIP_TRY(launch_cells(K, blocks, st));
        vlog(st, "cell kernel");
        HIP_TRY(hipEventRecord(c->ev[2], st));
        HIP_TRY(hipMemcpyAsync(&ndef, misc + 1, 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (ndef <= c->def_cap) break;
        // deferred buffer overflowed: grow it and run the batch again
        if (attempt == 3 || ndef > (1u << 30)) return fail(HSS_ENOMEM, "deferred-leaf buffer overflow");
        c->def_cap = ndef + ndef / 2;
        HIP_TRY(c->def.reserve((size_t)c->def_cap * sizeof(DefItem)));
    }
    HIP_TRY(launch_exact(K, ndef, st));
    vlog(st, "exact kernel");
    HIP_TRY(launch_accept(K.raw, K.nraw, raw_cap, limit, c->hits.as<HitOut>(), misc + 3, hit_cap, st));
    HIP_TRY(hipEventRecord(c->ev[3], st));
    cnt.resize((size_t)npos * 3);
    HIP_TRY(hipMemcpyAsync(cnt.data(), c->counters.p, (size_t)npos * 3 * 8, hipMemcpyDeviceToHost, st));
    uint32_t m4[4];
    HIP_TRY(hipMemcpyAsync(m4, misc, 16, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (m4[2] > raw_cap) return fail(HSS_EOVERFLOW, "raw fourth-power buffer overflow");
    if (m4[3] > hit_cap) return fail(HSS_EOVERFLOW, "accepted-hit buffer overflow");
    raw_count = m4[2];
    hits_out.resize(m4[3]);
    if (m4[3])
        HIP_TRY(hipMemcpy(hits_out.data(), c->hits.p, m4[3] * sizeof(HitOut), hipMemcpyDeviceToHost));
    float a = 0, b = 0, d = 0;
    (void)hipEventElapsedTime(&a, c->ev[0], c->ev[1]);
    (void)hipEventElapsedTime(&b, c->ev[1], c->ev[2]);
    (void)hipEventElapsedTime(&d, c->ev[2], c->ev[3]);
    c->stats.setup_ms = a;
    c->stats.cells_ms = b;
    c->stats.exact_ms = d;
    c->stats.tiles = ntiles;
    c->stats.left_entries = P.left_entries;
    c->stats.deferred = ndef;
    c->stats.raw_roots = raw_count;
    c->stats.total_ms = ms_since(t0);
    return HSS_OK;
}

extern "C" int hss_search(hss_ctx* c, uint64_t limit, const uint64_t* roots, uint32_t n_roots,
                          hss_root_result* per_root, uint64_t* candidates_total, hss_hit* hits, uint32_t hit_cap,
                          uint32_t* n_hits) {
    if (!c) return fail(HSS_EINVAL, "null context");
    if (n_roots && !roots) return fail(HSS_EINVAL, "roots is null");
    if (hit_cap && !hits) return fail(HSS_EINVAL, "hits is null");
    if (hipSetDevice(c->device) != hipSuccess) return fail(HSS_EDEVICE, "hipSetDevice failed");
    c->stats = hss_stats{};
    if (n_hits) *n_hits = 0;
    if (candidates_total) *candidates_total = 0;
    if (!n_roots) return HSS_OK;
    std::vector<unsigned long long> cnt;
    std::vector<HitOut> raw_hits;
    uint32_t raw_count = 0;
    int rc = run_batch(c, limit, roots, n_roots, cnt, raw_hits, raw_count);
    if (rc != HSS_OK) return rc;
    uint64_t total = 0, leaves = 0;
    for (uint32_t s = 0; s < n_roots; ++s) {
        total += cnt[s];
        leaves += cnt[n_roots + s];
        if (per_root) per_root[s] = hss_root_result{cnt[s], cnt[n_roots + s], cnt[2 * n_roots + s]};
    }
    c->stats.candidates = total;
    c->stats.leaves = leaves;
    if (candidates_total) *candidates_total = total;
    std::set<std::tuple<uint64_t, uint64_t, uint64_t>> seen;
    std::vector<HitOut> uniq;
    for (const HitOut& h : raw_hits)
        if (seen.insert({h.l, h.m, h.n}).second) uniq.push_back(h);
    std::sort(uniq.begin(), uniq.end(), [](const HitOut& a, const HitOut& b) {
        return std::tie(a.l, a.m, a.n) < std::tie(b.l, b.m, b.n);
    });
    uint32_t k = (uint32_t)uniq.size();
    for (uint32_t i = 0; i < k && i < hit_cap; ++i)
        hits[i] = hss_hit{uniq[i].l, uniq[i].m, uniq[i].n, uniq[i].area4_lo, uniq[i].area4_hi};
    if (n_hits) *n_hits = k;
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

extern "C" int hss_verify_triple(uint64_t l, uint64_t m, uint64_t n, uint64_t area_lo, uint64_t area_hi) {
    u128 area = ((u128)area_hi << 64) | area_lo;
    if (area >> 124) return 0;
    if (!l || !m || !n || gcd_u64(gcd_u64(l, m), n) != 1) return 0;
    return heron_matches(l, m, n, area * 4) ? 1 : 0;
}
