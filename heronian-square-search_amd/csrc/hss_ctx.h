// hss_ctx.h -- internal state of a context and the per-launch buffer sets
// shared by the synchronous search (hss_runtime.cpp) and the request server
// (hss_server.cpp).  Host code only; no part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/hss.h"
#include "hss_kernels.h"
#include "hss_plan.h"

namespace hss {

// thread-local error text behind hss_last_error()
int fail(int code, const std::string& msg);

#define HSS_HIP_TRY(x)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return ::hss::fail(e_ == hipErrorOutOfMemory ? HSS_ENOMEM : HSS_EDEVICE,       \
                               std::string(#x) + ": " + hipGetErrorString(e_));            \
    } while (0)

// grow-only device buffer
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t want);
    void release();
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// grow-only pinned host buffer (allocated pinned: no registration of existing memory)
struct HostBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t want);
    void release();
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// Counter arrays of one launch target: cand[i], leaves[i], pairs[i] for every
// counter index i < cap (PosDesc::out).
struct Counters {
    unsigned long long* cand = nullptr;
    unsigned long long* leaves = nullptr;
    unsigned long long* pairs = nullptr;
};

// Everything one launch of the search needs besides the resident tables.  A
// buffer set is reused only after its launch has completed.
struct LaunchBuffers {
    DevBuf header, pos, lefts, builds, tiles, wtiles, sorted, perm, btab, lfast, lre, lim, misc, def, raw, hits;
    HostBuf upload;   // pinned staging of the plan
    HostBuf upload2;  // pinned staging of the tiles when they are uploaded after the setup stage
    HostBuf h_misc;   // pinned copy of misc: [1] deferred, [2] raw, [3] hits, [6..13] per-kernel totals (4 x u64)
    uint32_t def_cap = 0;
    static constexpr int STAGES = 4;
    // start, setup done, window cells done, plain cells done, exact done, misc
    // read back, cells start (after the tile upload when it follows the setup)
    hipEvent_t ev[7] = {};
    Plan plan;
    uint32_t npos = 0;
    // The kernel chain (left-set sort and build with the counter reset; the
    // window cell kernel; the plain cell kernel; the exact and acceptance
    // kernels) captured as STAGES graphs; valid while the buffers' addresses
    // and def_cap match `layout`.
    hipGraph_t graph[STAGES] = {};
    hipGraphExec_t exec[STAGES] = {};
    std::vector<uintptr_t> layout;
    bool replayed = false;  // the last launch replayed the graph
    static constexpr uint32_t RAW_CAP = 1u << 16, HIT_CAP = 1u << 16;
    hipError_t create();
    void release();
    void drop_graph();
};

// Outcome of a completed launch.
struct LaunchResult {
    bool rerun = false;        // deferred buffer overflowed: grown, launch again
    uint32_t deferred = 0;
    uint32_t raw = 0;
    std::vector<HitOut> hits;  // accepted (not de-duplicated), HitOut::pos = plan position
    bool graph = false;        // the kernel chain ran as a graph replay
    uint64_t evaluated = 0;    // leaves evaluated by the fast path
    uint64_t win_evaluated = 0, plain_evaluated = 0;
    uint64_t win_cand = 0, plain_cand = 0;  // candidates counted by each cell kernel
    float setup_ms = 0, cells_ms = 0, exact_ms = 0;
    float win_ms = 0, plain_ms = 0;  // the two parts of cells_ms
};

}  // namespace hss

struct hss_server;

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
    hss::QR4 qr{};
    // resident tables
    uint32_t* d_off = nullptr;
    double* d_cre = nullptr;
    double* d_cim = nullptr;
    uint32_t* d_cres = nullptr;
    uint32_t* d_vbase = nullptr;
    hss::i128* d_vre = nullptr;
    hss::i128* d_vim = nullptr;
    // synchronous search path
    hss::LaunchBuffers sync;
    hss::DevBuf counters;
    hss_stats stats{};
    hss::Tuning tune{};
    int win_blocks_per_cu = 1;
    bool use_graphs = true;            // capture the kernel chain once, replay it per batch
    std::string graph_error;           // why capture was given up, if it was
    hipStream_t capture = nullptr;     // stream the chain is captured on (never executes work)
    hss_server* server = nullptr;  // a server drives the context while attached
    // no-allocation search (hss_search_noalloc): the device address of this
    // device's scratch array, and launch state whose device buffers are views
    // into it (events and the pinned read-back word are created with the context)
    void* scratch = nullptr;
    hss::LaunchBuffers noalloc;

    uint64_t waves() const { return (uint64_t)cus * blocks_per_cu * 4; }
    uint32_t grid_blocks() const { return (uint32_t)(cus * blocks_per_cu); }
    uint32_t win_grid_blocks() const { return (uint32_t)(cus * win_blocks_per_cu); }
};

namespace hss {

// Plan `n` positions into B.plan (host only).
int launch_plan(hss_ctx* c, LaunchBuffers& B, const PosReq* reqs, uint32_t n);
// Enqueue the planned launch on stream st: uploads, left-set build, reset of
// the positions' counters, cells, exact and acceptance kernels, read-back of
// the misc counters into pinned memory.  Asynchronous.
int launch_enqueue(hss_ctx* c, LaunchBuffers& B, const Counters& cnt, hipStream_t st);
// The same launch in two parts around plan_tiles: launch_plan_lefts, then
// launch_enqueue_setup (uploads the positions and left sets, runs the left-set
// sort and build and the counter reset), launch_plan_tiles (host, while the
// device sorts), launch_enqueue_cells (tiles, cell kernels, exact stage).
int launch_plan_lefts(hss_ctx* c, LaunchBuffers& B, const PosReq* reqs, uint32_t n);
int launch_plan_tiles(hss_ctx* c, LaunchBuffers& B);
int launch_enqueue_setup(hss_ctx* c, LaunchBuffers& B, const Counters& cnt, hipStream_t st);
int launch_enqueue_cells(hss_ctx* c, LaunchBuffers& B, const Counters& cnt, hipStream_t st);
// After B.ev[5] has completed: read the misc counters, detect overflow, copy
// accepted hits on `copy` (a stream with no work that writes B's buffers).
int launch_collect(hss_ctx* c, LaunchBuffers& B, hipStream_t copy, LaunchResult& out);
// The part of launch_collect that allocates nothing: misc counters and timings.
void launch_read_misc(const LaunchBuffers& B, LaunchResult& out);

}  // namespace hss
