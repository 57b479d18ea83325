/*
 * hss.h -- C ABI of the MI355X (gfx950) Gaussian-integer enumeration core for
 * the square-sided primitive Heronian triangle search.
 *
 * The reference has no in-process FFI: its replaceable unit is the
 * gpu_search_v42 executable, whose host loop (src/gpu_search_v42.cu:393-506)
 * prepares number-theory tables once, then launches search_kernel on batches of
 * at most 512 root positions and reads back a candidate counter and hits.
 * These entry points are that boundary, made callable: plain pointers and
 * integers, no C++ or torch types.  The executables gpu_search / search in
 * heronian-square-search_amd/ are thin C++ drivers over this ABI; a ctypes
 * binding is heronian-square-search_amd/hss.py.
 *
 * Ownership: a context owns all device memory; callers own every host buffer
 * they pass.  Errors: every function returns an int status (HSS_OK = 0) and
 * never aborts; hss_last_error() gives a thread-local message.  Threading: a
 * context is driven by one host thread at a time.
 */
#ifndef HSS_H
#define HSS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HSS_OK 0
#define HSS_EINVAL 1    /* bad argument (range, limit above the context's N_max, ...) */
#define HSS_ENOMEM 2    /* device or host allocation failed */
#define HSS_EDEVICE 3   /* HIP runtime error or kernel failure */
#define HSS_EOVERFLOW 4 /* more hits than the caller's buffer holds (count still reported) */
#define HSS_EAGAIN 5    /* a served request is not finished (wait timed out) */

typedef struct hss_ctx hss_ctx;

/* One accepted primitive triple: roots l <= m <= n, |4A| as two 64-bit halves. */
typedef struct {
    uint64_t l, m, n;
    uint64_t area4_lo, area4_hi;
} hss_hit;

/* Per-root result of a search call. */
typedef struct {
    uint64_t candidates; /* entries of the fourth-power test (reference counter) */
    uint64_t leaves;     /* Gaussian representations enumerated */
    uint64_t pairs;      /* (l, m) pairs that reached the enumeration */
} hss_root_result;

/* Timings (milliseconds, HIP events on the context's stream) and work of the
 * most recent search call. */
typedef struct {
    double plan_ms;    /* host: factor roots, signatures, tiles */
    double setup_ms;   /* uploads + left-set build kernel */
    double cells_ms;   /* cell_kernel (the enumeration) */
    double exact_ms;   /* exact kernel + acceptance kernel */
    double total_ms;   /* whole call, wall clock */
    uint64_t tiles;
    uint64_t left_entries;
    uint64_t deferred;  /* leaves settled by the exact kernel */
    uint64_t raw_roots; /* exact fourth powers before the acceptance rules */
    uint64_t leaves;
    uint64_t candidates;
    uint32_t grid_blocks;
    uint32_t block_threads;
    uint64_t window_tiles; /* tiles run by the window-mode kernel (range windows) */
    uint64_t sorted_lefts; /* left sets stored sorted by angle for it */
    uint64_t evaluated;    /* leaf orientations the fast path tested (the class rule and the
                              range windows skip the rest) */
    uint64_t window_evaluated; /* ... of them by the window-mode kernel */
    double window_ms;      /* the window-mode kernel's part of cells_ms */
    double plain_ms;       /* the plain kernel's part of cells_ms */
    uint64_t graph_launch; /* 1 when the kernel chain ran as a replay of its captured hipGraph */
    uint64_t window_candidates; /* candidates counted by the window-mode kernel */
    uint64_t plain_candidates;  /* ... by the plain kernel (the exact stage counts the rest) */
} hss_stats;

/*
 * Create a context for every LIMIT <= n_max (n_max <= 2^26) on HIP device
 * `device`: on-device smallest-prime-factor sieve, Gaussian splits of the
 * primes p = 1 (mod 4), and the half-set cache of every integer <= n_max, all
 * resident in HBM.  Replaces Search::prepare and the table upload of
 * src/gpu_search_v42.cu:354-432 (and src/search_v18.cpp:159-197, 285-309).
 */
int hss_ctx_create(uint64_t n_max, int device, hss_ctx** out);

/* Release a context (src/gpu_search_v42.cu:496-505); an attached server is
 * destroyed first (hss_server_destroy). */
void hss_ctx_destroy(hss_ctx* ctx);

/* Thread-local description of the last failure in this thread. */
const char* hss_last_error(void);

/* Largest LIMIT the context serves, and its device memory footprint in bytes. */
uint64_t hss_ctx_nmax(const hss_ctx* ctx);
uint64_t hss_ctx_device_bytes(const hss_ctx* ctx);

/*
 * Search root positions roots[0..n_roots) at LIMIT `limit` (<= n_max) in one
 * persistent launch: every pair (l, m), l <= m <= limit, is enumerated exactly
 * as the reference's per-pair loop (src/search_v18.cpp:329-479) and the
 * reference's candidate counter is reproduced bit for bit.  Roots may repeat
 * and come in any order.  per_root (nullable, n_roots entries) receives each
 * root's counters; *candidates_total (nullable) their sum.  Accepted hits
 * (de-duplicated on (l, m, n), sorted) go to hits[0..hit_cap) and their number
 * to *n_hits; HSS_EOVERFLOW when it exceeds hit_cap.  Replaces one
 * search_kernel launch + synchronize + read-back of src/gpu_search_v42.cu:457-481,
 * generalised from a contiguous batch to any list of roots.
 */
int hss_search(hss_ctx* ctx, uint64_t limit, const uint64_t* roots, uint32_t n_roots,
               hss_root_result* per_root, uint64_t* candidates_total, hss_hit* hits,
               uint32_t hit_cap, uint32_t* n_hits);

/* hss_search over the contiguous roots [first, last] (one reference batch). */
int hss_search_interval(hss_ctx* ctx, uint64_t limit, uint64_t first, uint64_t last,
                        uint64_t* candidates_total, hss_hit* hits, uint32_t hit_cap,
                        uint32_t* n_hits);

/*
 * hss_search without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory during a call (code under graph
 * capture, tight serving loops, harnesses that forbid it).  Same arguments,
 * results and errors as hss_search, but every per-call device buffer (plan,
 * left sets, tiles, counters, deferred leaves, hits) lives in one static
 * 256 MiB array per device, whose address the context looks up when it is
 * created; the call allocates no device or pinned host memory and launches
 * the kernels directly (no graph).  A batch whose buffers exceed the array, or
 * whose deferred leaves exceed its fixed capacity (2^20), fails with
 * HSS_ENOMEM and a message to split it; a batch of 512 root positions near
 * 7.5e6 at LIMIT 1e7 needs about 70 MiB.  Every call on a device shares that
 * device's array: calls that overlap on one device (two contexts, two
 * threads) race and must be serialised by the caller.
 */
int hss_search_noalloc(hss_ctx* ctx, uint64_t limit, const uint64_t* roots, uint32_t n_roots,
                       hss_root_result* per_root, uint64_t* candidates_total, hss_hit* hits,
                       uint32_t hit_cap, uint32_t* n_hits);

/*
 * Scheduling knobs (no counterpart in the reference, whose kernel has one
 * fixed launch shape, src/gpu_search_v42.cu:457-458), keyed on Gaussian
 * weight.  A tile -- one left set of a root against a range of partners --
 * runs in window mode (only the left entries whose leaves can pass the range
 * test, found from angle-sorted left sets) when the mean range-window fraction
 * of its partners is at most win_max_frac[c], c the weight class of the left
 * set: H = (w + 1) / 2 entries for its Gaussian weight w, c = floor(log2 H) - 3
 * for 9 <= H <= 4097 (2^(c+3) <= H < 2^(c+4); class 8 up to 4097); other left
 * sets, and classes with 0, evaluate every leaf.  Counts never depend on these
 * values; only speed does.
 */
#define HSS_WEIGHT_CLASSES 9

typedef struct {
    uint32_t tiles_per_wave; /* tile size: estimated work / (resident waves * this) */
    uint32_t reserved;
    double win_max_frac[HSS_WEIGHT_CLASSES];
    double win_item_cost;    /* planning cost of a window-mode leaf vs a plain leaf */
    double win_partner_cost; /* planning cost per partner entry of a window-mode tile */
    double plain_partner_cost; /* planning cost per partner entry of a plain tile (cell units) */
} hss_tuning;

/*
 * Graph replay of the kernel chain (left-set sort and build, counter reset,
 * window and plain cell kernels, exact and acceptance kernels): captured once
 * per buffer layout and replayed for every batch, sizes read from a device
 * header (enable = 1, the default; HSS_NO_GRAPH in the environment starts
 * with 0).  Returns the previous setting.
 */
int hss_ctx_use_graphs(hss_ctx* ctx, int enable);

int hss_ctx_get_tuning(const hss_ctx* ctx, hss_tuning* out);
int hss_ctx_set_tuning(hss_ctx* ctx, const hss_tuning* in);

/*
 * Autotuner: searches win_max_frac per weight class (coordinate descent over
 * {0, 0.05, 0.1, 0.2, 0.3, 0.45, 0.6, 0.8, 1}, classes in order of their share
 * of the work) and then tiles_per_wave over {8, 16, 32, 48, 64}, timing the device
 * part of hss_search (HIP events, best of `repeats`) on the calibration roots
 * at `limit`.  Every setting must give the same candidate total (else
 * HSS_EDEVICE).  The best tuning is applied to ctx and copied to *out
 * (nullable); *best_ms (nullable) receives its time.
 */
int hss_autotune(hss_ctx* ctx, uint64_t limit, const uint64_t* roots, uint32_t n_roots, uint32_t repeats,
                 hss_tuning* out, double* best_ms);

/* Timings and work counters of the last hss_search / hss_search_interval call. */
int hss_get_stats(const hss_ctx* ctx, hss_stats* out);

/*
 * Host-side exact check of a result line "l m n area" (area = |4A|/4): the
 * counterpart of scripts/verify.py:6-22 and of the 256-bit recheck in
 * src/search_v18.cpp:481-507.  Returns 1 when (l^2, m^2, n^2) is a
 * nondegenerate primitive triangle of that integer area, else 0.
 */
int hss_verify_triple(uint64_t l, uint64_t m, uint64_t n, uint64_t area_lo, uint64_t area_hi);

/*
 * Gaussian weight of a root side: w(l) = prod (4e + 1) over the prime powers
 * p^e || l with p = 1 (mod 4) (experiments/heavy/README.md:9; branch_weight of
 * experiments/heavy/gpu_search_v36.cu:51-55).  Host arithmetic, l <= 2^26;
 * 0 for l = 0 or l > 2^26.
 */
uint32_t hss_gaussian_weight(uint64_t l);

/*
 * Work estimate of the root positions [first, last] at `limit`: the planner's
 * count of (left entry, partner half-set entry) cells, the unit the server's
 * launch budget and the tile sizes are measured in.  Host arithmetic only.
 */
int hss_estimate(const hss_ctx* ctx, uint64_t limit, uint64_t first, uint64_t last, double* cells);

/*
 * Split the root positions [first, last] at `limit` into `parts` contiguous
 * intervals of about equal estimated work (hss_estimate's cells): interval i
 * is [bounds[i], bounds[i+1] - 1], bounds[0] = first, bounds[parts] = last + 1;
 * bounds[i] is one past the first position whose inclusive prefix of work
 * reaches i/parts of the total.  `bounds` holds parts + 1 values.  Replaces
 * the reference supervisor's fixed 12500-position interval per GPU
 * (scripts/gpu_queue.py:224-226) and its idle-GPU rebalancing
 * (scripts/gpu_rebalance.py:65-108) with a split by work.  Host arithmetic.
 */
int hss_split_by_work(const hss_ctx* ctx, uint64_t limit, uint64_t first, uint64_t last, uint32_t parts,
                      uint64_t* bounds);

/* ------------------------------------------------------------ served form
 *
 * A server drives one context from its own thread.  Clients on any thread
 * submit interval requests (LIMIT, FIRST_L, LAST_L) -- each is what one
 * gpu_search_v42 process did for its interval (src/gpu_search_v42.cu:610-632)
 * -- and wait for them.  Requests are admitted in arrival order while a
 * request slot and room in the device state pool (24 bytes per root position:
 * the request's per-position counters, resident on the device until it is
 * done) are free; every launch takes work from all admitted requests, so
 * requests that arrive while others run join the next launch and share the
 * device.  Each position's counts and hits are the same as a search of that
 * position alone.  The reference has no served form: its supervisor
 * (scripts/gpu_queue.py:65-98) runs one process per interval and GPU.
 */
typedef struct hss_server hss_server;

typedef struct {
    uint32_t max_slots;   /* requests admitted at once; 0: 256 */
    uint32_t reserved;
    uint64_t state_bytes; /* device state pool, 24 bytes per root position; 0: 64 MiB */
    double launch_work;   /* work per launch in planner cells (hss_estimate); 0: 3e10 */
} hss_server_config;

#define HSS_REQ_QUEUED 0 /* waiting for a slot and state memory */
#define HSS_REQ_ACTIVE 1 /* admitted: positions are being searched */
#define HSS_REQ_DONE 2
#define HSS_REQ_FAILED 3
#define HSS_REQ_CANCELED 4 /* stopped by hss_server_cancel: [first, resume) finished */

typedef struct {
    uint64_t id;
    uint64_t limit, first, last;
    uint64_t resume;     /* root sides [first, resume) are finished (a STATE file's resume) */
    uint64_t candidates; /* of the finished prefix [first, resume); the totals once done */
    uint64_t leaves;
    uint64_t pairs;
    uint32_t n_hits;     /* hits found so far (raw count; unique ones via hss_server_results) */
    uint32_t launches;   /* launches the request's positions were part of */
    int32_t state;       /* HSS_REQ_* */
    int32_t error;       /* status of a failed request */
    /* server clock, milliseconds since hss_server_create; -1 until reached */
    double submit_ms, admit_ms, start_ms, done_ms;
} hss_request_info;

typedef struct {
    uint64_t submitted, completed, failed;
    uint64_t launches, reruns;
    uint64_t shared_launches; /* launches that carried positions of two or more requests */
    uint64_t positions; /* root positions of completed requests */
    uint64_t leaves, candidates;
    double gpu_ms;      /* sum of launch durations (HIP events) */
    double uptime_ms;
    uint32_t queued, active, max_active;
    uint32_t pool_positions, pool_free;
    uint32_t reserved;
} hss_server_stats;

/* Attach a server to ctx (one per context; hss_search on ctx is refused while
 * it is attached). */
int hss_server_create(hss_ctx* ctx, const hss_server_config* cfg, hss_server** out);

/* Queue the request [first, last] at limit (first >= 1, last <= limit <= the
 * context's n_max, at most the pool's positions); *id receives its handle. */
int hss_server_submit(hss_server* srv, uint64_t limit, uint64_t first, uint64_t last, uint64_t* id);

/* Wait up to timeout_ms (< 0: no limit) for a request; fills *info (nullable)
 * either way.  HSS_OK when done, HSS_EAGAIN when still running, the request's
 * error when it failed. */
int hss_server_wait(hss_server* srv, uint64_t id, int64_t timeout_ms, hss_request_info* info);

/* Results of a request so far: per-position counters (cap entries, last-first+1
 * needed; zero for positions not finished yet) and the hits found so far
 * (de-duplicated, sorted); HSS_EOVERFLOW when a buffer is short (counts still
 * reported).  A writer of the reference's files appends these hits before it
 * writes a STATE resume taken from hss_server_wait. */
int hss_server_results(hss_server* srv, uint64_t id, hss_root_result* per_root, uint64_t cap,
                       hss_hit* hits, uint32_t hit_cap, uint32_t* n_hits);

/* Forget a finished request. */
int hss_server_release(hss_server* srv, uint64_t id);

/* Stop a request (the SIGTERM path of the reference's host loop,
 * src/gpu_search_v42.cu:460-494): positions not yet handed to a launch are
 * dropped; it ends HSS_REQ_CANCELED once the launches holding its other
 * positions finish, with resume = the end of its finished prefix. */
int hss_server_cancel(hss_server* srv, uint64_t id);

int hss_server_get_stats(hss_server* srv, hss_server_stats* out);

/* Finish every submitted request, stop the server thread and detach. */
void hss_server_destroy(hss_server* srv);

#ifdef __cplusplus
}
#endif
#endif
