// hss_kernels.hip -- gfx950 kernels of the Gaussian-integer enumeration core.
//
// Device-resident tables (built once per context for every integer <= N_max):
//   spf[i]            smallest prime factor (on-device sieve)
//   off[i]            start of i's half-set in the Gaussian cache (exclusive scan
//                     of H_i = (w(i)+1)/2, w(i) = prod(4e+1) over p = 1 (mod 4))
//   cre/cim/cres      the cache: for every i, the (w(i)+1)/2 Gaussian integers
//                     fixed(i) * prod pi_p^r conj(pi_p)^(4e_p - r) whose digit
//                     tuples are lexicographically <= the central tuple; FP64
//                     coordinates (exact: |coord| <= i^2 < 2^53), the area class
//                     (the class rule of hss_math.h) and residues mod 180 (QA;
//                     the kernels read only the class).  The last entry of each
//                     half-set is the real one.
//   pval              per split prime p and merged exponent E <= 2*floor(log_p N):
//                     pi^r conj(pi)^(4E-r), r = 0..4E, exact i128
//
// Reference counterparts: Search::prepare, prepare_gaussian_primes,
// flatten_prime_data, prepare_factor_sets (src/gpu_search_v42.cu:354-391,
// 525-607; src/search_v18.cpp:159-197, 285-309).  The reference stores
// free/central/raw lists per integer (3w+1)/2 entries; this cache keeps only
// the half-set (w+1)/2 and takes conjugates on the fly.
//
// The search kernels replace search_kernel (src/gpu_search_v42.cu:292-345):
// persistent grids pull tiles (left set x range of partner indices) from
// device work queues.  In cell_kernel each wavefront maps one partner
// half-set entry v per lane and walks the left set u uniformly (scalar loads),
// evaluating the cells u*v and u*conj(v) with the class rule and the FP64
// band of hss_math.h, which decide every candidate count exactly;
// cell_win_kernel evaluates only the left entries inside each partner's angle
// window (left sets sorted by angle).  In cell_kernel rows with counted terms
// are compacted with ballot/mbcnt into a per-wave LDS ring and root-tested a
// wave of terms at a time (cell_win_kernel tests them inline); the few
// survivors, and the leaves left undecided by the FP64 band, go to a global
// buffer that the exact kernel (k_exact) settles with 128/256-bit integer
// arithmetic.
#include <hip/hip_runtime.h>

#include <mutex>
#include <type_traits>

#include "hss_build.h"
#include "hss_kernels.h"

namespace hss {

// ------------------------------------------------------------- table build

__global__ void k_fill_identity(uint32_t* spf, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) spf[i] = (uint32_t)i;
}

// block b marks the multiples j >= p^2 of the b-th small prime p
__global__ void k_sieve_mark(uint32_t* spf, const uint32_t* primes, uint64_t n) {
    uint64_t p = primes[blockIdx.x];
    for (uint64_t j = p * p + (uint64_t)threadIdx.x * p; j <= n; j += (uint64_t)blockDim.x * p)
        atomicMin(&spf[j], (uint32_t)p);
}

// H_i = (w(i)+1)/2 for i <= n, H_0 = 0 and H_{n+1} = 0 (scan sentinel)
__global__ void k_half_counts(const uint32_t* spf, uint32_t* H, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n + 1) return;
    if (i == 0 || i == n + 1) {
        H[i] = 0;
        return;
    }
    uint32_t x = (uint32_t)i, w = 1;
    while (x > 1) {
        uint32_t p = spf[x], e = 0;
        while (x % p == 0) {
            x /= p;
            ++e;
        }
        if ((p & 3) == 1) w *= 4 * e + 1;
    }
    H[i] = (w + 1) / 2;
}

// exclusive scan of n u32 values, 4096 per block (256 threads x 16)
__global__ void k_scan_blocks(const uint32_t* in, uint32_t* out, uint32_t* block_sums, uint64_t n) {
    __shared__ uint32_t s[256];
    uint64_t base = (uint64_t)blockIdx.x * 4096 + threadIdx.x * 16;
    uint32_t v[16], sum = 0;
    for (int j = 0; j < 16; ++j) {
        v[j] = base + j < n ? in[base + j] : 0;
        sum += v[j];
    }
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        uint32_t t = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = s[threadIdx.x] - sum;
    for (int j = 0; j < 16; ++j) {
        if (base + j < n) out[base + j] = run;
        run += v[j];
    }
    if (threadIdx.x == 255) block_sums[blockIdx.x] = s[255];
}

__global__ void k_scan_sums(uint32_t* sums, uint32_t nblocks) {
    __shared__ uint32_t s[1024];
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nblocks; b0 += 1024) {
        uint32_t i = b0 + threadIdx.x;
        uint32_t v = i < nblocks ? sums[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {
            uint32_t t = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nblocks) sums[i] = carry + s[threadIdx.x] - v;
        uint32_t tot = s[1023];
        __syncthreads();
        carry += tot;
    }
}

__global__ void k_scan_add(uint32_t* out, const uint32_t* sums, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += sums[i / 4096];
}

// collect the split primes p = 1 (mod 4), p <= n (unordered) with their
// value-block sizes
__global__ void k_collect_split(const uint32_t* spf, uint64_t n, uint32_t* plist, uint32_t* psize,
                                uint32_t* count) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t p = 4 * i + 5;
    if (p > n || spf[p] != p) return;
    uint32_t a = 0;
    for (unsigned __int128 q = p; q <= n; q *= p) ++a;
    uint32_t Emax = 2 * a;
    uint32_t slot = atomicAdd(count, 1u);
    plist[slot] = (uint32_t)p;
    psize[slot] = Emax * (2 * Emax + 3);
}

// Gaussian split pi = a + bi of each collected prime by Cornacchia's method
__global__ void k_prime_split(const uint32_t* plist, uint32_t np, const uint32_t* vstart, uint32_t* vbase,
                              int64_t* pa, int64_t* pb) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    uint64_t p = plist[t];
    int64_t a, b;
    prime_split(p, a, b);
    pa[t] = a;
    pb[t] = b;
    vbase[p] = vstart[t];
}

// value v = pi^r conj(pi)^(4E-r) for every (prime, E, r)
__global__ void k_prime_values(const uint32_t* plist, uint32_t np, const uint32_t* vstart, uint32_t total,
                               const int64_t* pa, const int64_t* pb, i128* vre, i128* vim) {
    uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    uint32_t lo = 0, hi = np;  // last t with vstart[t] <= idx
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (vstart[mid] <= idx)
            lo = mid;
        else
            hi = mid;
    }
    uint32_t off = idx - vstart[lo];
    uint32_t E = 1;
    while (exp_offset(E + 1) <= off) ++E;
    uint32_t r = off - exp_offset(E);
    GInt v = gauss_value(pa[lo], pb[lo], E, r);
    vre[idx] = v.re;
    vim[idx] = v.im;
    (void)plist;
}

// one cache entry per thread
__global__ __launch_bounds__(256) void k_fill_cache(const uint32_t* spf, const uint32_t* off, uint64_t n, uint32_t total,
                             const uint32_t* vbase, const i128* vre, const i128* vim, double* cre, double* cim,
                             uint32_t* cres) {
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    uint64_t lo = 0, hi = n + 1;  // last i with off[i] <= e
    while (hi - lo > 1) {
        uint64_t mid = (lo + hi) / 2;
        if (off[mid] <= e)
            lo = mid;
        else
            hi = mid;
    }
    uint32_t i = (uint32_t)lo;
    uint32_t t = e - off[i];
    uint32_t ps[MAX_SPLIT], es[MAX_SPLIT], k = 0;
    int64_t fixed = 1;
    uint32_t x = i;
    while (x > 1) {
        uint32_t p = spf[x], a = 0;
        while (x % p == 0) {
            x /= p;
            ++a;
        }
        if ((p & 3) == 1) {
            ps[k] = p;
            es[k] = a;
            ++k;
        } else {
            for (uint32_t j = 0; j < 2 * a; ++j) fixed *= p;
        }
    }
    uint32_t dig[MAX_SPLIT];
    decode_rank(t, k, es, dig);
    const uint32_t cls = digits_class(k, ps, dig);
    int64_t re = fixed, im = 0;
    for (uint32_t j = 0; j < k; ++j) {
        uint32_t at = vbase[ps[j]] + exp_offset(es[j]) + dig[j];
        int64_t a = (int64_t)vre[at], b = (int64_t)vim[at];
        int64_t nr = re * a - im * b, ni = re * b + im * a;
        re = nr;
        im = ni;
    }
    cre[e] = (double)re;
    cim[e] = (double)im;
    int64_t r1 = re % (int64_t)QA, i1 = im % (int64_t)QA;
    cres[e] = (uint32_t)(r1 < 0 ? r1 + QA : r1) | cls << CRES_CLASS_SHIFT | (uint32_t)(i1 < 0 ? i1 + QA : i1) << 16;
}

// Gaussian value of rank t of a left set: fixed * prod of the merged prime
// values selected by the digits of t (decode_rank order)
__device__ inline GInt left_value(const LeftBuild& b, uint32_t t, const uint32_t* vbase, const i128* vre,
                                  const i128* vim) {
    uint32_t dig[MAX_SPLIT];
    decode_rank(t, b.k, b.E, dig);
    GInt z = {(i128)b.fixed, 0};
    for (uint32_t j = 0; j < b.k; ++j) {
        uint32_t at = vbase[b.p[j]] + exp_offset(b.E[j]) + dig[j];
        z = gmul(z, GInt{vre[at], vim[at]});
    }
    return z;
}

// Class of the entry of rank t of a left set (the class rule of hss_math.h).
__device__ inline uint32_t left_class(const LeftBuild& b, uint32_t t) {
    uint32_t dig[MAX_SPLIT];
    decode_rank(t, b.k, b.E, dig);
    return digits_class(b.k, b.p, dig);
}

// One block per left set (grid-stride over hdr->nleft).  Entry t0 holds rank
// perm[off + t0] for a sorted left set (window mode), rank t0 otherwise.
__global__ __launch_bounds__(256) void k_build_left(const BatchHeader* hdr, const LeftBuild* builds,
                                                    const LeftDesc* lefts, const uint32_t* vbase, const i128* vre,
                                                    const i128* vim, const uint32_t* perm, i128* lre, i128* lim,
                                                    LeftFast* lfast) {
    const uint32_t nleft = hdr->nleft;
    for (uint32_t li = blockIdx.x; li < nleft; li += gridDim.x) {
        const LeftBuild b = builds[li];
        const LeftDesc L = lefts[li];
        for (uint32_t t0 = threadIdx.x; t0 < L.H; t0 += blockDim.x) {
            const uint32_t e = L.off + t0;
            const uint32_t rank = L.bt != NO_TABLE ? perm[e] : t0;
            GInt z = left_value(b, rank, vbase, vre, vim);
            lre[e] = z.re;
            lim[e] = z.im;
            LeftFast f;
            f.re = i128_to_double(z.re);
            f.im = i128_to_double(z.im);
            f.cls = left_class(b, rank);
            f.pad = 0;
            lfast[e] = f;
        }
    }
}

// Angle modulo pi and class of the entry of rank t of a left set, from its
// digits: the value of digit d of prime p_j is pi_j^d conj(pi_j)^(4 E_j - d),
// of angle (2d - 4E_j) theta_j, and the fixed part is a positive integer, so
// the entry's angle is sum_j (2 d_j - 4 E_j) theta_j (theta_j modulo pi is
// enough: the multipliers are even).  FP64 error below 1e-12 rad (at most
// MAX_SPLIT terms with multipliers below 64 and angles below pi); the window
// budget (hss_math.h) allows 1e-7 for the FP32 rounding of the key.
__device__ inline float left_angle_class(const LeftBuild& b, uint32_t t, const double* th, uint32_t& cls) {
    uint32_t dig[MAX_SPLIT];
    decode_rank(t, b.k, b.E, dig);
    double a = 0.0;
    for (uint32_t j = 0; j < b.k; ++j) a = __builtin_fma((double)(2 * (int)dig[j] - 4 * (int)b.E[j]), th[j], a);
    a -= PI_D * __builtin_floor(a * (1.0 / PI_D));
    a = a < 0.0 ? a + PI_D : (a >= PI_D ? a - PI_D : a);
    cls = digits_class(b.k, b.p, dig);
    return (float)a;
}

// Bucket of a sorted left set's entry: its area group's slice of the WIN_NB
// buckets (WIN_NB >> group bits each), then its angle theta modulo pi.
__device__ inline uint32_t sort_bucket(uint32_t group, uint32_t gbits, float key) {
    const uint32_t nbg = WIN_NB >> gbits;
    return group * nbg + theta_bucket((double)key, nbg);
}

// One block per sorted left set (grid-stride over hdr->nsorted): order the
// non-central entries by (area group, theta) -- the group packs the entry's
// class bits that the left set's pairs can constrain (left_kmask), theta is the
// angle modulo pi from the same FP64 coordinates k_build_left stores -- write
// the permutation and the bucket table bt[b] = number of entries with bucket
// < b, b = 0..WIN_NB.  A counting sort on the buckets (whose starts are that
// table) followed by an insertion sort of each bucket by (key, index): entries
// of one left set share their norm, so their angles are distinct and spread
// over the buckets (a few per bucket), and the order equals a full sort by
// (group, key, index).
__global__ __launch_bounds__(256) void k_left_sort(const BatchHeader* hdr, const uint32_t* sorted,
                                                   const LeftBuild* builds, const LeftDesc* lefts,
                                                   const uint32_t* vbase, const i128* vre, const i128* vim,
                                                   uint32_t* perm, uint32_t* btab) {
    __shared__ float key[WIN_MAX_SORT];       // theta_u rounded to FP32 (the window's error budget allows it)
    __shared__ uint16_t order[WIN_MAX_SORT];  // entry indices in sorted order
    __shared__ uint32_t start[WIN_NB + 1];    // bucket counts, then bucket starts
    __shared__ uint32_t cursor[WIN_NB];
    __shared__ uint32_t part[256];
    __shared__ double th[MAX_SPLIT];          // angles of the left set's split primes, modulo pi
    const uint32_t tid = threadIdx.x;
    const uint32_t nsorted = hdr->nsorted;
    for (uint32_t bi = blockIdx.x; bi < nsorted; bi += gridDim.x) {  // uniform per block
        const uint32_t li = sorted[bi];
        const LeftBuild b = builds[li];
        const LeftDesc L = lefts[li];
        const uint32_t kmask = left_kmask(L);
        const uint32_t gbits = __popc(kmask);
        const uint32_t n = L.H - 1;  // sortable entries (L.H <= WIN_MAX_SORT + 1, planned)
        for (uint32_t bk = tid; bk <= WIN_NB; bk += blockDim.x) start[bk] = 0;
        // theta_j from the value of digit 2 E_j + 1, p^(2E_j - 1) pi_j^2: half its angle
        if (tid < b.k) {
            const uint32_t at = vbase[b.p[tid]] + exp_offset(b.E[tid]) + 2 * b.E[tid] + 1;
            th[tid] = 0.5 * atan2(i128_to_double(vim[at]), i128_to_double(vre[at]));
        }
        __syncthreads();
        for (uint32_t t = tid; t < n; t += blockDim.x) {
            uint32_t cls;
            const float k = left_angle_class(b, t, th, cls);
            key[t] = k;
            atomicAdd(&start[sort_bucket(area_group(cls, kmask), gbits, k)], 1u);
        }
        __syncthreads();
        // exclusive scan of the WIN_NB counts, WIN_NB / 256 per thread
        static_assert(WIN_NB == 4 * 256, "scan layout");
        uint32_t c[4], sum = 0;
        for (int j = 0; j < 4; ++j) {
            c[j] = start[4 * tid + j];
            sum += c[j];
        }
        part[tid] = sum;
        __syncthreads();
        for (uint32_t d = 1; d < 256; d <<= 1) {
            const uint32_t v = tid >= d ? part[tid - d] : 0u;
            __syncthreads();
            part[tid] += v;
            __syncthreads();
        }
        uint32_t run = part[tid] - sum;
        for (int j = 0; j < 4; ++j) {
            start[4 * tid + j] = run;
            cursor[4 * tid + j] = run;
            run += c[j];
        }
        if (tid == 0) start[WIN_NB] = n;
        __syncthreads();
        for (uint32_t t = tid; t < n; t += blockDim.x)
            order[atomicAdd(&cursor[sort_bucket(area_group(left_class(b, t), kmask), gbits, key[t])], 1u)] =
                (uint16_t)t;
        __syncthreads();
        for (uint32_t bk = tid; bk < WIN_NB; bk += blockDim.x) {
            const uint32_t s0 = start[bk], s1 = start[bk + 1];
            for (uint32_t i = s0 + 1; i < s1; ++i) {
                const uint16_t x = order[i];
                const float kx = key[x];
                uint32_t j = i;
                for (; j > s0; --j) {
                    const uint16_t y = order[j - 1];
                    const float ky = key[y];
                    if (ky < kx || (ky == kx && y < x)) break;
                    order[j] = y;
                }
                order[j] = x;
            }
        }
        __syncthreads();
        for (uint32_t t = tid; t < n; t += blockDim.x) perm[L.off + t] = order[t];
        if (tid == 0) perm[L.off + n] = n;  // the central entry stays last
        uint32_t* bt = btab + (size_t)L.bt * (WIN_NB + 1);
        for (uint32_t bk = tid; bk <= WIN_NB; bk += blockDim.x) bt[bk] = start[bk];
        __syncthreads();  // the arrays are rewritten by the next left set
    }
}

// ------------------------------------------------------------ cell kernel

static constexpr int WG = 256;
static constexpr int WAVES = WG / 64;
static constexpr int QCAP = 128;  // root queue ring (a power of two)

// Inclusive prefix sum / maximum over the wavefront with DPP: shifts by 1, 2,
// 4, 8 within each row of 16 lanes, then the last lane of rows 0 and 2 into
// rows 1 and 3, and the last lane of row 1 into rows 2 and 3 (lanes a DPP move
// does not reach keep the identity 0).
template <int CTRL, int ROW_MASK>
__device__ inline uint32_t dpp_shr(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, false);
}

__device__ inline uint32_t wave_incl_scan(uint32_t v) {
    v += dpp_shr<0x111, 0xf>(v);  // row_shr:1
    v += dpp_shr<0x112, 0xf>(v);  // row_shr:2
    v += dpp_shr<0x114, 0xf>(v);  // row_shr:4
    v += dpp_shr<0x118, 0xf>(v);  // row_shr:8
    v += dpp_shr<0x142, 0xa>(v);  // row_bcast:15 into rows 1, 3
    v += dpp_shr<0x143, 0xc>(v);  // row_bcast:31 into rows 2, 3
    return v;
}

__device__ inline uint32_t wave_incl_max(uint32_t v) {
    v = max(v, dpp_shr<0x111, 0xf>(v));
    v = max(v, dpp_shr<0x112, 0xf>(v));
    v = max(v, dpp_shr<0x114, 0xf>(v));
    v = max(v, dpp_shr<0x118, 0xf>(v));
    v = max(v, dpp_shr<0x142, 0xa>(v));
    v = max(v, dpp_shr<0x143, 0xc>(v));
    return v;
}

// Owners of flattened items.  Lane j owns items [start_j, start_j + len_j),
// starts ascending with j.  For the step of items [i0, i0 + 64), the owner of
// item i0 + lane is the last lane whose nonempty run starts at or before it:
// starts inside the step are marked in a per-wave LDS array (65 words; slot 64
// takes the lanes without a mark) and a prefix maximum spreads them; carry is
// the owner of item i0 - 1 (uniform, updated to the owner of item i0 + 63).
__device__ inline uint32_t step_owner(uint32_t* mark, uint32_t start, uint32_t len, uint32_t i0, uint32_t& carry) {
    const uint32_t lane = __lane_id();
    mark[lane] = 0;
    __builtin_amdgcn_wave_barrier();
    const uint32_t d = start - i0;
    mark[len != 0 && d < 64 ? d : 64u] = lane + 1;
    __builtin_amdgcn_wave_barrier();
    const uint32_t m = wave_incl_max(mark[lane]);
    __builtin_amdgcn_wave_barrier();
    const uint32_t o = m ? m - 1 : carry;
    carry = __builtin_amdgcn_readlane(o, 63);
    return o;
}

__device__ inline uint64_t wave_sum64(uint64_t v) {
    for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s, 64);
    return v;
}

__device__ inline uint32_t lane_rank(uint64_t mask, uint32_t base = 0u) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, base));
}

__device__ inline bool divides(uint32_t k, uint32_t inv, uint32_t thr) { return k * inv <= thr; }

// Element i of an LDS array of T through a 32-bit byte offset (one 24-bit
// multiply-add): plain indexing with a stride that is not a power of two
// leaves 64-bit address arithmetic in the emitted code.
template <typename T>
__device__ inline T& lds_at(T* base, uint32_t i) {
    return *(T*)((char*)base + __umul24(i, (uint32_t)sizeof(T)));
}

// Append the lanes with pred set to the global deferred buffer.
__device__ inline void push_global(const CellParams& P, bool pred, const DefItem& it) {
    uint64_t mask = __builtin_amdgcn_ballot_w64(pred);
    if (!mask) return;
    // Every lane executes the atomic (lane 0 adds the count, the others 0) and the
    // base is read with readfirstlane: no divergent branch may precede a
    // cross-lane read inside the kernel's loops, or the compiler is free to let
    // the lanes drift into different loop iterations.
    uint32_t got = atomicAdd(P.ndef, __lane_id() == 0 ? (uint32_t)__popcll(mask) : 0u);
    uint32_t base = __builtin_amdgcn_readfirstlane(got);
    uint32_t slot = base + lane_rank(mask);
    if (pred && slot < P.def_cap) P.def[slot] = it;
}

// Per-wave LDS ring of leaves with counted terms awaiting the root-proximity
// test (leaves the FP64 band left undecided go straight to the deferred
// buffer).  Entry: one row of cell_kernel -- D0 = Re(u w) and D1 = Re(u conj(w))
// for the leaves u v and u conj(v) in their orientation (w = v or -i v) -- and
// the packed id (u | partner entry << 17, k | counted terms << 28), the terms as
// bits 0-3 for lm - 2|D0|, lm + 2|D0|, lm - 2|D1|, lm + 2|D1|, never all clear
// (lm = l^4+m^4); bits 2-3 belong to the conjugate leaf.  Slots head .. head +
// n - 1 (mod QCAP) are live.  One 24-byte entry per slot: one address per push
// or pop.
struct QEntry {
    double x, y;
    uint32_t lo, hi;
};

struct RootQueue {
    QEntry* e;
    uint32_t head, n;
    // terms root-tested so far: every counted term of a queued leaf is tested
    // exactly once, so this is their candidate count (a scalar add per round
    // instead of a count per leaf)
    uint64_t tested;
};

struct TileCtx {
    uint32_t left;
    uint32_t q;      // m = q * k
    double l4_d;     // FP64 l^4
    double tol2;     // root-test tolerance of counted second-sign terms (root_tol2)
};


// One round over the oldest min(n, 64) queued leaves, one per lane: the
// lowest remaining term is root-tested (a survivor goes to the deferred
// buffer) and the leaf is queued again while terms remain, so every round
// tests a full wave of terms.  FULL: the queue holds at least 64 entries
// (every lane has one; the lane masks fold away).
template <bool FULL>
__device__ inline void root_round(const CellParams& P, RootQueue& q, const TileCtx& tc) {
    const uint32_t lane = __lane_id();
    const uint32_t count = FULL ? 64u : (q.n < 64u ? q.n : 64u);
    const bool in = FULL || lane < count;
    __builtin_amdgcn_wave_barrier();
    // full rounds keep head a multiple of 64 (the queue starts, and every
    // flush ends, with head 0), so head + lane needs no wrap
    // (full rounds: head is 0 or 64, so the lane's slot address plus 0 or 64 entries)
    const QEntry qe = FULL ? *(const QEntry*)((const char*)&lds_at(q.e, lane) + (q.head ? 64u * sizeof(QEntry) : 0u))
                           : lds_at(q.e, (q.head + lane) & (uint32_t)(QCAP - 1));
    __builtin_amdgcn_wave_barrier();
    q.head = (q.head + count) & (uint32_t)(QCAP - 1);
    q.n -= count;
    q.tested += count;
    const uint32_t k = qe.hi & 0x7FFFFFFu;
    const uint32_t terms = in ? qe.hi >> 28 : 0u;  // every queued entry has a counted term
    const uint32_t t = (uint32_t)__builtin_ctz(terms | 16u);  // 4 on lanes without an entry
    const double lm = lm_of(tc.q, k, tc.l4_d);
    const double a = 2.0 * __builtin_fabs(t < 2u ? qe.x : qe.y);
    const double zz = (t & 1u) ? lm + a : lm - a;
    const bool pass = in & root_near_integer(zz);
    if (__builtin_amdgcn_ballot_w64(pass)) {
        const uint32_t ui = qe.lo & MAX_LEFT_ENTRIES, jj = qe.lo >> ID_U_BITS;
        const uint32_t vidx = pass ? P.off[k] + jj : 0u;
        push_global(P, pass, DefItem{tc.left, ui, vidx, k | (t >= 2u ? DEF_CONJ : 0u)});
    }
    const uint32_t rest = terms & (terms - 1u);  // terms left after the lowest
    const uint64_t again = __builtin_amdgcn_ballot_w64(rest != 0u);
    if (again) {
        if (rest != 0u)
            lds_at(q.e, lane_rank(again, q.head + q.n) & (uint32_t)(QCAP - 1)) =
                QEntry{qe.x, qe.y, qe.lo, (qe.hi & 0x0FFFFFFFu) | rest << 28};
        __builtin_amdgcn_wave_barrier();
        q.n += (uint32_t)__popcll(again);
    }
}

// Append the lanes with pred set to the queue; full waves of entries are
// drained (the queue never holds more than 127 live entries).
__device__ inline void root_push(const CellParams& P, RootQueue& q, const TileCtx& tc, bool pred, double X, double Y,
                                 uint32_t idlo, uint32_t idhi) {
    const uint64_t mask = __builtin_amdgcn_ballot_w64(pred);
    if (!mask) return;
    if (pred) lds_at(q.e, lane_rank(mask, q.head + q.n) & (uint32_t)(QCAP - 1)) = QEntry{X, Y, idlo, idhi};
    q.n += (uint32_t)__popcll(mask);
    while (q.n >= 64) root_round<true>(P, q, tc);
}

__device__ inline void root_flush(const CellParams& P, RootQueue& q, const TileCtx& tc) {
    while (q.n) root_round<false>(P, q, tc);
    q.head = 0;  // empty: restart at slot 0 (full rounds rely on head = 0 or 64)
}

// One partner entry v (lane data).
struct Partner {
    double c, d;
    uint32_t vidx, kj, jj;
    uint32_t cls;           // area class of v (the class rule of hss_math.h)
    bool vc;                // v is the central (real) entry of its half-set
};

// The band of the pair (l, m = q k).
__device__ inline Band pair_band(const LeftDesc& L, const PosDesc& S, uint32_t k) {
    return make_band((double)L.q * (double)k, S.l4_d, S.l2_d, S.n4_d);
}

// The class bits the pair (l, q k) constrains (area_constrained) and whether
// the pair is even (l or m even; q is odd): cm | even << 2.  Whether 3 or 5
// divides m: the divisibility tests k * 3^-1 <= (2^32 - 1) / 3 and
// k * 5^-1 <= (2^32 - 1) / 5 modulo 2^32 (or q's own factor).
__device__ inline uint32_t pair_area_info(const LeftDesc& L, const PosDesc& S, uint32_t k) {
    const bool m3z = L.q3 == 0u || k * 0xAAAAAAABu <= 0x55555555u;
    const bool m5z = L.q5 == 0u || k * 0xCCCCCCCDu <= 0x33333333u;
    const uint32_t cm = area_constrained(S.l3 == 0u || m3z, S.l5 != 0u && !m5z);
    return cm | ((S.l_even || !(k & 1u)) ? 4u : 0u);
}

// Per-k record of the widened bounds (the narrowed ones are w - tau) and the
// pair's area info: one per lane of a partner chunk, shared by the k's
// half-set entries.
struct KCtx {
    double mid, half_w, h2_w, tau, lm;
    uint32_t info;  // pair_area_info | KCTX_TWICE
    uint32_t pad;
};

// An even pair with no constrained class bit (cm = 0): every leaf passes the
// area tests in both orientations, so cell_kernel runs its half-set entries
// twice, once per orientation; every other pair has at most one orientation
// per leaf.
static constexpr uint32_t KCTX_TWICE = 8;

__device__ inline void load_partner(const CellParams& P, bool act, uint32_t kj, uint32_t jj, uint32_t Hj,
                                    uint32_t oj, Partner& v) {
    v.kj = kj;
    v.jj = jj;
    v.vidx = act ? oj + jj : 0u;
    v.vc = jj + 1 == Hj;
    v.c = P.cre[v.vidx];
    v.d = P.cim[v.vidx];
    v.cls = (P.cres[v.vidx] >> CRES_CLASS_SHIFT) & 3u;
}

// The leaf c_u * v of the left set's central (real) entry with partner entry
// v, in its one orientation o: by the class rule orientation 1 (d = c_u Re v)
// counts only when v's class is 0 on the constrained bits cm, orientation 2
// (d = c_u Im v) only on even pairs with class 3 on cm.  With cm != 0 at most
// one of them holds; an even pair with cm = 0 runs its entries twice, force =
// 1 and 2 naming the orientation of each run (force = 0: by class).  A real v
// makes the leaf real: orientation 1 has area 0, orientation 2 has d = 0 and
// counts once when lo <= 0.  Counted terms are root-tested here; an undecided
// leaf goes to the exact stage with its orientation.  Returns whether the leaf
// had an orientation to test.
template <bool ONE>
__device__ inline bool central_win(const CellParams& P, const TileCtx& tc, bool act, double cu, uint32_t ui,
                                   const Partner& v, const Band& b, uint32_t cm, bool even, uint32_t force,
                                   uint32_t& cnt) {
    const uint32_t c = v.cls & cm;
    const uint32_t o = force ? force : (c == 0u ? 1u : ((even && c == cm) ? 2u : 0u));
    const bool ok = act && o != 0u;
    const double D = cu * (o == 2u ? v.d : v.c);
    const bool real2 = ok && o == 2u && v.vc;  // -lo - tau >= 0 <=> lo <= -tau: decided; |lo| <= tau: undecided
    const double h2n = b.h2_w - b.tau;
    bool u;
    const uint32_t f1 = o1_decide<ONE>(ok && !v.vc, D, b, u);
    const bool und = u | (real2 && b.h2_w >= 0.0 && h2n < 0.0);
    if (__builtin_amdgcn_ballot_w64(und)) push_global(P, und, DefItem{tc.left, ui, v.vidx, v.kj | o << 27 | DEF_FULL});
    const uint32_t f = und ? 0u : f1;
    cnt += (uint32_t)__popc(f) + ((!und && real2 && h2n >= 0.0) ? 1u : 0u);
    const double a = __builtin_fabs(D);
    bool pass = false;
    if (__builtin_amdgcn_ballot_w64(f & 1u)) pass = ((f & 1u) != 0u) & root_near_integer(__builtin_fma(-2.0, a, b.lm));
    if (!ONE && __builtin_amdgcn_ballot_w64(f & 2u))
        pass |= ((f & 2u) != 0u) & root_near_integer(__builtin_fma(2.0, a, b.lm), tc.tol2);
    if (__builtin_amdgcn_ballot_w64(pass)) push_global(P, pass, DefItem{tc.left, ui, v.vidx, v.kj});
    return (ok && !v.vc) || real2;  // the leaf had an orientation to test
}

// Partner chunks of a tile: per parity of k, 64 consecutive k of that parity
// (k = kb + 2 lane) are filtered (on_chunk(k) is called with each lane's k
// first and returns how many times k's half-set entries run, 1 or 2) and their
// half-set entries flattened; f is called for every 64 flattened entries with
// the lane's entry (act, kj, jj, Hj; with two runs jj < Hj = twice the
// half-set size), the lane j of its k and the offset oj of k's half-set.  pairs
// and leaves count the pairs that reach the enumeration and their leaves.
template <typename C, typename F>
__device__ inline void for_partner_chunks(const CellParams& P, const TileDesc& T, const LeftDesc& L, const PosDesc& S,
                                          uint32_t* mark, uint64_t& pairs, uint64_t& leaves, C&& on_chunk, F&& f) {
    const uint32_t lane = __lane_id();
    for (uint32_t par = 0; par < 2; ++par) {
        if (S.l_even && par == 0u) continue;  // l and m both even: no pair
        const uint32_t kstart = T.k0 + ((T.k0 & 1u) != par ? 1u : 0u);
        for (uint32_t kb = kstart; kb <= T.k1; kb += 128) {
            const uint32_t k = kb + 2u * lane;
            const uint32_t kq = k <= T.k1 ? k : T.k1;  // in-bounds index for every lane
            bool ok = k <= T.k1 && !(S.l_even && !(k & 1u));
            for (uint32_t j = 0; j < S.nsplit; ++j) ok = ok && !divides(k, S.split_inv[j], S.split_thr[j]);
            for (uint32_t j = 0; j < S.ninert; ++j) ok = ok && !divides(k, S.inert_inv[j], S.inert_thr[j]);
            const uint32_t ok0 = P.off[kq];
            uint32_t Hk = P.off[kq + 1] - ok0;
            Hk = (ok && !((L.flags & LEFT_TRIVIAL) && Hk == 1)) ? Hk : 0u;
            pairs += Hk ? 1 : 0;
            // leaves of the pair: 2H - 1 per non-central entry of k, H for its central one
            leaves += Hk ? (uint64_t)Hk * (2u * L.H - 1u) - (L.H - 1u) : 0ull;
            Hk *= on_chunk(k);
            const uint32_t incl = wave_incl_scan(Hk);
            const uint32_t excl = incl - Hk;
            const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
            uint32_t carry = 0;
            for (uint32_t e0 = 0; e0 < total; e0 += 64) {
                const uint32_t e = e0 + lane;
                const uint32_t j = step_owner(mark, excl, Hk, e0, carry);
                const uint32_t Hj = (uint32_t)__shfl((int)Hk, (int)j, 64);
                const uint32_t xj = (uint32_t)__shfl((int)excl, (int)j, 64);
                const uint32_t oj = (uint32_t)__shfl((int)ok0, (int)j, 64);  // the owner's half-set offset
                const uint32_t jj = e - xj;
                // an entry beyond the owner's run never happens; the test keeps
                // every index in bounds regardless
                const bool act = e < total && jj < Hj;
                f(act, kb + 2u * j, act ? jj : 0u, Hj, j, oj);
            }
        }
    }
}

// Tile bookkeeping shared by both cell kernels: per-position counters and the
// kernel's own totals (kstat[0]: evaluated leaf orientations, kstat[2]:
// candidates).  cand = the lanes' direct counts cnt plus a count kept per
// wave (the root queue's tested terms, or the window units' counted terms).
__device__ inline void tile_done(const CellParams& P, const PosDesc& S, uint32_t cnt, uint64_t tested,
                                 uint64_t leaves, uint64_t pairs, uint64_t evald, unsigned long long* kstat) {
    const uint32_t lane = __lane_id();
    const uint64_t cand = wave_sum64((uint64_t)cnt) + tested;
    leaves = wave_sum64(leaves);
    pairs = wave_sum64(pairs);
    evald = wave_sum64(evald);
    // all lanes execute the atomics; only lane 0 contributes
    atomicAdd(P.cand + S.out, lane == 0 ? (unsigned long long)cand : 0ull);
    atomicAdd(P.leaves + S.out, lane == 0 ? (unsigned long long)leaves : 0ull);
    atomicAdd(P.pairs + S.out, lane == 0 ? (unsigned long long)pairs : 0ull);
    atomicAdd(kstat, lane == 0 ? (unsigned long long)evald : 0ull);
    atomicAdd(kstat + 2, lane == 0 ? (unsigned long long)cand : 0ull);
}

// The batch arrays the cell kernels only read, as separate restrict-qualified
// arguments: uniform reads of them (tiles, left sets, positions, the left
// entries walked by a whole wave) become scalar loads.
#define HSS_CELL_ARGS                                                                                   \
    const CellParams P, const BatchHeader* __restrict__ hdr, const TileDesc* __restrict__ tiles,       \
        const LeftDesc* __restrict__ lefts, const PosDesc* __restrict__ pos,                            \
        const LeftFast* __restrict__ lfast

__global__ __launch_bounds__(WG, 6) void cell_kernel(HSS_CELL_ARGS) {
    __shared__ QEntry s_q[WAVES][QCAP];
    __shared__ uint32_t s_mark[WAVES][65];
    __shared__ KCtx s_kctx[WAVES][64];
    __shared__ uint32_t s_ncls[WAVES][32];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    RootQueue rq{s_q[wave], 0, 0, 0};
    uint32_t* mark = s_mark[wave];
    KCtx* kctx = s_kctx[wave];
    uint32_t* ncls = s_ncls[wave];
    for (;;) {
        // next tile: all lanes execute the atomic (only lane 0 adds 1), see push_global
        const uint32_t t = __builtin_amdgcn_readfirstlane(atomicAdd(P.next_tile, lane == 0 ? 1u : 0u));
        if (t >= hdr->ntiles) break;
        const TileDesc T = tiles[t];
        const LeftDesc L = lefts[T.left];
        const PosDesc& S = pos[L.pos];
        const TileCtx tc{T.left, L.q, S.l4_d, root_tol2(S.n4_d)};
        const LeftFast* lf = lfast + L.off;
        const bool one = __builtin_amdgcn_readfirstlane(tile_one_term(L.q, T.k0, S.l4_d, S.l2_d, S.n4_d) ? 1 : 0) != 0;
        uint32_t cnt = 0;
        uint64_t leaves = 0, pairs = 0, evald = 0;
        // for the count of leaf orientations the rows test: with n(x, msk) the
        // non-central entries u with (class of u ^ x) & msk == 0, ncls[x + 4 msk +
        // 16 both] = n(x, msk) + (both ? n(x ^ 3, msk) : 0) -- the rows of a
        // partner entry of class x test n(x, msk) entries in orientation 1 and,
        // when both orientations can count (both), n(x ^ 3, msk) in orientation 2
        if (L.H > 1) {
            uint64_t h01 = 0, h23 = 0;  // class counts, 32 bits each
            for (uint32_t i = (uint32_t)lane; i + 1 < L.H; i += 64) {
                const uint32_t c = lf[i].cls;
                h01 += c == 0u ? 1ull : (c == 1u ? 1ull << 32 : 0ull);
                h23 += c == 2u ? 1ull : (c == 3u ? 1ull << 32 : 0ull);
            }
            h01 = wave_sum64(h01);
            h23 = wave_sum64(h23);
            const uint32_t msk = ((uint32_t)lane >> 2) & 3u;
            const auto n_of = [&](uint32_t x) {
                uint32_t n = 0;
                n += (x & msk) == 0u ? (uint32_t)h01 : 0u;
                n += ((1u ^ x) & msk) == 0u ? (uint32_t)(h01 >> 32) : 0u;
                n += ((2u ^ x) & msk) == 0u ? (uint32_t)h23 : 0u;
                n += ((3u ^ x) & msk) == 0u ? (uint32_t)(h23 >> 32) : 0u;
                return n;
            };
            const uint32_t x = (uint32_t)lane & 3u;
            const uint32_t n = n_of(x) + (((uint32_t)lane & 16u) ? n_of(x ^ 3u) : 0u);
            __builtin_amdgcn_wave_barrier();
            if (lane < 32) ncls[lane] = n;
            __builtin_amdgcn_wave_barrier();
        }
        // the pair's band and area info once per k (a k's half-set entries share them)
        const auto on_chunk = [&](uint32_t k) {
            const Band c = pair_band(L, S, k);
            const uint32_t info = pair_area_info(L, S, k);
            const bool twice = (info & 3u) == 0u && (info & 4u) != 0u;
            __builtin_amdgcn_wave_barrier();
            lds_at(kctx, lane) = KCtx{c.mid, c.half_w, c.h2_w, c.tau, c.lm, info | (twice ? KCTX_TWICE : 0u), 0u};
            __builtin_amdgcn_wave_barrier();
            return twice ? 2u : 1u;
        };
        for_partner_chunks(P, T, L, S, mark, pairs, leaves, on_chunk,
                           [&](bool act, uint32_t kj, uint32_t jj, uint32_t Hj, uint32_t j, uint32_t oj) {
            const KCtx K = lds_at(kctx, j);
            // a pair run twice: orientation 1 on the first run of its entries, 2 on the second
            const bool twice = (K.info & KCTX_TWICE) != 0u;
            const uint32_t Hv = twice ? Hj >> 1 : Hj;
            const uint32_t force = twice ? (jj < Hv ? 1u : 2u) : 0u;
            const uint32_t jv = jj < Hv ? jj : jj - Hv;
            Partner v;
            load_partner(P, act, kj, jv, Hv, oj, v);
            const Band b{K.mid, K.half_w, K.h2_w, K.tau, K.lm};
            const uint32_t cm = K.info & 3u;
            const bool even = (K.info & 4u) != 0u;
            // with cm != 0 a leaf passes the area tests in at most one
            // orientation: 1 when its class is 0 on cm, 2 (even pairs) when 3
            // on cm.  As compares of c = (class of u v) & cmsk: orientation 1
            // when c == t1, 2 when c == t2 (a forced run has cmsk = 0 and
            // matches one of them; 4 never matches).
            const uint32_t cmsk = force ? 0u : cm;
            const uint32_t t1 = force == 2u ? 4u : 0u;
            const uint32_t t2 = force == 2u ? 0u : ((force == 0u && even && cm != 0u) ? cm : 4u);
            const uint32_t jpk = jv << ID_U_BITS;
            // rows of non-central u: the leaves u v and, unless v is real,
            // u conj(v) (same class: conjugation keeps it), both in the
            // orientation of their class.  Orientation 2 of a leaf is
            // orientation 1 of the leaf times -i, so each row evaluates
            // D = Re(u w) and Re(u conj(w)) with w = v or -i v; instantiated for
            // tiles with and without second-sign terms
            const auto rows = [&](auto one_c) {
                constexpr bool ONE = decltype(one_c)::value;
                for (uint32_t ui = 0; ui + 1 < L.H; ++ui) {
                    const LeftFast u = lf[ui];
                    const uint32_t c = (u.cls ^ v.cls) & cmsk;
                    const bool is2 = c == t2;
                    const bool live = act && (c == t1 || is2);
                    const double ce = is2 ? v.d : v.c, de = is2 ? -v.c : v.d;
                    const double t = u.im * de;
                    const double D0 = __builtin_fma(u.re, ce, -t), D1 = __builtin_fma(u.re, ce, t);
                    bool u0, u1;
                    const uint32_t f0 = o1_decide<ONE>(live, D0, b, u0);
                    const uint32_t f1 = o1_decide<ONE>(live && !v.vc, D1, b, u1);
                    if (__builtin_amdgcn_ballot_w64(u0 | u1)) {  // rare: the leaf in its orientation only
                        const uint32_t ko = kj | (is2 ? DEF_O2 : DEF_O1) | DEF_FULL;
                        push_global(P, u0, DefItem{tc.left, ui, v.vidx, ko});
                        push_global(P, u1, DefItem{tc.left, ui, v.vidx, ko | DEF_CONJ});
                    }
                    const uint32_t f = f0 | f1 << 2;
                    root_push(P, rq, tc, f != 0u, D0, D1, ui | jpk, kj | f << 28);
                }
            };
            const double cu = lf[L.H - 1].re;  // the central entry is real
            // leaf orientations the rows test: entries matching t1, then t2 (a
            // forced run has cmsk = 0: all H - 1 rows in its one orientation)
            if (L.H > 1) {
                const uint32_t nrow = ncls[v.cls | cmsk << 2 | (t2 == cm && cmsk != 0u ? 16u : 0u)];
                evald += act ? (uint64_t)(nrow << (v.vc ? 0 : 1)) : 0ull;
            }
            bool cev;
            if (one) {
                rows(std::true_type{});
                cev = central_win<true>(P, tc, act, cu, L.H - 1, v, b, cm, even, force, cnt);
            } else {
                rows(std::false_type{});
                cev = central_win<false>(P, tc, act, cu, L.H - 1, v, b, cm, even, force, cnt);
            }
            evald += cev ? 1u : 0u;
        });
        root_flush(P, rq, tc);
        tile_done(P, S, cnt, rq.tested, leaves, pairs, evald, P.evaluated + 1);
        rq.tested = 0;
    }
}

// ------------------------------------------------------ window-mode kernel
//
// Tiles whose left set is stored sorted by (area group, theta) (hss_math.h:
// the class rule and orientation windows).  Per chunk of 64 partner entries
// (one per lane, mapped as in cell_kernel):
//   phase 1 (lane = partner entry v): the leaf of the left set's central entry
//     with v (class rule, both orientations); the pair's band, orientation-1
//     window and the angle of v go to a per-lane LDS record.  v has one
//     virtual partner w = i^rho v per orientation (rho = 1, orientation 2,
//     on even pairs only) and class variant j (the class bits the pair leaves
//     free); the virtual partners of the chunk are numbered by a prefix sum;
//   phase 2 (lane = virtual partner, 64 at a time): the index ranges of the
//     one area group whose leaves u w and u conj(w) pass the area tests,
//     inside the orientation-1 window shifted by the angle of w, cut into
//     units of WIN_UNIT entries (a second per-lane LDS record);
//   phase 3 (lane = one unit): the flattened units of those 64 virtual
//     partners, 64 at a time; a lane evaluates its unit's leaves: X = Re(u w),
//     the FP64 band of orientation 1 and the root test of each counted term.
// Leaves outside the ranges count 0 exactly (wrong orientation, wrong class
// or outside the window) and are skipped; a leaf inside them with an
// undecided band goes to the exact stage with its orientation only.

// Pair of a chunk lane's k (computed once per k, read by its entries in all
// three phases).  Records are 16-byte aligned and a multiple of 16 bytes
// long, so that every wide LDS access is aligned.
struct alignas(16) WinK {
    double mid, half_w, h2_w, tau, lm;  // the pair's band (hss_math.h Band)
    uint32_t wa, wla, wb;               // orientation-1 window arcs in angle units (Window1F)
    uint32_t info;                      // cm | even << 2 | window arcs << 3
    uint32_t kj;                        // k
    uint32_t pad;
};
static_assert(sizeof(WinK) == 64, "LDS budget of cell_win_kernel");

// Partner entry of a chunk (phase 1 -> 2, 3).
struct alignas(16) WinPart {
    double c, d;                   // v
    uint32_t thv;                  // angle of v modulo pi, angle units
    uint32_t vidx;                 // v's cache entry
    uint32_t info;                 // class | real v << 2 | force << 3 | k's chunk lane << 5
    uint32_t pad;
};
static_assert(sizeof(WinPart) == 32, "LDS budget of cell_win_kernel");

// Virtual partner of a batch (phase 2 -> 3): its ranges and partner lane.
// 64 + 32 + 48 bytes per lane: four blocks of four waves fit a CU's 160 KB of LDS.
struct alignas(16) WinVp {
    uint32_t se[8];  // range p (0-3 leaves u w, 4-7 u conj(w)): start | end << 16 (left-set entry indices)
    uint32_t uc[4];  // units before ranges 2i (low half) and 2i + 1; range 0's (always 0) holds
                     // the partner lane | rho << 6 | the lane of its k << 7
};
static_assert(sizeof(WinVp) == 48, "LDS budget of cell_win_kernel");

// Phase 3: the flattened units of a batch of virtual partners.
template <bool ONE>
__device__ inline void win_units(const CellParams& P, const TileCtx& tc, const LeftFast* lf, const WinK* kws,
                                 const WinPart* part, const WinVp* vps, uint32_t* mark, uint32_t uex, uint32_t units,
                                 uint32_t nunits, uint64_t& wcnt) {
    const uint32_t lane = __lane_id();
    uint32_t carry = 0;
    for (uint32_t w0 = 0; w0 < nunits; w0 += 64) {
        const uint32_t wi = w0 + lane;
        const uint32_t o = step_owner(mark, uex, units, w0, carry);
        const uint32_t first = (uint32_t)__shfl((int)uex, (int)o, 64);
        const bool liveu = wi < nunits;
        const uint32_t wq = liveu ? wi - first : 0u;
        const WinVp& Q = *(const WinVp*)((const char*)vps + __umul24(o, (uint32_t)sizeof(WinVp)));
        const uint32_t pr = Q.uc[0] & 0xFFFFu;
        // the unit's range pc: the number of ranges p >= 1 whose units-before
        // uc[p] <= wq (uc never decreases), two halfwords at a time: every
        // count is below 2^15 (8 ranges of at most 2048 units) and so is pr,
        // so (2^15 + wq) - uc[p] in each half keeps its guard bit 15 exactly
        // when uc[p] <= wq and never borrows from the next half; ucp = uc[pc]
        // read back by its halfword address (range 0 starts at unit 0)
        static_assert(8 * ((WIN_MAX_SORT + WIN_UNIT - 1) / WIN_UNIT) < 0x8000u, "unit counts below 2^15");
        const uint32_t wg = (wq | wq << 16) | 0x80008000u;
        uint32_t pc = __popc((wg - Q.uc[0]) & 0x80000000u);  // its low half is pr
#pragma unroll
        for (int i = 1; i < 4; ++i) pc += __popc((wg - Q.uc[i]) & 0x80008000u);
        const uint32_t ucp = pc ? (uint32_t)*(const uint16_t*)((const char*)Q.uc + 2u * pc) : 0u;
        const uint32_t se = Q.se[pc];
        const uint32_t ui0 = (se & 0xFFFFu) + (wq - ucp) * WIN_UNIT;
        const uint32_t uend = liveu ? se >> 16 : 0u;
        const bool cj = pc >= 4;
        const WinPart& V = *(const WinPart*)((const char*)part + __umul24(pr & 63u, (uint32_t)sizeof(WinPart)));
        const WinK& K = *(const WinK*)((const char*)kws + __umul24((pr >> 7) & 63u, (uint32_t)sizeof(WinK)));
        const bool rho = (pr >> 6) & 1u;
        // w = i^rho v; u conj(w) is u w with Im w negated: one sign per unit, none per leaf
        const double vc = V.c, vd = V.d;
        const double ce = rho ? -vd : vc, dw = rho ? vc : vd;
        const double de = cj ? -dw : dw;
        const uint32_t kj = K.kj;
        const Band b{K.mid, K.half_w, K.h2_w, K.tau, K.lm};
        const uint32_t kinfo = kj | (rho ? DEF_O2 : DEF_O1) | (cj ? DEF_CONJ : 0u);
        const uint32_t vidx = V.vidx;
        const uint32_t nlive = uend > ui0 ? uend - ui0 : 0u;  // live entries of the unit
        // the unit's entries from one base: ui0 <= nsort, and the left entries
        // carry WIN_UNIT entries of padding, so every load is in bounds;
        // entries past uend are masked by live
        const LeftFast* up = (const LeftFast*)((const char*)lf + __umul24(ui0, (uint32_t)sizeof(LeftFast)));
        double AX[WIN_UNIT];
        bool F1[WIN_UNIT], F2[WIN_UNIT];  // counted terms 1 and 2 (o1_decide's bits, kept as lane masks)
        bool U[WIN_UNIT];
#pragma unroll
        for (uint32_t s = 0; s < WIN_UNIT; ++s) {
            const LeftFast u = up[s];
            const double X = __builtin_fma(u.re, ce, -(u.im * de));
            AX[s] = __builtin_fabs(X);
            const uint32_t f = o1_decide<ONE>(s < nlive, X, b, U[s]);
            F1[s] = (f & 1u) != 0u;
            F2[s] = (f & 2u) != 0u;
        }
        bool uany = false;
#pragma unroll
        for (uint32_t s = 0; s < WIN_UNIT; ++s) uany |= U[s];
        if (__builtin_amdgcn_ballot_w64(uany)) {  // rare
#pragma unroll
            for (uint32_t s = 0; s < WIN_UNIT; ++s)
                push_global(P, U[s], DefItem{tc.left, ui0 + s, vidx, kinfo | DEF_FULL});
        }
        // the candidates: counted terms summed over the wave (scalar popcounts of the lane masks)
        bool pass[WIN_UNIT];
        bool pany = false;
        const double tol1 = root_tol1(b.lm);
#pragma unroll
        for (uint32_t s = 0; s < WIN_UNIT; ++s) {
            wcnt += (uint64_t)__popcll(__builtin_amdgcn_ballot_w64(F1[s]));
            if constexpr (!ONE) wcnt += (uint64_t)__popcll(__builtin_amdgcn_ballot_w64(F2[s]));
            // both operands evaluated (&, not &&): nearly every evaluated leaf
            // has a counted term, so a per-leaf branch would only add mask work
            pass[s] = F1[s] & root_near_integer(__builtin_fma(-2.0, AX[s], b.lm), tol1);
        }
        if constexpr (!ONE) {
            bool any2 = false;
#pragma unroll
            for (uint32_t s = 0; s < WIN_UNIT; ++s) any2 |= F2[s];
            if (__builtin_amdgcn_ballot_w64(any2)) {
#pragma unroll
                for (uint32_t s = 0; s < WIN_UNIT; ++s)
                    pass[s] |= F2[s] & root_near_integer(__builtin_fma(2.0, AX[s], b.lm), tc.tol2);
            }
        }
#pragma unroll
        for (uint32_t s = 0; s < WIN_UNIT; ++s) pany |= pass[s];
        if (__builtin_amdgcn_ballot_w64(pany)) {  // rare: near-integer fourth roots
#pragma unroll
            for (uint32_t s = 0; s < WIN_UNIT; ++s) push_global(P, pass[s], DefItem{tc.left, ui0 + s, vidx, kinfo});
        }
    }
}

__global__ __launch_bounds__(WG) void cell_win_kernel(HSS_CELL_ARGS) {
    __shared__ WinK s_k[WAVES][64];
    __shared__ WinPart s_part[WAVES][64];
    __shared__ WinVp s_vp[WAVES][64];
    __shared__ uint32_t s_mark[WAVES][65];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    WinK* kws = s_k[wave];
    WinPart* part = s_part[wave];
    WinVp* vps = s_vp[wave];
    uint32_t* mark = s_mark[wave];
    for (;;) {
        const uint32_t t = __builtin_amdgcn_readfirstlane(atomicAdd(P.next_wtile, lane == 0 ? 1u : 0u));
        if (t >= hdr->nwtiles) break;
        const TileDesc T = tiles[t];
        const LeftDesc L = lefts[T.left];
        const PosDesc& S = pos[L.pos];
        const TileCtx tc{T.left, L.q, S.l4_d, root_tol2(S.n4_d)};
        const LeftFast* lf = lfast + L.off;
        const uint32_t* bt = P.btab + (size_t)L.bt * (WIN_NB + 1);
        const uint32_t nsort = L.H - 1;
        const uint32_t kmask = left_kmask(L);
        const uint32_t nbg = WIN_NB >> __popc(kmask);
        const uint32_t bsh = 32u - (uint32_t)__builtin_ctz(nbg);  // bucket of an angle: its top bits
        const double cu = lf[nsort].re;  // the central entry is real
        const bool one = __builtin_amdgcn_readfirstlane(tile_one_term(L.q, T.k0, S.l4_d, S.l2_d, S.n4_d) ? 1 : 0) != 0;
        uint32_t cnt = 0;     // per lane: candidates of the central entry's leaves
        uint64_t wcnt = 0;    // per wave: candidates of the units
        uint64_t leaves = 0, pairs = 0, evald = 0;
        // per k: the pair's band, area info and orientation-1 window; an even
        // pair with cm = 0 runs its entries twice, one orientation per run
        const auto on_chunk = [&](uint32_t k) {
            const Band b = pair_band(L, S, k);
            const uint32_t ainfo = pair_area_info(L, S, k);
            const double md = (double)L.q * (double)k;  // exact: m < 2^26
            const double m2 = md * md;                  // exact: m^2 < 2^52
            Window1 w;
            make_window_o1((S.l4_d + m2 * m2) - S.n4_d, S.l4_d, b.tau, 2.0 * S.l2_d * m2, w);
            Window1F wf;
            window_o1_fx(w, wf);
            __builtin_amdgcn_wave_barrier();
            lds_at(kws, lane) = WinK{b.mid, b.half_w, b.h2_w, b.tau, b.lm, wf.a, wf.la, wf.b,
                                     ainfo | (uint32_t)(w.n & 3) << 3, k, 0u};
            __builtin_amdgcn_wave_barrier();
            return ((ainfo & 3u) == 0u && (ainfo & 4u) != 0u) ? 2u : 1u;
        };
        for_partner_chunks(P, T, L, S, mark, pairs, leaves, on_chunk,
                           [&](bool act, uint32_t kj, uint32_t jj, uint32_t Hj, uint32_t j, uint32_t oj) {
            // ---- phase 1: lane = partner entry
            const WinK& K = lds_at(kws, j);
            const uint32_t kinf = K.info;
            const uint32_t cm = kinf & 3u;
            const bool even = (kinf & 4u) != 0u;
            const bool twice = cm == 0u && even;
            const uint32_t Hv = twice ? Hj >> 1 : Hj;
            const uint32_t force = twice ? (jj < Hv ? 1u : 2u) : 0u;
            Partner v;
            load_partner(P, act, kj, jj < Hv ? jj : jj - Hv, Hv, oj, v);
            const Band b{K.mid, K.half_w, K.h2_w, K.tau, K.lm};
            evald += (one ? central_win<true>(P, tc, act, cu, nsort, v, b, cm, even, force, cnt)
                          : central_win<false>(P, tc, act, cu, nsort, v, b, cm, even, force, cnt))
                         ? 1u
                         : 0u;
            // virtual partners: class variants j times the orientations of this run
            const uint32_t nvp = act ? allowed_count(cm, kmask) << ((even && !force) ? 1 : 0) : 0u;
            __builtin_amdgcn_wave_barrier();
            lds_at(part, lane) = WinPart{v.c, v.d, theta_fx_of_f(v.c, v.d), v.vidx,
                                         v.cls | (v.vc ? 4u : 0u) | force << 3 | j << 5, 0u};
            const uint32_t vincl = wave_incl_scan(nvp);
            const uint32_t vex = vincl - nvp;
            const uint32_t nvps = __builtin_amdgcn_readlane(vincl, 63);
            __builtin_amdgcn_wave_barrier();
            // ---- phase 2: lane = virtual partner, 64 at a time
            uint32_t vcarry = 0;
            for (uint32_t v0 = 0; v0 < nvps; v0 += 64) {
                const uint32_t vi = v0 + (uint32_t)lane;
                const uint32_t o = step_owner(mark, vex, nvp, v0, vcarry);
                const bool livev = vi < nvps;
                // every lane executes the cross-lane read (a source lane may not be live)
                const uint32_t ovex = (uint32_t)__shfl((int)vex, (int)o, 64);
                const uint32_t q = livev ? vi - ovex : 0u;
                const WinPart& V = *(const WinPart*)((const char*)part + __umul24(o, (uint32_t)sizeof(WinPart)));
                const uint32_t info = V.info;
                const uint32_t kl = (info >> 5) & 63u;
                const WinK& VK = *(const WinK*)((const char*)kws + __umul24(kl, (uint32_t)sizeof(WinK)));
                const uint32_t vkinf = VK.info;
                const uint32_t vcls = info & 3u, vcm = vkinf & 3u;
                const bool veven = (vkinf >> 2) & 1u, vreal = (info >> 2) & 1u;
                const uint32_t vforce = (info >> 3) & 3u;
                const bool both = veven && !vforce;
                const uint32_t rho = vforce ? vforce - 1u : (both ? (q & 1u) : 0u), jv = both ? q >> 1 : q;
                const uint32_t g = area_group(allowed_class(vcls ^ (rho ? 3u : 0u), vcm, kmask, jv), kmask);
                const uint32_t g0 = g * nbg;
                const BucketGroup G{bt, g0, bsh, bucket_at(bt, g0), bucket_at(bt, g0 + nbg)};
                Window1F vw;
                vw.a = VK.wa;
                vw.la = VK.wla;
                vw.b = VK.wb;
                const uint32_t wn = (vkinf >> 3) & 3u;
                vw.n = wn == 3u ? -1 : (int)wn;
                // w = i^rho v: its angle modulo pi wraps with the unsigned sum
                const uint32_t thw = V.thv + (rho ? ANGLE_HALF_PI : 0u);
                uint32_t rs[8], re[8];
                const uint32_t n0 = window_o1_ranges(vw, G, thw, false, rs, re);
                const uint32_t n1 = window_o1_ranges(vw, G, thw, true, rs + 4, re + 4);
                const bool live0 = livev && n0 != 0, live1 = livev && !vreal && n1 != 0;
                __builtin_amdgcn_wave_barrier();
                WinVp& R = vps[lane];
                uint32_t items = 0, units = 0, ucs[8];
#pragma unroll
                for (int p = 0; p < 8; ++p) {
                    const uint32_t len = (p < 4 ? live0 : live1) ? re[p] - rs[p] : 0u;
                    R.se[p] = rs[p] | (rs[p] + len) << 16;
                    ucs[p] = p ? units : (o | rho << 6 | kl << 7);
                    items += len;
                    units += (len + WIN_UNIT - 1) / WIN_UNIT;
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) R.uc[i] = ucs[2 * i] | ucs[2 * i + 1] << 16;
                evald += items;
                const uint32_t ucl = wave_incl_scan(units);
                const uint32_t uex = ucl - units;  // first unit of this lane's virtual partner
                const uint32_t nunits = __builtin_amdgcn_readlane(ucl, 63);
                __builtin_amdgcn_wave_barrier();
                // ---- phase 3: lane = unit of WIN_UNIT consecutive entries of one range
                if (one)
                    win_units<true>(P, tc, lf, kws, part, vps, mark, uex, units, nunits, wcnt);
                else
                    win_units<false>(P, tc, lf, kws, part, vps, mark, uex, units, nunits, wcnt);
                __builtin_amdgcn_wave_barrier();
            }
        });
        tile_done(P, S, cnt, wcnt, leaves, pairs, evald, P.evaluated);
    }
}

// Exact settlement of deferred leaves: full reference chain with 128-bit
// integers; counts are added only for leaves the fast path left undecided.
// The item count is read on the device (an overflowed buffer holds def_cap
// items; the host then runs the batch again with a larger buffer).
__global__ __launch_bounds__(256) void k_exact(const CellParams P) {
    const uint32_t n = *P.ndef < P.def_cap ? *P.ndef : P.def_cap;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const DefItem it = P.def[i];
        const LeftDesc L = P.lefts[it.left];
        const PosDesc& S = P.pos[L.pos];
        const uint32_t k = it.kflags & DEF_K_MASK;
        const bool conj = (it.kflags & DEF_CONJ) != 0, full = (it.kflags & DEF_FULL) != 0;
        const uint32_t om = (it.kflags >> 27) & 3u;
        const uint64_t m = (uint64_t)L.q * k;
        ExactCtx e;
        make_exact_ctx(S.l, m, S.limit, e);
        i128 zre, zim;
        exact_product(P.lre[L.off + it.ui], P.lim[L.off + it.ui], (int64_t)P.cre[it.vidx], (int64_t)P.cim[it.vidx],
                      conj, zre, zim);
        uint32_t c = exact_leaf(zre, zim, e, P.qr, [&](uint64_t r, u128 a4) {
            uint32_t slot = atomicAdd(P.nraw, 1u);
            if (slot < P.raw_cap) P.raw[slot] = HitOut{S.l, m, r, (uint64_t)a4, (uint64_t)(a4 >> 64), L.pos, 0u};
        }, full && om ? om : 3u);
        if (full && c) atomicAdd(P.cand + S.out, (unsigned long long)c);
    }
}

// Acceptance of the raw exact fourth powers: gcd, 5 | lmn, ordering and the
// 256-bit Heron recheck against the position's own LIMIT; accepted triples
// are appended to the hit list.
__global__ __launch_bounds__(256) void k_accept(const PosDesc* pos, const HitOut* raw, const uint32_t* nraw, uint32_t raw_cap, HitOut* hits,
                         uint32_t* nhits, uint32_t hit_cap) {
    uint32_t n = *nraw < raw_cap ? *nraw : raw_cap;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        HitOut h = raw[i];
        u128 a4 = ((u128)h.area4_hi << 64) | h.area4_lo;
        if (!accept_hit(h.l, h.m, h.n, pos[h.pos].limit, a4)) continue;
        uint32_t slot = atomicAdd(nhits, 1u);
        if (slot < hit_cap) hits[slot] = h;
    }
}

// Start of a batch: zero the work-queue heads and the deferred / raw / hit
// counters (misc[0..15]) and the counters of the batch's positions.  A kernel
// rather than memset nodes, so the launch sequence can be captured in a graph.
__global__ void k_batch_reset(const BatchHeader* hdr, const PosDesc* pos, unsigned long long* cand,
                              unsigned long long* leaves, unsigned long long* pairs, uint32_t* misc) {
    const uint32_t npos = hdr->npos;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 16) misc[i] = 0;
    for (; i < npos; i += gridDim.x * blockDim.x) {
        const uint32_t o = pos[i].out;
        cand[o] = 0;
        leaves[o] = 0;
        pairs[o] = 0;
    }
}

// ------------------------------------------------------------ launchers

#define HSS_CHECK(x)                                  \
    do {                                              \
        hipError_t err_ = (x);                        \
        if (err_ != hipSuccess) return err_;          \
    } while (0)

hipError_t launch_build_tables(const BuildBuffers& B, hipStream_t st) {
    uint64_t n = B.n;
    uint32_t g = (uint32_t)((n + 256) / 256);
    hipLaunchKernelGGL(k_fill_identity, dim3(g), dim3(256), 0, st, B.spf, n);
    if (B.nsmall) hipLaunchKernelGGL(k_sieve_mark, dim3(B.nsmall), dim3(1024), 0, st, B.spf, B.small_primes, n);
    uint32_t g2 = (uint32_t)((n + 2 + 255) / 256);
    hipLaunchKernelGGL(k_half_counts, dim3(g2), dim3(256), 0, st, (const uint32_t*)B.spf, B.H, n);
    HSS_CHECK(hipGetLastError());
    return hipSuccess;
}

hipError_t launch_scan(const uint32_t* in, uint32_t* out, uint32_t* sums, uint64_t n, hipStream_t st) {
    uint32_t nb = (uint32_t)((n + 4095) / 4096);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, st, in, out, sums, n);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, st, sums, nb);
    hipLaunchKernelGGL(k_scan_add, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, out, (const uint32_t*)sums, n);
    return hipGetLastError();
}

hipError_t launch_collect_split(const uint32_t* spf, uint64_t n, uint32_t* plist, uint32_t* psize,
                                uint32_t* count, hipStream_t st) {
    uint64_t cnt = n >= 5 ? (n - 5) / 4 + 1 : 0;
    if (!cnt) return hipSuccess;
    hipLaunchKernelGGL(k_collect_split, dim3((uint32_t)((cnt + 255) / 256)), dim3(256), 0, st, spf, n, plist, psize,
                       count);
    return hipGetLastError();
}

hipError_t launch_prime_values(const uint32_t* plist, uint32_t np, const uint32_t* vstart, uint32_t total,
                               uint32_t* vbase, int64_t* pa, int64_t* pb, i128* vre, i128* vim, hipStream_t st) {
    if (!np) return hipSuccess;
    hipLaunchKernelGGL(k_prime_split, dim3((np + 255) / 256), dim3(256), 0, st, plist, np, vstart, vbase, pa, pb);
    hipLaunchKernelGGL(k_prime_values, dim3((total + 255) / 256), dim3(256), 0, st, plist, np, vstart, total,
                       (const int64_t*)pa, (const int64_t*)pb, vre, vim);
    return hipGetLastError();
}

hipError_t launch_fill_cache(const uint32_t* spf, const uint32_t* off, uint64_t n, uint32_t total,
                             const uint32_t* vbase, const i128* vre, const i128* vim, double* cre, double* cim,
                             uint32_t* cres, hipStream_t st) {
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_fill_cache, dim3((total + 255) / 256), dim3(256), 0, st, spf, off, n, total, vbase, vre,
                       vim, cre, cim, cres);
    return hipGetLastError();
}

hipError_t launch_build_left(const BatchHeader* hdr, const LeftBuild* builds, const LeftDesc* lefts,
                             const uint32_t* vbase, const i128* vre, const i128* vim, const uint32_t* sorted,
                             uint32_t* perm, uint32_t* btab, i128* lre, i128* lim, LeftFast* lfast, hipStream_t st) {
    // the left-set count is known on the device only (the launch may be
    // captured once and replayed), so the grids are fixed: large enough that
    // most blocks take one left set and the hardware hands blocks to CUs as
    // they free up (a left set's cost varies by two orders of magnitude);
    // blocks past the count exit at once.  Measured on the 1536-root LIMIT
    // 1e7 batch: 6.6 ms for the pair against 7.5 ms with 1024 and 2048 blocks
    hipLaunchKernelGGL(k_left_sort, dim3(16384), dim3(256), 0, st, hdr, sorted, builds, lefts, vbase, vre, vim, perm,
                       btab);
    hipLaunchKernelGGL(k_build_left, dim3(16384), dim3(256), 0, st, hdr, builds, lefts, vbase, vre, vim,
                       (const uint32_t*)perm, lre, lim, lfast);
    return hipGetLastError();
}

hipError_t launch_cells(const CellParams& P, uint32_t blocks, hipStream_t st) {
    hipLaunchKernelGGL(cell_kernel, dim3(blocks), dim3(WG), 0, st, P, P.hdr, P.tiles, P.lefts, P.pos, P.lfast);
    return hipGetLastError();
}

hipError_t launch_cells_win(const CellParams& P, uint32_t blocks, hipStream_t st) {
    hipLaunchKernelGGL(cell_win_kernel, dim3(blocks), dim3(WG), 0, st, P, P.hdr, P.wtiles, P.lefts, P.pos, P.lfast);
    return hipGetLastError();
}

hipError_t launch_exact(const CellParams& P, uint32_t blocks, hipStream_t st) {
    hipLaunchKernelGGL(k_exact, dim3(blocks), dim3(256), 0, st, P);
    return hipGetLastError();
}

hipError_t launch_accept(const PosDesc* pos, const HitOut* raw, const uint32_t* nraw, uint32_t raw_cap, HitOut* hits,
                         uint32_t* nhits, uint32_t hit_cap, hipStream_t st) {
    hipLaunchKernelGGL(k_accept, dim3(16), dim3(256), 0, st, pos, raw, nraw, raw_cap, hits, nhits, hit_cap);
    return hipGetLastError();
}

hipError_t launch_batch_reset(const BatchHeader* hdr, const PosDesc* pos, unsigned long long* cand,
                              unsigned long long* leaves, unsigned long long* pairs, uint32_t* misc, hipStream_t st) {
    hipLaunchKernelGGL(k_batch_reset, dim3(256), dim3(256), 0, st, hdr, pos, cand, leaves, pairs, misc);
    return hipGetLastError();
}

hipError_t cell_occupancy(int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, cell_kernel, WG, 0);
}

hipError_t cell_win_occupancy(int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, cell_win_kernel, WG, 0);
}

// ------------------------------------------------ no-allocation scratch
//
// Every per-launch buffer of hss_search_noalloc (plan, left sets, tiles,
// counters, deferred leaves, hits) is carved from this array, so that entry
// point allocates no device memory.  Each device holds its own copy for the
// life of the process.  In host code `scratch` names a host-side stand-in:
// only the address hipGetSymbolAddress returns may reach a kernel or a copy.
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

hipError_t noalloc_scratch(int device, void** ptr) {
    static constexpr int MAX_DEVICES = 64;
    static std::mutex mu;
    static void* addr[MAX_DEVICES] = {};
    if (!ptr || device < 0 || device >= MAX_DEVICES) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> guard(mu);
    if (!addr[device]) {
        int current = 0;
        hipError_t e = hipGetDevice(&current);
        if (e != hipSuccess) return e;
        e = hipSetDevice(device);
        if (e != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        const hipError_t back = hipSetDevice(current);
        if (e != hipSuccess) return e;
        if (back != hipSuccess) return back;
        if (!p) return hipErrorInvalidSymbol;
        addr[device] = p;
    }
    *ptr = addr[device];
    return hipSuccess;
}

}  // namespace hss
