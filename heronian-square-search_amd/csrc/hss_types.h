// hss_types.h -- plain data shared by the host runtime and the gfx950 kernels.
#pragma once

#include <stdint.h>

#include "hss_math.h"

namespace hss {

static constexpr int MAX_SPLIT = 8;  // distinct primes = 1 (mod 4) of an integer <= 2^26
static constexpr int MAX_INERT = 8;  // distinct primes = 3 (mod 4)
// Window mode: left sets with WIN_MIN_SORT <= H <= WIN_MAX_SORT + 1 are stored
// sorted by the angle of their entries modulo pi/2 (the central entry stays
// last) with a table of WIN_NB + 1 bucket starts.
static constexpr uint32_t WIN_NB = 1024;
static_assert(WIN_NB / 4 == 256, "WIN_GAP (hss_math.h) is two buckets of the coarsest group slice");
static constexpr uint32_t WIN_MAX_SORT = 8192;
static constexpr uint32_t WIN_MIN_SORT = 8;
// Window-mode leaves are evaluated in units of WIN_UNIT consecutive sorted
// entries; a unit's loads may run up to WIN_UNIT - 1 entries past the end of
// its left set, so the left-entry arrays carry WIN_UNIT entries of padding.
static constexpr uint32_t WIN_UNIT = 4;
static constexpr uint32_t NO_TABLE = 0xFFFFFFFFu;
// cache residue word: re mod 180 (bits 0-7), area class (bits 8-9), im mod 180 << 16
static constexpr uint32_t CRES_CLASS_SHIFT = 8;
// Leaves queued for the root test are named by (entry u, partner entry within
// its half-set, k, conjugate, counted terms) packed into 64 bits: u below
// 2^17, the partner entry below 2^15 (the planner and the table build check),
// k below 2^27 (k <= N_max <= 2^26).
static constexpr uint32_t ID_U_BITS = 17;
static constexpr uint32_t MAX_LEFT_ENTRIES = (1u << ID_U_BITS) - 1;
static constexpr uint32_t MAX_HALF_SET = (1u << (32 - ID_U_BITS)) - 1;

// One root position l of a batch (uniform across a tile).  Positions of one
// batch may carry different LIMITs.
struct PosDesc {
    uint64_t l;
    uint64_t limit;   // LIMIT of this position: partners and third sides <= limit
    double n4_d;      // FP64 value of limit^4
    double l4_d;      // FP64 value of l^4
    double l2_d;      // l^2 (exact)
    uint32_t l3, l5;  // l mod 3, l mod 5
    uint32_t l_even;
    uint32_t nsplit, ninert;
    uint32_t out;     // index of this position's counters (candidates, leaves, pairs)
    // divisibility of a partner index k by the split / inert primes of l:
    // p | k  <=>  k * inv (mod 2^32) <= thr
    uint32_t split_inv[MAX_SPLIT], split_thr[MAX_SPLIT];
    uint32_t inert_inv[MAX_INERT], inert_thr[MAX_INERT];
};

// One left set: the half-set of Gaussian products for root l and one
// signature of shared split primes (hss_runtime.cpp, plan_batch).
struct LeftDesc {
    uint32_t pos;      // position index in the batch
    uint32_t off;      // first entry in the left arrays
    uint32_t H;        // entries; entry H-1 is the real (central) one
    uint32_t q;        // product of the shared prime powers: m = q * k
    uint32_t flags;    // LEFT_TRIVIAL | area key mask << 1 (left_kmask)
    uint32_t q3, q5;   // q mod 3, q mod 5
    uint32_t bt;       // bucket table index of a sorted left set, else NO_TABLE
};

// LeftDesc::flags: l has no split prime (partners without one are skipped);
// bits 1-2: the class bits any pair of the left set constrains (AREA_C3 when
// 3 does not divide l, AREA_C5 when 5 does not), the group key of a sorted set.
static constexpr uint32_t LEFT_TRIVIAL = 1;
HSS_HD inline uint32_t left_kmask(const LeftDesc& L) { return (L.flags >> 1) & 3u; }
HSS_HD inline uint32_t left_flags(bool trivial, uint64_t l) {
    return (trivial ? LEFT_TRIVIAL : 0u) | ((l % 3 ? AREA_C3 : 0u) | (l % 5 ? AREA_C5 : 0u)) << 1;
}

// Build recipe of one left set: fixed part and merged exponents of l's split primes.
struct LeftBuild {
    uint64_t fixed;
    uint32_t k;
    uint32_t p[MAX_SPLIT];
    uint32_t E[MAX_SPLIT];
};

// One unit of scheduled work: a left set against partner indices k in [k0, k1].
struct TileDesc {
    uint32_t left;
    uint32_t k0, k1;
    uint32_t pad;
};

// FP64 view of a left entry used by the hot loop (one scalar load).
struct LeftFast {
    double re, im;  // FP64 values (exact below 2^53, else correctly rounded)
    uint32_t cls;   // area class (the class rule of hss_math.h)
    uint32_t pad;
};

// Sizes of one batch, read by the kernels from device memory so that the
// launch sequence (fixed grids) can be captured once as a graph and replayed
// for batches of any size.
struct BatchHeader {
    uint32_t npos, nleft, nsorted, ntiles, nwtiles;
    uint32_t pad[3];
};

// A leaf sent to the exact kernel.  kflags = k | omask << 27 | conj << 30 |
// full << 31, where full = 1 means the fast path left the leaf's count
// undecided and the exact kernel counts the terms of the orientations in omask
// (bit 0: orientation 1, bit 1: orientation 2; 0 means both).
static constexpr uint32_t DEF_K_MASK = (1u << 27) - 1;
static constexpr uint32_t DEF_O1 = 1u << 27, DEF_O2 = 2u << 27, DEF_CONJ = 1u << 30, DEF_FULL = 1u << 31;

struct DefItem {
    uint32_t left;
    uint32_t ui;    // left entry index within the left set
    uint32_t vidx;  // global half-set entry of the partner
    uint32_t kflags;
};

}  // namespace hss
