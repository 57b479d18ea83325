// hss_kernels.h -- launch interface between the host runtime and the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "hss_types.h"

namespace hss {

struct BuildBuffers {
    uint64_t n;
    uint32_t* spf;                // [n+1]
    uint32_t* H;                  // [n+2]
    const uint32_t* small_primes;  // primes <= sqrt(n)
    uint32_t nsmall;
};

struct CellParams {
    // context tables
    const uint32_t* off;
    const double* cre;
    const double* cim;
    const uint32_t* cres;  // (re mod 180) | class << 8 | (im mod 180) << 16
    // batch
    const PosDesc* pos;
    const LeftDesc* lefts;
    const BatchHeader* hdr;   // batch sizes (device memory)
    const TileDesc* tiles;    // plain mode, hdr->ntiles
    const TileDesc* wtiles;   // window mode, hdr->nwtiles
    const uint32_t* btab;     // bucket tables of the sorted left sets, WIN_NB + 1 each
    const LeftFast* lfast;
    const i128* lre;
    const i128* lim;
    uint32_t* next_tile;
    uint32_t* next_wtile;
    unsigned long long* cand;    // per counter index (PosDesc::out)
    unsigned long long* leaves;  // per counter index
    unsigned long long* pairs;   // per counter index: pairs reaching the enumeration
    // per kernel ([0] + i: window kernel, [1] + i: plain kernel): [i = 0] leaf
    // orientations the fast path tested, [i = 2] candidates it counted
    unsigned long long* evaluated;
    DefItem* def;     // leaves for the exact kernel
    uint32_t* ndef;
    uint32_t def_cap;
    HitOut* raw;      // exact fourth powers before the acceptance rules
    uint32_t* nraw;
    uint32_t raw_cap;
    QR4 qr;
};

hipError_t launch_build_tables(const BuildBuffers& B, hipStream_t st);
hipError_t launch_scan(const uint32_t* in, uint32_t* out, uint32_t* sums, uint64_t n, hipStream_t st);
hipError_t launch_collect_split(const uint32_t* spf, uint64_t n, uint32_t* plist, uint32_t* psize,
                                uint32_t* count, hipStream_t st);
hipError_t launch_prime_values(const uint32_t* plist, uint32_t np, const uint32_t* vstart, uint32_t total,
                               uint32_t* vbase, int64_t* pa, int64_t* pb, i128* vre, i128* vim, hipStream_t st);
hipError_t launch_fill_cache(const uint32_t* spf, const uint32_t* off, uint64_t n, uint32_t total,
                             const uint32_t* vbase, const i128* vre, const i128* vim, double* cre, double* cim,
                             uint32_t* cres, hipStream_t st);
// Fixed-grid launches (sizes from hdr): capturable once and replayable.
hipError_t launch_build_left(const BatchHeader* hdr, const LeftBuild* builds, const LeftDesc* lefts,
                             const uint32_t* vbase, const i128* vre, const i128* vim, const uint32_t* sorted,
                             uint32_t* perm, uint32_t* btab, i128* lre, i128* lim, LeftFast* lfast, hipStream_t st);
hipError_t launch_cells(const CellParams& P, uint32_t blocks, hipStream_t st);
hipError_t launch_cells_win(const CellParams& P, uint32_t blocks, hipStream_t st);
hipError_t launch_exact(const CellParams& P, uint32_t blocks, hipStream_t st);
hipError_t launch_accept(const PosDesc* pos, const HitOut* raw, const uint32_t* nraw, uint32_t raw_cap, HitOut* hits,
                         uint32_t* nhits, uint32_t hit_cap, hipStream_t st);
hipError_t launch_batch_reset(const BatchHeader* hdr, const PosDesc* pos, unsigned long long* cand,
                              unsigned long long* leaves, unsigned long long* pairs, uint32_t* misc, hipStream_t st);
hipError_t cell_occupancy(int* blocks_per_cu);
hipError_t cell_win_occupancy(int* blocks_per_cu);

// Device memory of the no-allocation search (hss_search_noalloc): a static
// array of this many bytes in every device's copy of the code object.
static constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)256 << 20;
// Deferred-leaf capacity of that search (16 bytes each, inside the array).
static constexpr uint32_t NOALLOC_DEF_CAP = 1u << 20;
// Device address of that array on `device` (looked up once per device, with
// the device made current for the lookup and restored afterwards).
hipError_t noalloc_scratch(int device, void** ptr);

}  // namespace hss
