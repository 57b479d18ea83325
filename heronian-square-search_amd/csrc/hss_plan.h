// hss_plan.h -- host-side batch planning (no device code).
//
// For a root l with split primes p_j^(e_j) every partner m in [l, LIMIT] has a
// unique signature sigma = (v_{p_j}(m))_j.  With q_sigma = prod p_j^(v_{p_j}(m))
// the partner factors as m = q_sigma * k where k is coprime to every split prime
// of l, so the pair's Gaussian representations are exactly
//     { 2 * u * v : u in G(l; E_j = e_j + v_{p_j}(m)),  v in G(k) }
// with G(.) the full product set of Gaussian prime values and the fixed part.
// A left set is the half-set of G(l; E) for one signature; partners are the
// indices k in [ceil(l / q), floor(LIMIT / q)].  Reference counterpart: the term
// merge of src/search_v18.cpp:340-361 / src/gpu_search_v42.cu:318-338, which the
// reference performs per pair; here it is done once per (root, signature).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "hss_types.h"

namespace hss {

struct Plan {
    std::vector<PosDesc> pos;
    std::vector<LeftDesc> lefts;
    std::vector<LeftBuild> builds;
    std::vector<TileDesc> tiles;
    std::vector<uint32_t> kmin, kmax;  // partner range per left set
    uint64_t left_entries = 0;
    double est_cells = 0;   // upper bound of (left entry, partner entry) cells
};

// primes up to and including `upto`
std::vector<uint32_t> primes_upto(uint32_t upto);

// Multiplicative inverse of an odd 32-bit value modulo 2^32.
uint32_t inv_mod_2_32(uint32_t p);

// Plan one batch.  `off` is the host copy of the half-set offsets (n_max + 2
// entries); `waves` the number of resident wavefronts used to size tiles.
bool plan_batch(uint64_t limit, const uint64_t* roots, uint32_t n_roots, const std::vector<uint32_t>& small_primes,
                const uint32_t* off, uint64_t n_max, uint64_t waves, Plan& out, std::string& err);

}  // namespace hss
