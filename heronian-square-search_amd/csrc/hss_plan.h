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

// Scheduling knobs of one batch, keyed on Gaussian weight.  A left set of
// H = (w + 1) / 2 entries (w the Gaussian weight of l with the shared prime
// powers merged) falls in weight class min(floor(log2 H) - 3, 8) (H in
// [9, 8193]: classes 0..8; smaller or larger left sets never use window
// mode).  A tile
// runs in window mode (cell_win_kernel) when the mean range-window fraction of
// its partner range is at most win_max_frac[class]; otherwise in plain mode
// (cell_kernel).  Counts never depend on these values, only speed does.
static constexpr int WEIGHT_CLASSES = 9;

inline int weight_class(uint32_t H) {
    if (H < WIN_MIN_SORT + 1 || H > WIN_MAX_SORT + 1) return -1;
    int c = 31 - __builtin_clz(H) - 3;
    return c < 0 ? 0 : (c >= WEIGHT_CLASSES ? WEIGHT_CLASSES - 1 : c);
}

// Defaults on MI355X, from calibration roots drawn like the uneven LIMIT 1e7
// workload with another seed: the window cost weights by a sweep (window
// item 0.5 and partner entry 8 against 1.4 and 4: 4.5 % less kernel time on
// those roots), then win_max_frac by hss_autotune on the whole calibration
// batch under those weights (3.5 % less again), then tiles_per_wave 48 by a
// sweep of the whole search's time (1.3 % less than 32: finer tiles cost
// planning time, which the device's left-set stage now hides).
struct Tuning {
    uint32_t tiles_per_wave = 48;  // tile size: est. work / (waves * tiles_per_wave)
    double win_max_frac[WEIGHT_CLASSES] = {0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0.8};
    double win_item_cost = 0.5;   // cost of a window item relative to a plain leaf (tile order)
    double win_partner_cost = 8;  // per partner entry of a window tile, in plain-cell units
    double plain_partner_cost = 2.5;  // per partner entry of a plain tile, in plain-cell units
};

struct Plan {
    std::vector<PosDesc> pos;
    std::vector<LeftDesc> lefts;
    std::vector<LeftBuild> builds;
    std::vector<TileDesc> tiles;    // plain mode, heaviest first
    std::vector<TileDesc> wtiles;   // window mode, heaviest first
    std::vector<uint32_t> sorted;   // left sets stored sorted; LeftDesc::bt = index here
    std::vector<uint32_t> kmin, kmax;  // partner range per left set
    uint64_t left_entries = 0;
    double est_cells = 0;   // upper bound of (left entry, partner entry) cells
    double est_work = 0;    // estimated work after window restriction (plain-leaf units / 2)
    double win_cells = 0;   // cells of the window-mode tiles
    // per left set, from plan_lefts for plan_tiles: window mode possible,
    // its window-fraction threshold, estimated cost per partner entry
    std::vector<uint8_t> eligible;
    std::vector<double> fmax, ecost;
};

// primes up to and including `upto`
std::vector<uint32_t> primes_upto(uint32_t upto);

// Multiplicative inverse of an odd 32-bit value modulo 2^32.
uint32_t inv_mod_2_32(uint32_t p);

// One root position of a batch: root side l searched against partners and
// third sides <= limit; `out` indexes the counters the position's counts go to.
struct PosReq {
    uint64_t l, limit;
    uint32_t out;
};

// Plan one batch.  `off` is the host copy of the half-set offsets (n_max + 2
// entries); `waves` the number of resident wavefronts used to size tiles.
// Positions may repeat (the same l twice, or at two LIMITs): each is its own
// unit of work.
bool plan_batch(const PosReq* reqs, uint32_t n, const std::vector<uint32_t>& small_primes, const uint32_t* off,
                uint64_t n_max, uint64_t waves, const Tuning& tune, Plan& out, std::string& err);

// The two halves of plan_batch: positions, left sets and their window
// eligibility (every eligible left set is listed in `sorted`), then the tiles.
// The left-set build and sort may run on the device while plan_tiles runs.
bool plan_lefts(const PosReq* reqs, uint32_t n, const std::vector<uint32_t>& small_primes, const uint32_t* off,
                uint64_t n_max, const Tuning& tune, Plan& P, std::string& err);
bool plan_tiles(const uint32_t* off, uint64_t waves, const Tuning& tune, Plan& P);

// Mean range-window fraction over partners m = q*k, k in [k0, k1], of root l at
// limit (sampled at a few k).
double tile_window_fraction(uint64_t l, uint64_t limit, uint64_t q, uint32_t k0, uint32_t k1, int samples);

// Gaussian weight w(l) = prod (4e + 1) over the prime powers p^e || l with
// p = 1 (mod 4) (experiments/heavy/README.md:9); 0 if l cannot be factored by
// `primes` (needs every prime <= sqrt(l)).
uint32_t gaussian_weight(uint64_t l, const std::vector<uint32_t>& primes);

// Work estimate of one root position in cells (left entry x partner half-set
// entry), the same measure plan_batch sizes tiles with; -1 if l cannot be
// planned (l outside [1, limit], limit > n_max, too many distinct primes).
double estimate_position(uint64_t l, uint64_t limit, const std::vector<uint32_t>& small_primes, const uint32_t* off,
                         uint64_t n_max);

// Contiguous split of the root positions [first, last] into `parts` intervals
// of about equal estimated work: part i is [bounds[i], bounds[i+1] - 1],
// bounds[0] = first, bounds[parts] = last + 1, and bounds[i] is one past the
// first position whose inclusive prefix of work reaches i/parts of the total
// (so a part is empty only when a single position outweighs a share).  False with
// err set when an input is out of range or a position cannot be planned.
bool split_by_work(uint64_t limit, uint64_t first, uint64_t last, uint32_t parts,
                   const std::vector<uint32_t>& small_primes, const uint32_t* off, uint64_t n_max, uint64_t* bounds,
                   std::string& err);

}  // namespace hss
