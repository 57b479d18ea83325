// hss_plan.cpp -- host-side batch planning: root factorisation, shared-prime
// signatures, left-set recipes and tiles (see hss_plan.h).
#include "hss_plan.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>

namespace hss {

std::vector<uint32_t> primes_upto(uint32_t upto) {
    std::vector<uint8_t> comp(upto + 1, 0);
    std::vector<uint32_t> out;
    for (uint32_t i = 2; i <= upto; ++i) {
        if (comp[i]) continue;
        out.push_back(i);
        for (uint64_t j = (uint64_t)i * i; j <= upto; j += i) comp[j] = 1;
    }
    return out;
}

uint32_t inv_mod_2_32(uint32_t p) {
    uint32_t x = p;
    for (int i = 0; i < 5; ++i) x *= 2u - p * x;
    return x;
}

namespace {

struct RootFactors {
    uint64_t fixed = 1;
    uint32_t ns = 0, ni = 0;
    uint32_t sp[MAX_SPLIT], se[MAX_SPLIT];
    uint32_t ip[MAX_INERT];
};

bool factor_root(uint64_t l, const std::vector<uint32_t>& primes, RootFactors& f) {
    uint64_t x = l;
    auto take = [&](uint64_t p, uint32_t a) -> bool {
        if (p % 4 == 1) {
            if (f.ns == MAX_SPLIT) return false;
            f.sp[f.ns] = (uint32_t)p;
            f.se[f.ns] = a;
            ++f.ns;
        } else {
            for (uint32_t t = 0; t < 2 * a; ++t) f.fixed *= p;
            if (p % 4 == 3) {
                if (f.ni == MAX_INERT) return false;
                f.ip[f.ni++] = (uint32_t)p;
            }
        }
        return true;
    };
    for (uint32_t p : primes) {
        if ((uint64_t)p * p > x) break;
        if (x % p) continue;
        uint32_t a = 0;
        while (x % p == 0) {
            x /= p;
            ++a;
        }
        if (!take(p, a)) return false;
    }
    if (x > 1 && !take(x, 1)) return false;
    return true;
}

// partners' half-set entries with index in [a, b]
inline uint64_t entries(const uint32_t* off, uint32_t a, uint32_t b) { return (uint64_t)off[b + 1] - off[a]; }

// Run f(begin, end, chunk) over `chunks` contiguous pieces of [0, n) on up to
// 16 host threads (pieces taken dynamically; results are merged by the caller
// in chunk order, so the plan does not depend on the thread count).
template <typename F>
void parallel_chunks(uint32_t n, uint32_t chunks, F&& f) {
    const uint32_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const uint32_t nt = std::min(hw, chunks);
    std::atomic<uint32_t> next{0};
    auto run = [&] {
        for (uint32_t c; (c = next.fetch_add(1)) < chunks;)
            f((uint32_t)((uint64_t)n * c / chunks), (uint32_t)((uint64_t)n * (c + 1) / chunks), c);
    };
    if (nt <= 1) {
        run();
        return;
    }
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < nt; ++t) pool.emplace_back(run);
    run();
    for (std::thread& t : pool) t.join();
}

// Visit every signature of root l: exponent vectors s_j >= 0 over l's split
// primes with q = prod p_j^s_j <= limit.  visit(q, sexp) is called once each.
template <typename Visit>
void for_each_signature(const RootFactors& f, uint64_t limit, Visit&& visit) {
    uint32_t sexp[MAX_SPLIT] = {0};
    uint64_t qpow[MAX_SPLIT + 1];
    for (uint32_t j = 0; j <= f.ns; ++j) qpow[j] = 1;
    for (;;) {
        visit(qpow[f.ns], sexp);
        // increment: last position first
        int j = (int)f.ns - 1;
        while (j >= 0) {
            uint64_t nq = qpow[j + 1] * f.sp[j];
            if (nq <= limit) {
                ++sexp[j];
                qpow[j + 1] = nq;
                for (uint32_t t = (uint32_t)j + 1; t < f.ns; ++t) {
                    sexp[t] = 0;
                    qpow[t + 1] = qpow[t];
                }
                break;
            }
            --j;
        }
        if (j < 0) break;
    }
}

uint64_t left_size(const RootFactors& f, const uint32_t* sexp) {
    uint64_t W = 1;
    for (uint32_t j = 0; j < f.ns; ++j) W *= 4ull * (f.se[j] + sexp[j]) + 1;
    return (W + 1) / 2;
}

}  // namespace

uint32_t gaussian_weight(uint64_t l, const std::vector<uint32_t>& primes) {
    if (!l) return 0;
    RootFactors f;
    if (!factor_root(l, primes, f)) return 0;
    uint32_t w = 1;
    for (uint32_t j = 0; j < f.ns; ++j) w *= 4 * f.se[j] + 1;
    return w;
}

double estimate_position(uint64_t l, uint64_t limit, const std::vector<uint32_t>& small_primes, const uint32_t* off,
                         uint64_t n_max) {
    if (limit < 1 || limit > n_max || l < 1 || l > limit) return -1.0;
    RootFactors f;
    if (!factor_root(l, small_primes, f)) return -1.0;
    double est = 0;
    for_each_signature(f, limit, [&](uint64_t q, const uint32_t* sexp) {
        uint64_t kmin = (l + q - 1) / q, kmax = limit / q;
        if (kmin <= kmax) est += (double)left_size(f, sexp) * (double)entries(off, (uint32_t)kmin, (uint32_t)kmax);
    });
    return est;
}

bool split_by_work(uint64_t limit, uint64_t first, uint64_t last, uint32_t parts,
                   const std::vector<uint32_t>& small_primes, const uint32_t* off, uint64_t n_max, uint64_t* bounds,
                   std::string& err) {
    if (!parts || !bounds) {
        err = "need at least one part";
        return false;
    }
    if (limit < 1 || limit > n_max || first < 1 || first > last || last > limit) {
        err = "interval outside [1, limit] or limit outside [1, n_max]";
        return false;
    }
    const uint64_t n = last - first + 1;
    std::vector<double> work(n);
    std::atomic<bool> bad{false};
    const uint32_t chunks = (uint32_t)std::min<uint64_t>(256, (n + 255) / 256);
    parallel_chunks((uint32_t)n, chunks, [&](uint32_t b, uint32_t e, uint32_t) {
        for (uint32_t i = b; i < e; ++i) {
            work[i] = estimate_position(first + i, limit, small_primes, off, n_max);
            if (work[i] < 0) bad = true;
        }
    });
    if (bad) {
        err = "position cannot be planned";
        return false;
    }
    double total = 0;
    for (double w : work) total += w;
    bounds[0] = first;
    double acc = 0;
    uint64_t l = first;
    for (uint32_t i = 1; i < parts; ++i) {
        const double target = total * i / parts;
        while (l <= last && acc + work[l - first] < target) acc += work[l++ - first];
        // l is the first position whose inclusive prefix reaches the target
        bounds[i] = l <= last ? l + 1 : last + 1;
        if (bounds[i] < bounds[i - 1]) bounds[i] = bounds[i - 1];
    }
    bounds[parts] = last + 1;
    return true;
}

double tile_window_fraction(uint64_t l, uint64_t limit, uint64_t q, uint32_t k0, uint32_t k1, int samples) {
    const double l2 = (double)(l * l), l4 = l2 * l2;
    const double nd = (double)(limit * limit), n4 = nd * nd;
    double sum = 0;
    for (int i = 0; i < samples; ++i) {
        const uint64_t k = samples == 1 ? k0 : k0 + (uint64_t)(k1 - k0) * i / (samples - 1);
        const uint64_t m = q * k;
        const double m2 = (double)(m * m), m4 = m2 * m2;
        const double lo = (l4 + m4) - n4, tau = (4.0 * (l2 * m2) + n4) * 0x1p-44;
        Window w;
        make_window(lo, l4, tau, 2.0 * l2 * m2, w);
        sum += window_fraction(w);
    }
    return sum / samples;
}

bool plan_lefts(const PosReq* reqs, uint32_t n_roots, const std::vector<uint32_t>& small_primes, const uint32_t* off,
                uint64_t n_max, const Tuning& tune, Plan& P, std::string& err) {
    P = Plan();
    P.pos.resize(n_roots);
    // positions: factorisation, signatures and left-set recipes, in chunks
    struct PosChunk {
        std::vector<LeftDesc> lefts;
        std::vector<LeftBuild> builds;
        std::vector<uint32_t> kmin, kmax;
        uint64_t entries = 0;
        double cells = 0;
        std::string err;
    };
    const uint32_t pchunks = std::max(1u, std::min(64u, n_roots / 16));
    std::vector<PosChunk> pc(pchunks);
    parallel_chunks(n_roots, pchunks, [&](uint32_t b, uint32_t e, uint32_t c) {
        PosChunk& C = pc[c];
        for (uint32_t s = b; s < e && C.err.empty(); ++s) {
            const uint64_t l = reqs[s].l, limit = reqs[s].limit;
            if (limit < 1 || limit > n_max) {
                C.err = "limit outside [1, n_max]";
                break;
            }
            if (l < 1 || l > limit) {
                C.err = "root outside [1, limit]";
                break;
            }
            RootFactors f;
            if (!factor_root(l, small_primes, f)) {
                C.err = "root has too many distinct prime factors";
                break;
            }
            PosDesc& D = P.pos[s];
            D = PosDesc();
            D.l = l;
            D.limit = limit;
            const double nd = (double)(limit * limit);
            D.n4_d = nd * nd;
            const double l2 = (double)(l * l);
            D.l2_d = l2;
            D.l4_d = l2 * l2;
            D.l3 = (uint32_t)(l % 3);
            D.l5 = (uint32_t)(l % 5);
            D.l_even = (l & 1) ? 0u : 1u;
            D.nsplit = f.ns;
            D.ninert = f.ni;
            D.out = reqs[s].out;
            for (uint32_t j = 0; j < f.ns; ++j) {
                D.split_inv[j] = inv_mod_2_32(f.sp[j]);
                D.split_thr[j] = 0xFFFFFFFFu / f.sp[j];
            }
            for (uint32_t j = 0; j < f.ni; ++j) {
                D.inert_inv[j] = inv_mod_2_32(f.ip[j]);
                D.inert_thr[j] = 0xFFFFFFFFu / f.ip[j];
            }
            for_each_signature(f, limit, [&](uint64_t q, const uint32_t* sexp) {
                uint64_t kmin = (l + q - 1) / q, kmax = limit / q;
                if (kmin > kmax) return;
                LeftBuild B{};
                B.fixed = f.fixed;
                B.k = f.ns;
                for (uint32_t j = 0; j < f.ns; ++j) {
                    B.p[j] = f.sp[j];
                    B.E[j] = f.se[j] + sexp[j];
                }
                const uint64_t H = left_size(f, sexp);
                if (H > MAX_LEFT_ENTRIES) {
                    if (C.err.empty()) C.err = "left set larger than 2^17 - 1 entries";
                    return;
                }
                LeftDesc L{};
                L.pos = s;
                L.off = (uint32_t)C.entries;  // relative to the chunk until merged
                L.H = (uint32_t)H;
                L.q = (uint32_t)q;
                L.flags = left_flags(f.ns == 0, l);
                L.q3 = (uint32_t)(q % 3);
                L.q5 = (uint32_t)(q % 5);
                C.entries += L.H;
                C.lefts.push_back(L);
                C.builds.push_back(B);
                C.kmin.push_back((uint32_t)kmin);
                C.kmax.push_back((uint32_t)kmax);
                C.cells += (double)L.H * (double)entries(off, (uint32_t)kmin, (uint32_t)kmax);
            });
        }
    });
    // merge: each chunk's left sets go to their slice, entry offsets shifted
    // by the entries of the chunks before it
    std::vector<uint32_t> first(pchunks + 1, 0);
    std::vector<uint64_t> eoff(pchunks, 0);
    for (uint32_t c = 0; c < pchunks; ++c) {
        const PosChunk& C = pc[c];
        if (!C.err.empty()) {
            err = C.err;
            return false;
        }
        if (P.left_entries + C.entries > 0xFFFFFFFFull) {
            err = "batch too large: split it";
            return false;
        }
        eoff[c] = P.left_entries;
        P.left_entries += C.entries;
        P.est_cells += C.cells;
        first[c + 1] = first[c] + (uint32_t)C.lefts.size();
    }
    P.lefts.resize(first[pchunks]);
    P.builds.resize(first[pchunks]);
    P.kmin.resize(first[pchunks]);
    P.kmax.resize(first[pchunks]);
    parallel_chunks(pchunks, pchunks, [&](uint32_t b, uint32_t e, uint32_t) {
        for (uint32_t c = b; c < e; ++c) {
            const PosChunk& C = pc[c];
            for (size_t i = 0; i < C.lefts.size(); ++i) {
                LeftDesc L = C.lefts[i];
                L.off += (uint32_t)eoff[c];
                P.lefts[first[c] + i] = L;
            }
            std::copy(C.builds.begin(), C.builds.end(), P.builds.begin() + first[c]);
            std::copy(C.kmin.begin(), C.kmin.end(), P.kmin.begin() + first[c]);
            std::copy(C.kmax.begin(), C.kmax.end(), P.kmax.begin() + first[c]);
        }
    });
    // Window mode per left set: eligible by size, then per tile by the
    // window fraction of its partner range.  Work is estimated in plain-cell
    // units per partner entry: a plain tile H + plain_partner_cost (the cells
    // plus the flattening, partner loads and central leaf of each entry), a
    // window-mode tile H f win_item_cost + win_partner_cost.
    const uint32_t nl = (uint32_t)P.lefts.size();
    const uint32_t lchunks = std::max(1u, std::min(256u, nl / 64));
    std::vector<uint8_t>& eligible = P.eligible;
    std::vector<double>& ecost = P.ecost;  // cost per partner entry of the whole left set
    std::vector<double>& fmax = P.fmax;
    eligible.assign(nl, 0);
    ecost.assign(nl, 0.0);
    fmax.assign(nl, 0.0);
    parallel_chunks(nl, lchunks, [&](uint32_t b, uint32_t e, uint32_t) {
        for (uint32_t li = b; li < e; ++li) {
            const LeftDesc& L = P.lefts[li];
            const uint32_t H = L.H;
            const double plain = (double)H + tune.plain_partner_cost;
            ecost[li] = plain;
            const int wc = weight_class(H);
            if (wc < 0 || !(tune.win_max_frac[wc] > 0.0)) continue;
            eligible[li] = 1;
            fmax[li] = tune.win_max_frac[wc];
            const double f =
                tile_window_fraction(P.pos[L.pos].l, P.pos[L.pos].limit, L.q, P.kmin[li], P.kmax[li], 9);
            if (f <= fmax[li]) ecost[li] = std::min(plain, (double)H * f * tune.win_item_cost + tune.win_partner_cost);
        }
    });
    // every left set that window mode may use is stored sorted (its plain
    // tiles read the entries in that order just as well), so the sort can
    // start before the tiles are planned
    for (uint32_t li = 0; li < nl; ++li) {
        P.lefts[li].bt = NO_TABLE;
        if (eligible[li]) {
            P.lefts[li].bt = (uint32_t)P.sorted.size();
            P.sorted.push_back(li);
        }
    }
    return true;
}

bool plan_tiles(const uint32_t* off, uint64_t waves, const Tuning& tune, Plan& P) {
    const uint32_t nl = (uint32_t)P.lefts.size();
    const uint32_t lchunks = std::max(1u, std::min(256u, nl / 64));
    const std::vector<uint8_t>& eligible = P.eligible;
    const std::vector<double>& ecost = P.ecost;
    const std::vector<double>& fmax = P.fmax;
    P.tiles.clear();
    P.wtiles.clear();
    P.est_work = P.win_cells = 0;
    double total = 0;
    for (uint32_t li = 0; li < nl; ++li)
        total += (double)entries(off, P.kmin[li], P.kmax[li]) * ecost[li];
    // tile size: a few tiles per resident wavefront, bounded
    const uint32_t tpw = tune.tiles_per_wave ? tune.tiles_per_wave : 8;
    double target = total / (double)(waves ? waves * tpw : tpw);
    target = std::min(std::max(target, 8192.0), 4194304.0);
    // heaviest tiles first, per mode: a counting sort on the cost quantised to
    // 1/8 octaves (load balance needs the order only roughly; linear time)
    constexpr int KEYS = 8 * 64;
    auto key = [](double c) {
        const int k = c > 1.0 ? (int)(8.0 * std::log2(c)) : 0;
        return (uint32_t)(KEYS - 1 - std::min(k, KEYS - 1));  // descending cost
    };
    struct Cost {
        uint32_t key;  // sort key of the tile's cost
        uint32_t i;    // the tile in its chunk
        bool win;
    };
    struct TileChunk {
        std::vector<TileDesc> tiles;
        std::vector<Cost> costs;
        double win_cells = 0, work = 0;
    };
    std::vector<TileChunk> tc(lchunks);
    parallel_chunks(nl, lchunks, [&](uint32_t b, uint32_t e, uint32_t c) {
        TileChunk& C = tc[c];
        for (uint32_t li = b; li < e; ++li) {
            const LeftDesc& L = P.lefts[li];
            uint32_t k0 = P.kmin[li], kmax = P.kmax[li];
            uint64_t budget = (uint64_t)std::max(target / ecost[li], 256.0);
            while (k0 <= kmax) {
                // largest kend in [k0, kmax] with entries(k0, kend) <= budget, at least 128 indices;
                // entries(k0, .) grows with kend and a tile spans a small part
                // of [k0, kmax], so probe k0 + 127, k0 + 255, ... (cache lines
                // near k0) until one is over budget, then bisect the last gap
                uint32_t lo = k0, hi = kmax;
                for (uint64_t step = 128;; step *= 2) {
                    const uint64_t probe = (uint64_t)k0 + step - 1;
                    if (probe >= hi) break;
                    if (entries(off, k0, (uint32_t)probe) > budget) {
                        hi = (uint32_t)probe - 1;
                        break;
                    }
                    lo = (uint32_t)probe;
                }
                while (lo < hi) {
                    uint32_t mid = lo + (hi - lo + 1) / 2;
                    if (entries(off, k0, mid) <= budget)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                uint32_t kend = std::max(lo, std::min(kmax, k0 + 127u));
                const uint64_t ne = entries(off, k0, kend);
                const double cells = (double)L.H * (double)ne;
                // plain mode runs every row of the left set for each step of 64
                // partner entries, and each chunk of 64 partner indices of one
                // parity takes a step of its own: a tile of a few entries over a
                // large left set costs its steps (H rows each), not its cells
                const uint64_t per_parity = ((uint64_t)kend - k0 + 2) / 2;
                const uint64_t chunks = (P.pos[L.pos].l_even ? 1u : 2u) * ((per_parity + 63) / 64);
                const uint64_t steps = std::max((ne + 63) / 64, std::min(chunks, ne));
                double cost = (double)L.H * 64.0 * (double)steps + (double)ne * tune.plain_partner_cost;
                bool win = false;
                if (eligible[li]) {
                    const double f = tile_window_fraction(P.pos[L.pos].l, P.pos[L.pos].limit, L.q, k0, kend, 3);
                    if (f <= fmax[li]) {
                        win = true;
                        cost = cells * f * tune.win_item_cost +
                               (double)entries(off, k0, kend) * tune.win_partner_cost;
                        C.win_cells += cells;
                    }
                }
                C.work += cost;
                C.costs.push_back({key(cost), (uint32_t)C.tiles.size(), win});
                C.tiles.push_back(TileDesc{li, k0, kend, 0});
                if (kend == kmax) break;
                k0 = kend + 1;
            }
        }
    });
    // the counting sort per mode, placed straight from the chunks (within a
    // key: chunk order, then tile order)
    std::vector<uint32_t> start[2] = {std::vector<uint32_t>(KEYS + 1, 0), std::vector<uint32_t>(KEYS + 1, 0)};
    for (const TileChunk& C : tc) {
        for (const Cost& x : C.costs) ++start[x.win][x.key + 1];
        P.win_cells += C.win_cells;
        P.est_work += C.work;
    }
    for (int m = 0; m < 2; ++m)
        for (int k = 0; k < KEYS; ++k) start[m][k + 1] += start[m][k];
    P.tiles.resize(start[0][KEYS]);
    P.wtiles.resize(start[1][KEYS]);
    for (const TileChunk& C : tc)
        for (const Cost& x : C.costs) (x.win ? P.wtiles : P.tiles)[start[x.win][x.key]++] = C.tiles[x.i];
    return true;
}

bool plan_batch(const PosReq* reqs, uint32_t n_roots, const std::vector<uint32_t>& small_primes, const uint32_t* off,
                uint64_t n_max, uint64_t waves, const Tuning& tune, Plan& P, std::string& err) {
    return plan_lefts(reqs, n_roots, small_primes, off, n_max, tune, P, err) && plan_tiles(off, waves, tune, P);
}

}  // namespace hss
