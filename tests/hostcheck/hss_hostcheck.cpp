// hss_hostcheck.cpp -- CPU execution of the kernels' arithmetic, for tests only.
//
// TEST INFRASTRUCTURE.  The product (libhss.so, gpu_search, search) never links
// or loads this library.  It runs, sequentially on the host, exactly the
// per-entry and per-cell functions that the gfx950 kernels run (hss_build.h,
// hss_math.h) over the same plan (hss_plan.cpp): the half-set cache build, the
// left-set build, the cell evaluation with its FP64 range band and root
// proximity filter, and the exact stage.  Its counts must equal the oracle's;
// that checks the algorithm and the error bounds on the CPU, so that a GPU run
// only has to confirm the parallel mapping.
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "hss_build.h"
#include "hss_plan.h"
#include "hss_sched.h"

using namespace hss;

namespace {

struct HostTables {
    uint64_t n = 0;
    std::vector<uint32_t> spf, off;
    std::vector<double> cre, cim;
    std::vector<uint32_t> cres;
    std::vector<int64_t> pa, pb;  // per integer p (split primes only)
    std::vector<uint32_t> small;
};

void factor(const HostTables& T, uint32_t i, int64_t& fixed, uint32_t* ps, uint32_t* es, uint32_t& k) {
    fixed = 1;
    k = 0;
    uint32_t x = i;
    while (x > 1) {
        uint32_t p = T.spf[x], a = 0;
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
}

void build(HostTables& T, uint64_t n) {
    T.n = n;
    T.spf.assign(n + 1, 0);
    for (uint64_t i = 0; i <= n; ++i) T.spf[i] = (uint32_t)i;
    for (uint64_t p = 2; p * p <= n; ++p)
        if (T.spf[p] == p)
            for (uint64_t j = p * p; j <= n; j += p) T.spf[j] = std::min<uint32_t>(T.spf[j], (uint32_t)p);
    T.pa.assign(n + 1, 0);
    T.pb.assign(n + 1, 0);
    for (uint64_t p = 5; p <= n; p += 4)
        if (T.spf[p] == p) prime_split(p, T.pa[p], T.pb[p]);
    T.off.assign(n + 2, 0);
    for (uint64_t i = 1; i <= n; ++i) {
        int64_t fixed;
        uint32_t ps[MAX_SPLIT], es[MAX_SPLIT], k;
        factor(T, (uint32_t)i, fixed, ps, es, k);
        uint32_t w = 1;
        for (uint32_t j = 0; j < k; ++j) w *= 4 * es[j] + 1;
        T.off[i + 1] = T.off[i] + (w + 1) / 2;
    }
    uint32_t total = T.off[n + 1];
    T.cre.resize(total);
    T.cim.resize(total);
    T.cres.resize(total);
    for (uint64_t i = 1; i <= n; ++i) {
        int64_t fixed;
        uint32_t ps[MAX_SPLIT], es[MAX_SPLIT], k;
        factor(T, (uint32_t)i, fixed, ps, es, k);
        for (uint32_t e = T.off[i]; e < T.off[i + 1]; ++e) {
            uint32_t dig[MAX_SPLIT];
            decode_rank(e - T.off[i], k, es, dig);
            const uint32_t cls = digits_class(k, ps, dig);
            int64_t re = fixed, im = 0;
            for (uint32_t j = 0; j < k; ++j) {
                GInt v = gauss_value(T.pa[ps[j]], T.pb[ps[j]], es[j], dig[j]);
                int64_t a = (int64_t)v.re, b = (int64_t)v.im;
                int64_t nr = re * a - im * b, ni = re * b + im * a;
                re = nr;
                im = ni;
            }
            T.cre[e] = (double)re;
            T.cim[e] = (double)im;
            int64_t r1 = re % (int64_t)QA, i1 = im % (int64_t)QA;
            T.cres[e] = (uint32_t)(r1 < 0 ? r1 + QA : r1) | cls << CRES_CLASS_SHIFT |
                        (uint32_t)(i1 < 0 ? i1 + QA : i1) << 16;
        }
    }
    uint32_t root = 2;
    while ((uint64_t)root * root <= n) ++root;
    T.small = primes_upto(root);
}

HostTables g_tables;

}  // namespace

extern "C" {

typedef struct {
    uint64_t l, m, n, area4_lo, area4_hi;
} hc_hit;

typedef struct {
    uint64_t deferred_full;  // leaves left undecided by the FP64 band
    uint64_t deferred_root;  // counted terms near an exact fourth root
    uint64_t raw_roots;
    uint64_t cells;          // leaves evaluated
    uint64_t window_tiles;
    uint64_t window_items;   // leaves evaluated in window-mode tiles
    uint64_t class_mismatch; // leaves whose class-rule area decision differs from the residue test
    uint64_t window_units;   // units of WIN_UNIT entries the window kernel runs for those leaves
} hc_stats;

static Tuning g_tune;

// Window-mode knobs of the planner (the device context's hss_tuning): weight
// classes >= first_class get threshold frac, the others 0 (window mode off).
void hc_set_tuning(double frac, uint32_t first_class, uint32_t tiles_per_wave) {
    for (int i = 0; i < WEIGHT_CLASSES; ++i) g_tune.win_max_frac[i] = (uint32_t)i >= first_class ? frac : 0.0;
    g_tune.tiles_per_wave = tiles_per_wave;
}

// The planner's default tuning (the device context's defaults), or one window
// threshold per weight class (WEIGHT_CLASSES values).
void hc_set_tuning_classes(const double* fracs, uint32_t tiles_per_wave) {
    const Tuning d;
    for (int i = 0; i < WEIGHT_CLASSES; ++i) g_tune.win_max_frac[i] = fracs ? fracs[i] : d.win_max_frac[i];
    g_tune.tiles_per_wave = tiles_per_wave ? tiles_per_wave : d.tiles_per_wave;
}

// Search root positions (roots[s] at limits[s]) with the kernels' arithmetic,
// plain tiles as cell_kernel and window tiles as cell_win_kernel.  Counters
// are indexed by position.  Tables are rebuilt when nmax changes.  Returns 0
// on success.
int hc_search_mixed(uint64_t nmax, const uint64_t* roots, const uint64_t* limits, uint32_t n_roots, uint64_t* cand,
                    uint64_t* leaves, uint64_t* pairs, hc_hit* hits, uint32_t hit_cap, uint32_t* n_hits,
                    hc_stats* st) {
    if (g_tables.n != nmax) build(g_tables, nmax);
    const HostTables& T = g_tables;
    Plan P;
    std::string err;
    std::vector<PosReq> reqs(n_roots);
    for (uint32_t s = 0; s < n_roots; ++s) reqs[s] = PosReq{roots[s], limits[s], s};
    if (!plan_batch(reqs.data(), n_roots, T.small, T.off.data(), nmax, 4096, g_tune, P, err)) return 1;
    // left sets; sorted ones (window mode) ordered by psi with the central entry last
    std::vector<i128> lre(P.left_entries), lim(P.left_entries);
    std::vector<LeftFast> lfast(P.left_entries);
    std::vector<uint32_t> lcls(P.left_entries);  // class of each stored left entry
    std::vector<uint32_t> lres(P.left_entries);  // (re mod 180) | (im mod 180) << 16 (class-rule check only)
    std::vector<uint32_t> btab(P.sorted.size() * (WIN_NB + 1));
    for (size_t li = 0; li < P.lefts.size(); ++li) {
        const LeftBuild& b = P.builds[li];
        const LeftDesc& L = P.lefts[li];
        std::vector<GInt> val(L.H);
        std::vector<uint32_t> cls(L.H);
        for (uint32_t t = 0; t < L.H; ++t) {
            uint32_t dig[MAX_SPLIT];
            decode_rank(t, b.k, b.E, dig);
            GInt z = {(i128)b.fixed, 0};
            for (uint32_t j = 0; j < b.k; ++j) z = gmul(z, gauss_value(T.pa[b.p[j]], T.pb[b.p[j]], b.E[j], dig[j]));
            val[t] = z;
            cls[t] = digits_class(b.k, b.p, dig);
        }
        std::vector<uint32_t> perm(L.H);
        for (uint32_t t = 0; t < L.H; ++t) perm[t] = t;
        if (L.bt != NO_TABLE) {
            // k_left_sort: (area group, theta) keys, FP32 angles, group slices of the buckets
            // (the device sums the primes' angles over the digits instead of taking the
            // product's angle: the same to 1e-12 rad before the FP32 rounding; either
            // order gives exact counts, the window covers the key's error)
            const uint32_t n = L.H - 1, kmask = left_kmask(L), gbits = __builtin_popcount(kmask);
            const uint32_t nbg = WIN_NB >> gbits;
            std::vector<float> key(n);
            std::vector<uint32_t> bucket(n);
            for (uint32_t t = 0; t < n; ++t) {
                key[t] = (float)theta_of(i128_to_double(val[t].re), i128_to_double(val[t].im));
                bucket[t] = area_group(cls[t], kmask) * nbg + theta_bucket((double)key[t], nbg);
            }
            std::sort(perm.begin(), perm.begin() + n, [&](uint32_t x, uint32_t y) {
                return std::tie(bucket[x], key[x], x) < std::tie(bucket[y], key[y], y);
            });
            uint32_t* bt = btab.data() + (size_t)L.bt * (WIN_NB + 1);
            for (uint32_t bk = 0; bk <= WIN_NB; ++bk) {
                uint32_t c = 0;
                while (c < n && bucket[perm[c]] < bk) ++c;
                bt[bk] = bk == WIN_NB ? n : c;
            }
        }
        for (uint32_t t = 0; t < L.H; ++t) {
            const GInt& z = val[perm[t]];
            lre[L.off + t] = z.re;
            lim[L.off + t] = z.im;
            LeftFast f;
            f.re = i128_to_double(z.re);
            f.im = i128_to_double(z.im);
            f.cls = cls[perm[t]];
            f.pad = 0;
            lfast[L.off + t] = f;
            lcls[L.off + t] = cls[perm[t]];
            lres[L.off + t] = smod128(z.re, QA) | smod128(z.im, QA) << 16;
        }
    }
    std::fill(cand, cand + n_roots, 0);
    std::fill(leaves, leaves + n_roots, 0);
    std::fill(pairs, pairs + n_roots, 0);
    hc_stats S{};
    struct Def {
        uint32_t left, ui, vidx, k, conj, full, omask;
    };
    std::vector<Def> defs;
    // The class rule against the reference's area test on residues modulo 180
    // (src/search_v18.cpp:441-453: 4h | area4 / 2, h = (9 if 3 | l m, else 3) *
    // (5 if 5 divides neither, else 1)) for the leaf u * v or u * conj(v) (cj):
    // orientation 1 tests Im z, orientation 2 Re z.  A real leaf (both factors
    // central) has orientation-1 area 0 and is not compared there.
    auto check_class = [&](const LeftDesc& L, const PosDesc& D, uint32_t ui, uint32_t vidx, uint32_t k, uint32_t cj,
                           bool both_real) {
        const uint32_t ua = lres[L.off + ui] & 0xFFFFu, ub = lres[L.off + ui] >> 16;
        const uint32_t c180 = T.cres[vidx] & 0xFFu, d180 = T.cres[vidx] >> 16;
        const uint32_t pac = ua * c180, pbd = ub * d180, pad = ua * d180, pbc = ub * c180;
        const uint32_t vX = cj ? pac + pbd : pac + QA * QA - pbd;
        const uint32_t vY = cj ? pbc + QA * QA - pad : pad + pbc;
        const uint64_t m = (uint64_t)L.q * k;
        const bool need9 = D.l3 == 0 || m % 3 == 0, need5 = D.l5 != 0 && m % 5 != 0;
        const uint32_t h4 = 4u * (need9 ? 9u : 3u) * (need5 ? 5u : 1u);
        const uint32_t cm = area_constrained(need9, need5);
        const uint32_t c = (lcls[L.off + ui] ^ ((T.cres[vidx] >> CRES_CLASS_SHIFT) & 3u)) & cm;
        const bool even = D.l_even || !(k & 1u);
        const bool rule1 = c == 0u, rule2 = even && c == cm;
        if ((!both_real && rule1 != (vY % h4 == 0)) || rule2 != (vX % h4 == 0)) ++S.class_mismatch;
    };
    // a decided leaf's counted terms f (bits: orientation 1 term 1, 2 on |X|,
    // orientation 2 on |Y|) go to the root test; one survivor defers the leaf
    auto root_test = [&](uint32_t li, uint32_t ui, uint32_t vidx, uint32_t k, uint32_t cj, uint32_t f, double X,
                         double Y, double lm) {
        const double ax = std::fabs(X), ay = std::fabs(Y);
        bool pass = false;
        for (uint32_t t = 0; t < 4; ++t)
            pass |= ((f >> t) & 1u) && root_near_integer(__builtin_fma((t & 1u) ? 2.0 : -2.0, t < 2u ? ax : ay, lm));
        if (pass) {
            defs.push_back({li, ui, vidx, k, cj, 0, 3});
            ++S.deferred_root;
        }
    };
    for (int mode = 0; mode < 2; ++mode) {
        const std::vector<TileDesc>& tiles = mode == 0 ? P.tiles : P.wtiles;
        for (const TileDesc& Tl : tiles) {
            const LeftDesc& L = P.lefts[Tl.left];
            const PosDesc& D = P.pos[L.pos];
            S.window_tiles += mode;
            const bool one = tile_one_term(L.q, Tl.k0, D.l4_d, D.l2_d, D.n4_d);
            const uint32_t kmask = left_kmask(L);
            const uint32_t nbg = WIN_NB >> __builtin_popcount(kmask);
            const uint32_t* bt = L.bt != NO_TABLE ? btab.data() + (size_t)L.bt * (WIN_NB + 1) : nullptr;
            for (uint32_t k = Tl.k0; k <= Tl.k1; ++k) {
                bool ok = !(D.l_even && !(k & 1u));
                for (uint32_t j = 0; j < D.nsplit; ++j) ok = ok && !(k * D.split_inv[j] <= D.split_thr[j]);
                for (uint32_t j = 0; j < D.ninert; ++j) ok = ok && !(k * D.inert_inv[j] <= D.inert_thr[j]);
                if (!ok) continue;
                uint32_t Hk = T.off[k + 1] - T.off[k];
                if ((L.flags & LEFT_TRIVIAL) && Hk == 1) continue;
                pairs[D.out] += 1;
                const uint64_t m = (uint64_t)L.q * k;
                const double m2 = (double)(m * m);
                const Band bd = make_band((double)m, D.l4_d, D.l2_d, D.n4_d);
                const bool need9 = D.l3 == 0 || m % 3 == 0, need5 = D.l5 != 0 && m % 5 != 0;
                const uint32_t cm = area_constrained(need9, need5);
                const bool even = D.l_even || !(k & 1u);
                // both kernels run the entries of an even pair with cm = 0 twice,
                // orientation 1 (force 1) then 2 (force 2)
                const bool twice = cm == 0u && even;
                for (uint32_t run = 0; run < (twice ? 2u : 1u); ++run)
                for (uint32_t jj = 0; jj < Hk; ++jj) {
                    const uint32_t force = twice ? run + 1 : 0u;
                    const uint32_t vidx = T.off[k] + jj;
                    const bool vc = jj + 1 == Hk;
                    const double c = T.cre[vidx], d = T.cim[vidx];
                    const uint32_t vcls = (T.cres[vidx] >> CRES_CLASS_SHIFT) & 3u;
                    if (run == 0) leaves[D.out] += vc ? L.H : 2ull * L.H - 1;
                    {
                        // central_win: the central u by the class rule (both kernels)
                        ++S.cells;
                        if (run == 0) check_class(L, D, L.H - 1, vidx, k, 0, vc);
                        const double cu = lfast[L.off + L.H - 1].re;
                        const uint32_t cc = vcls & cm;
                        const uint32_t o = force ? force : (cc == 0u ? 1u : ((even && cc == cm) ? 2u : 0u));
                        const double Dc = cu * (o == 2u ? d : c);
                        const bool real2 = o == 2u && vc;
                        const double h2n = bd.h2_w - bd.tau;
                        bool u1;
                        const bool live = o != 0u && !vc;
                        const uint32_t f1 = one ? o1_decide<true>(live, Dc, bd, u1) : o1_decide<false>(live, Dc, bd, u1);
                        const bool und = u1 || (real2 && bd.h2_w >= 0.0 && h2n < 0.0);
                        if (und) {
                            defs.push_back({Tl.left, L.H - 1, vidx, k, 0, 1, o});
                            ++S.deferred_full;
                        } else {
                            cand[D.out] += __builtin_popcount(f1) + ((real2 && h2n >= 0.0) ? 1u : 0u);
                            root_test(Tl.left, L.H - 1, vidx, k, 0, f1, Dc, 0.0, bd.lm);
                        }
                    }
                    if (mode == 0) {
                        // cell_kernel rows: every non-central u, leaves u*v and (v not real)
                        // u*conj(v), each in the one orientation its class allows (or the
                        // run's forced one): D = Re(u w), Re(u conj(w)) with w = v or -i v
                        const uint32_t cmsk = force ? 0u : cm;
                        const uint32_t t1 = force == 2u ? 4u : 0u;
                        const uint32_t t2 = force == 2u ? 0u : ((force == 0u && even && cm != 0u) ? cm : 4u);
                        for (uint32_t ui = 0; ui + 1 < L.H; ++ui) {
                            const LeftFast& u = lfast[L.off + ui];
                            const uint32_t cr = (u.cls ^ vcls) & cmsk;
                            const bool is2 = cr == t2, live = cr == t1 || is2;
                            if (run == 0)
                                for (uint32_t cj = 0; cj < (vc ? 1u : 2u); ++cj) check_class(L, D, ui, vidx, k, cj, false);
                            const double ce = is2 ? d : c, de = is2 ? -c : d;
                            const double tt = u.im * de;
                            for (uint32_t cj = 0; cj < (vc ? 1u : 2u); ++cj) {
                                ++S.cells;
                                if (!live) continue;
                                const double Dw = cj ? __builtin_fma(u.re, ce, tt) : __builtin_fma(u.re, ce, -tt);
                                bool und;
                                const uint32_t f = one ? o1_decide<true>(true, Dw, bd, und) : o1_decide<false>(true, Dw, bd, und);
                                if (und) {
                                    defs.push_back({Tl.left, ui, vidx, k, cj, 1, is2 ? 2u : 1u});
                                    ++S.deferred_full;
                                    continue;
                                }
                                cand[D.out] += __builtin_popcount(f);
                                root_test(Tl.left, ui, vidx, k, cj, f, Dw, 0.0, bd.lm);
                            }
                        }
                        continue;
                    }
                    // cell_win_kernel: passes over (class variant j, orientation rho)
                    Window1 w;
                    make_window_o1((D.l4_d + m2 * m2) - D.n4_d, D.l4_d, bd.tau, 2.0 * D.l2_d * m2, w);
                    Window1F wf;
                    window_o1_fx(w, wf);
                    const uint32_t thv = theta_fx_of_f(c, d);
                    const uint32_t bsh = 32u - (uint32_t)__builtin_ctz(nbg);
                    for (uint32_t j = 0; j < allowed_count(cm, kmask); ++j)
                        for (uint32_t rho = force ? force - 1 : 0u; rho < (force ? force : (even ? 2u : 1u)); ++rho) {
                            const uint32_t g = area_group(allowed_class(vcls ^ (rho ? 3u : 0u), cm, kmask, j), kmask);
                            const uint32_t thw = thv + (rho ? ANGLE_HALF_PI : 0u);
                            const double ce = rho ? -d : c, de0 = rho ? c : d;
                            for (uint32_t cj = 0; cj < (vc ? 1u : 2u); ++cj) {
                                uint32_t rs[4], re[4];
                                const BucketGroup G{bt, g * nbg, bsh, bt[g * nbg], bt[g * nbg + nbg]};
                                window_o1_ranges(wf, G, thw, cj != 0, rs, re);
                                const double de = cj ? -de0 : de0;
                                for (int p = 0; p < 4; ++p) S.window_units += (re[p] - rs[p] + WIN_UNIT - 1) / WIN_UNIT;
                                for (int p = 0; p < 4; ++p)
                                    for (uint32_t ui = rs[p]; ui < re[p]; ++ui) {
                                        ++S.window_items;
                                        ++S.cells;
                                        const LeftFast& u = lfast[L.off + ui];
                                        const double X = __builtin_fma(u.re, ce, -(u.im * de));
                                        bool und;
                                        const uint32_t f = one ? o1_decide<true>(true, X, bd, und)
                                                               : o1_decide<false>(true, X, bd, und);
                                        // every leaf in a range passes its orientation's area test
                                        {
                                            const uint32_t cr = (lcls[L.off + ui] ^ vcls) & cm;
                                            if (cr != (rho ? cm : 0u)) ++S.class_mismatch;
                                        }
                                        if (und) {
                                            defs.push_back({Tl.left, ui, vidx, k, cj, 1, rho ? 2u : 1u});
                                            ++S.deferred_full;
                                            continue;
                                        }
                                        cand[D.out] += __builtin_popcount(f);
                                        root_test(Tl.left, ui, vidx, k, cj, f, X, 0.0, bd.lm);
                                    }
                            }
                        }
                }
            }
        }
    }
    // exact stage + acceptance
    QR4 qr = make_qr4();
    std::set<std::tuple<uint64_t, uint64_t, uint64_t>> seen;
    std::vector<hc_hit> got;
    for (const Def& df : defs) {
        const LeftDesc& L = P.lefts[df.left];
        const PosDesc& D = P.pos[L.pos];
        uint64_t m = (uint64_t)L.q * df.k;
        ExactCtx e;
        make_exact_ctx(D.l, m, D.limit, e);
        i128 zre, zim;
        exact_product(lre[L.off + df.ui], lim[L.off + df.ui], (int64_t)T.cre[df.vidx], (int64_t)T.cim[df.vidx],
                      df.conj != 0, zre, zim);
        uint32_t c = exact_leaf(zre, zim, e, qr, [&](uint64_t r, u128 a4) {
            ++S.raw_roots;
            if (accept_hit(D.l, m, r, D.limit, a4) && seen.insert({D.l, m, r}).second)
                got.push_back(hc_hit{D.l, m, r, (uint64_t)a4, (uint64_t)(a4 >> 64)});
        }, df.full ? df.omask : 3u);
        if (df.full) cand[D.out] += c;
    }
    std::sort(got.begin(), got.end(),
              [](const hc_hit& a, const hc_hit& b) { return std::tie(a.l, a.m, a.n) < std::tie(b.l, b.m, b.n); });
    for (uint32_t i = 0; i < got.size() && i < hit_cap; ++i) hits[i] = got[i];
    *n_hits = (uint32_t)got.size();
    if (st) *st = S;
    return 0;
}

// All positions at one limit.
int hc_search(uint64_t nmax, uint64_t limit, const uint64_t* roots, uint32_t n_roots, uint64_t* cand,
              uint64_t* leaves, uint64_t* pairs, hc_hit* hits, uint32_t hit_cap, uint32_t* n_hits, hc_stats* st) {
    std::vector<uint64_t> limits(n_roots, limit);
    return hc_search_mixed(nmax, roots, limits.data(), n_roots, cand, leaves, pairs, hits, hit_cap, n_hits, st);
}

// The class rule (hss_math.h, value_class in hss_build.h) on `samples` random
// products z of values pi^r conj(pi)^(4E - r) of split primes p != 5 (2 or 3
// factors, E <= 2; |z| < 2^88): Re z odd, 4 | Im z, and exactly one of Re z, Im z divisible
// by 3 -- Im iff class bit 0 is clear -- and by 5 -- Im iff class bit 1 is
// clear.  Also z times the conjugate of a product w of two values (E = 1)
// has the xor of their classes (|z conj w| < 2^118).  Returns the number of violated statements.
uint64_t hc_class_rule_mismatches(uint64_t samples, uint64_t seed) {
    static const uint64_t ps[] = {13, 17, 29, 37, 41, 53, 61, 73, 89, 97, 101, 109, 113, 137, 149, 157};
    constexpr uint32_t NP = sizeof(ps) / sizeof(ps[0]);
    int64_t pa[NP], pb[NP];
    for (uint32_t i = 0; i < NP; ++i) prime_split(ps[i], pa[i], pb[i]);
    uint64_t x = seed | 1;
    auto next = [&](uint32_t mod) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return (uint32_t)(x % mod);
    };
    auto mod = [](i128 v, int64_t q) { return (int64_t)(v % q + (v % q < 0 ? q : 0)); };
    uint64_t bad = 0;
    for (uint64_t t = 0; t < samples; ++t) {
        GInt z{1, 0}, w{1, 0};
        uint32_t cz = 0, cw = 0;
        const uint32_t nf = 2 + next(2);
        for (uint32_t f = 0; f < nf; ++f) {
            const uint32_t i = next(NP), E = 1 + next(2), r = next(4 * E + 1);
            z = gmul(z, gauss_value(pa[i], pb[i], E, r));
            cz ^= value_class(ps[i], r);
        }
        for (uint32_t f = 0; f < 2; ++f) {
            const uint32_t i = next(NP), r = next(5);
            w = gmul(w, gauss_value(pa[i], pb[i], 1, r));
            cw ^= value_class(ps[i], r);
        }
        bad += mod(z.re, 2) != 1;
        bad += mod(z.im, 4) != 0;
        const bool re3 = mod(z.re, 3) == 0, im3 = mod(z.im, 3) == 0;
        bad += re3 == im3;
        bad += im3 != ((cz & AREA_C3) == 0);
        const bool re5 = mod(z.re, 5) == 0, im5 = mod(z.im, 5) == 0;
        bad += re5 == im5;
        bad += im5 != ((cz & AREA_C5) == 0);
        // z * conj(w): class cz ^ cw
        const GInt y = gmul(z, GInt{w.re, -w.im});
        const uint32_t cy = cz ^ cw;
        bad += (mod(y.im, 3) == 0) != ((cy & AREA_C3) == 0);
        bad += (mod(y.im, 5) == 0) != ((cy & AREA_C5) == 0);
        bad += (mod(y.re, 3) == 0) == (mod(y.im, 3) == 0);
    }
    return bad;
}

// The device's fourth-root candidate with its two hardware reciprocal square
// roots replaced by correctly rounded ones carrying relative errors dt, dr:
// the refined r (root4_refine), and whether root_candidate_ok accepts it.
double hc_root_refined(double zz, double dt, double dr) {
    const double t = (1.0 / std::sqrt(zz)) * (1.0 + dt);
    const double r = (1.0 / std::sqrt(t)) * (1.0 + dr);
    return root4_refine(zz, t, r);
}

int hc_root_candidate_ok(double zz, double r) { return root_candidate_ok(zz, r) ? 1 : 0; }

// Partners k in [k0, k1] of root l at limit (signature factor q) whose pair
// context has h2_w >= 0 although tile_one_term(q, k0) holds: the kernels then
// skip second-sign terms that could count.  Must be 0.
// With any = 1: partners in [k0, k1] with h2_w >= 0, whatever the rule says.
uint64_t hc_one_term_violations(uint64_t l, uint64_t limit, uint32_t q, uint32_t k0, uint32_t k1, int any) {
    const double nd = (double)(limit * limit), l2 = (double)(l * l);
    const double n4 = nd * nd, l4 = l2 * l2;
    if (!any && !tile_one_term(q, k0, l4, l2, n4)) return 0;
    uint64_t bad = 0;
    for (uint32_t k = k0; k <= k1; ++k) {
        bad += make_band((double)q * (double)k, l4, l2, n4).h2_w >= 0.0;
    }
    return bad;
}

// Smallest k with tile_one_term(q, k) for root l at limit (k_max + 1 if none).
uint32_t hc_one_term_first(uint64_t l, uint64_t limit, uint32_t q, uint32_t k_max) {
    const double nd = (double)(limit * limit), l2 = (double)(l * l);
    const double n4 = nd * nd, l4 = l2 * l2;
    uint32_t lo = 1, hi = k_max + 1;  // tile_one_term is monotone in k
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (tile_one_term(q, mid, l4, l2, n4)) hi = mid; else lo = mid + 1;
    }
    return lo;
}

// Mean range-window fraction of root l over partners q*k, k in [k0, k1].
double hc_window_fraction(uint64_t l, uint64_t limit, uint64_t q, uint32_t k0, uint32_t k1) {
    return tile_window_fraction(l, limit, q, k0, k1, 33);
}

// The planner's split of [first, last] into parts of about equal work
// (bounds: parts + 1 values); 0 on success.
int hc_split_by_work(uint64_t nmax, uint64_t limit, uint64_t first, uint64_t last, uint32_t parts, uint64_t* bounds) {
    if (g_tables.n != nmax) build(g_tables, nmax);
    std::string err;
    return split_by_work(limit, first, last, parts, g_tables.small, g_tables.off.data(), nmax, bounds, err) ? 0 : 1;
}

// The planner's tiles of a batch of positions at one limit, planned for
// `waves` resident waves under the current tuning: the tiles of every left
// set (in either list) must cover its partner range [kmin, kmax] exactly once.
// out: [0] plain tiles, [1] window tiles, [2] left sets, [3] left sets whose
// tiles leave a gap, overlap or run past the range, [4] tiles with k1 < k0.
// 0 on success (the plan was made).
int hc_plan_cover(uint64_t nmax, uint64_t limit, const uint64_t* roots, uint32_t n_roots, uint64_t waves,
                  uint64_t* out) {
    if (g_tables.n != nmax) build(g_tables, nmax);
    std::vector<PosReq> reqs(n_roots);
    for (uint32_t s = 0; s < n_roots; ++s) reqs[s] = PosReq{roots[s], limit, s};
    Plan P;
    std::string err;
    if (!plan_batch(reqs.data(), n_roots, g_tables.small, g_tables.off.data(), nmax, waves, g_tune, P, err)) return 1;
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> by_left(P.lefts.size());
    uint64_t reversed = 0;
    for (const std::vector<TileDesc>* list : {&P.tiles, &P.wtiles})
        for (const TileDesc& t : *list) {
            reversed += t.k1 < t.k0 ? 1 : 0;
            by_left[t.left].push_back({t.k0, t.k1});
        }
    uint64_t bad = 0;
    for (size_t li = 0; li < P.lefts.size(); ++li) {
        std::vector<std::pair<uint32_t, uint32_t>>& r = by_left[li];
        std::sort(r.begin(), r.end());
        bool ok = !r.empty() && r.front().first == P.kmin[li] && r.back().second == P.kmax[li];
        for (size_t i = 1; ok && i < r.size(); ++i) ok = r[i].first == r[i - 1].second + 1;
        bad += ok ? 0 : 1;
    }
    out[0] = P.tiles.size();
    out[1] = P.wtiles.size();
    out[2] = P.lefts.size();
    out[3] = bad;
    out[4] = reversed;
    return 0;
}

// Planner work estimate of one position (cells), or -1.
double hc_estimate(uint64_t nmax, uint64_t l, uint64_t limit) {
    if (g_tables.n != nmax) build(g_tables, nmax);
    return estimate_position(l, limit, g_tables.small, g_tables.off.data(), nmax);
}

// Cache entry e of the host tables (for cross-checks of the table layout).
int hc_cache_entry(uint64_t nmax, uint32_t i, uint32_t t, double* re, double* im, uint32_t* r180, uint32_t* H) {
    if (g_tables.n != nmax) build(g_tables, nmax);
    const HostTables& T = g_tables;
    if (i < 1 || i > nmax) return 1;
    *H = T.off[i + 1] - T.off[i];
    if (t >= *H) return 1;
    *re = T.cre[T.off[i] + t];
    *im = T.cim[T.off[i] + t];
    *r180 = T.cres[T.off[i] + t];
    return 0;
}

// ---- request scheduler of the served form (hss_sched.h), driven from tests

void* hc_sched_new(uint32_t max_slots, uint64_t pool_positions, double launch_work, uint32_t max_launch_positions) {
    SchedConfig c;
    c.max_slots = max_slots;
    c.pool_positions = pool_positions;
    c.launch_work = launch_work;
    c.max_launch_positions = max_launch_positions;
    return new Sched(c);
}

void hc_sched_free(void* h) { delete static_cast<Sched*>(h); }

int hc_sched_enqueue(void* h, uint64_t id, uint32_t npos) { return static_cast<Sched*>(h)->enqueue(id, npos) ? 1 : 0; }

// admitted ids into out[0..cap); returns their number
uint32_t hc_sched_admit(void* h, uint64_t* out, uint32_t cap) {
    std::vector<uint64_t> a = static_cast<Sched*>(h)->admit();
    for (uint32_t i = 0; i < a.size() && i < cap; ++i) out[i] = a[i];
    return (uint32_t)a.size();
}

void hc_sched_set_costs(void* h, uint64_t id, const double* cost, uint32_t n) {
    static_cast<Sched*>(h)->set_costs(id, std::vector<double>(cost, cost + n));
}

// segments of the next launch; returns their number
uint32_t hc_sched_next(void* h, uint64_t* ids, uint32_t* begin, uint32_t* end, uint32_t cap) {
    std::vector<Segment> s = static_cast<Sched*>(h)->next_launch();
    for (uint32_t i = 0; i < s.size() && i < cap; ++i) {
        ids[i] = s[i].id;
        begin[i] = s[i].begin;
        end[i] = s[i].end;
    }
    return (uint32_t)s.size();
}

uint32_t hc_sched_finished(void* h, const uint64_t* ids, const uint32_t* begin, const uint32_t* end, uint32_t n,
                           uint64_t* out, uint32_t cap) {
    std::vector<Segment> s(n);
    for (uint32_t i = 0; i < n; ++i) s[i] = Segment{ids[i], begin[i], end[i]};
    std::vector<uint64_t> d = static_cast<Sched*>(h)->finished(s);
    for (uint32_t i = 0; i < d.size() && i < cap; ++i) out[i] = d[i];
    return (uint32_t)d.size();
}

void hc_sched_retire(void* h, uint64_t id) { static_cast<Sched*>(h)->retire(id); }
int hc_sched_cancel(void* h, uint64_t id) { return static_cast<Sched*>(h)->cancel(id); }
uint32_t hc_sched_pool_offset(void* h, uint64_t id) { return static_cast<Sched*>(h)->pool_offset(id); }
uint32_t hc_sched_prefix(void* h, uint64_t id) { return static_cast<Sched*>(h)->prefix(id); }
uint64_t hc_sched_pool_free(void* h) { return static_cast<Sched*>(h)->pool_free(); }
uint32_t hc_sched_queued(void* h) { return (uint32_t)static_cast<Sched*>(h)->queued(); }
uint32_t hc_sched_active(void* h) { return (uint32_t)static_cast<Sched*>(h)->active(); }

}  // extern "C"
