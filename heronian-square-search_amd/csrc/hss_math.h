// hss_math.h -- arithmetic shared by the gfx950 kernels and the host-side checker.
//
// Problem recap (docs/ALGORITHM.en.md:25-107 of the reference).  For root sides
// l <= m and LIMIT N every Gaussian integer z = D + 4A*i of norm (2 l^2 m^2)^2 is
// enumerated once per conjugation class; for each orientation (d, area4) in
// {(Re z, Im z), (Im z, Re z) unless Re z = +-Im z} and each sign of d the
// reference (src/search_v18.cpp:436-479) counts a "candidate" when
//     area4 != 0, area4 = 0 (mod 8), lo <= d <= hi  (lo = l^4+m^4-N^4, hi = l^4),
//     |area4| = 0 (mod 3), 3 | lm  => |area4| = 0 (mod 9),
//     5 divides neither l nor m   => |area4| = 0 (mod 5),
// (the negative sign only when d != 0), and then accepts the triple when
// zz = l^4 + m^4 - d passes the fourth-power residue tests modulo 16, 625, 9, 7,
// 13, 17, is an exact fourth power n^4, gcd(l, m, n) = 1, 5 | l*m*n and
// m <= n <= N.
//
// Fast path.  Every z is 2*u*v (or 2*u*conj(v)) with u from a "left set" of
// the root l and v from the half-set of a partner k; u and v carry FP64
// coordinates and a two-bit area class.  Writing z/2 = X + iY:
//   * the area tests need X or Y modulo 4 and modulo h = (3 or 9)*(1 or 5),
//     decided exactly from the classes of u and v and the parity of l m (the
//     class rule below);
//   * the range test is decided from FP64 estimates whose error is provably
//     below tau/16; an estimate within tau of a bound is "undecided" and the
//     whole leaf goes to the exact stage;
//   * a counted term's zz is estimated in FP64 (relative error < 2^-49 because
//     zz >= m^4); unless it is that close to a fourth power n^4 it cannot be
//     one, so only such terms reach the exact stage (root_near_integer: on the
//     host a fourth root within ROOT_EPS of an integer, on the device the
//     single candidate n with |n^4 - zz| <= 2^-47 zz).
// The exact stage (exact_leaf) restates the reference chain with 128-bit
// integers and is the only place where residues mod 16/625/9/7/13/17, the
// integer fourth root and the acceptance rules are evaluated.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define HSS_HD __host__ __device__
#else
#define HSS_HD
#endif

namespace hss {

typedef __int128 i128;
typedef unsigned __int128 u128;

static constexpr uint32_t QA = 180;        // area residue modulus 4 * 9 * 5
// Root proximity band.  For a counted term zz >= m^4 >= (l^4 + m^4)/2; the FP64
// estimate carries |err| <= 2^-52 (l^4+m^4) [l^4 + m^4] + 2^-50 l^2 m^2 [2|D|,
// left values possibly rounded] + 2^-53 zz [subtraction] < 2^-49.3 zz, so the
// fourth root is off by < 2^-51.3 n plus two sqrt roundings: < 3e-8 for
// n <= 2^26.  1e-6 keeps a 30x margin; random zz pass with probability 2e-6.
// (Host test; the device form in root_near_integer has its own bound.)
static constexpr double ROOT_EPS = 1e-6;

// ------------------------------------------------------------------ residues

HSS_HD inline uint32_t two64_mod(uint32_t q) {
    uint64_t r = ((uint64_t)1 << 32) % q;
    return (uint32_t)((r * r) % q);
}

// residue of a non-negative value below 2^128 modulo a small modulus q
HSS_HD inline uint32_t umod128(u128 v, uint32_t q) {
    uint64_t hi = (uint64_t)(v >> 64), lo = (uint64_t)v;
    uint64_t t = (hi % q) * two64_mod(q) + lo % q;
    return (uint32_t)(t % q);
}

// residue in [0, q) of a signed 128-bit value
HSS_HD inline uint32_t smod128(i128 v, uint32_t q) {
    if (v >= 0) return umod128((u128)v, q);
    uint32_t r = umod128((u128)(-v), q);
    return r ? q - r : 0;
}

// fourth-power residue sets as bit masks: bit r set iff r = x^4 (mod q)
struct QR4 {
    uint32_t m625[20];
    uint32_t m9, m7, m13, m17;
};

inline QR4 make_qr4() {
    QR4 q{};
    for (uint32_t x = 0; x < 625; ++x) {
        uint32_t r = (x * x % 625) * (x * x % 625) % 625;
        q.m625[r >> 5] |= 1u << (r & 31);
    }
    auto small = [](uint32_t m) {
        uint32_t b = 0;
        for (uint32_t x = 0; x < m; ++x) b |= 1u << ((x * x % m) * (x * x % m) % m);
        return b;
    };
    q.m9 = small(9);
    q.m7 = small(7);
    q.m13 = small(13);
    q.m17 = small(17);
    return q;
}

HSS_HD inline bool bit_of(const uint32_t* words, uint32_t i) { return (words[i >> 5] >> (i & 31)) & 1u; }

// ------------------------------------------------- per-(l, m) fast-path state

// The range test of a pair (l, m) as a band on |D| for D = X or Y of a leaf
// (the halved bounds on |2D|: halving is exact), with lo = l^4 + m^4 - N^4 and
// hi = l^4:
//   term 1 (d = +|2D|): lo+ <= |2D| <= hi, i.e. ||D| - mid| <= half with
//     mid = (lo+ + hi) / 4, half = (hi - lo+) / 4, lo+ = max(lo, 0);
//   term 2 (d = -|2D|): |2D| <= -lo.
// Both are widened by tau / 2 (half_w, h2_w); the narrowed bounds are w - tau.
// An estimate between the two is undecided.  mid, half and the distance add at
// most five roundings of 2^-52 N^4 <= tau / 256 each to the tau / 16 budget,
// still far inside the tau / 2 band.
struct Band {
    double mid, half_w, h2_w;
    double tau;  // undecided band: FP64 errors stay below tau / 16
    double lm;   // FP64 value of l^4 + m^4 (the zz estimates are lm -+ 2|D|)
};

// FP64 value of l^4 + m^4
HSS_HD inline double lm_of(uint64_t m, double l4_d) {
    const double md = (double)(uint32_t)m;  // exact: m < 2^26
    const double m2 = md * md;
    return l4_d + m2 * m2;
}

// the same for m = q * k given as factors (their FP64 product is exact: m < 2^26)
HSS_HD inline double lm_of(uint32_t q, uint32_t k, double l4_d) {
    const double md = (double)q * (double)k;
    const double m2 = md * md;
    return l4_d + m2 * m2;
}

// The band of the pair (l, m): l4_d, n4_d the FP64 values of l^4 and N^4,
// l2_d = l^2, md = m as an FP64 value (exact: m < 2^26).
HSS_HD inline Band make_band(double md, double l4_d, double l2_d, double n4_d) {
    const double m2 = md * md;  // exact: m^2 < 2^52
    const double m4 = m2 * m2;
    const double lo = (l4_d + m4) - n4_d;
    const double hi = l4_d;
    const double tau = (4.0 * (l2_d * m2) + n4_d) * 0x1p-44;
    const double lop = lo > 0.0 ? lo : 0.0;
    Band b;
    b.mid = (lop + hi) * 0.25;
    b.half_w = (hi - lop) * 0.25 + tau * 0.5;
    b.h2_w = (-lo + tau) * 0.5;
    b.tau = tau;
    b.lm = l4_d + m4;
    return b;
}

// Do all pairs (l, q k), k >= k0, have lo > tau?  Then h2_w = (tau - lo)/2 < 0
// and the second-sign terms (term2 of both orientations: d = -|2D|, i.e. zz =
// l^4 + m^4 + |2D| <= N^4) can neither count nor be undecided for any leaf.
// lo - tau grows with m (by 4m^3 - 2^-41 l^2 m > 0 per unit), so lo > 2 tau at
// m0 = q k0 leaves a margin of tau (far above the FP64 roundings of lo and
// tau) for every larger m.
HSS_HD inline bool tile_one_term(uint32_t q, uint32_t k0, double l4_d, double l2_d, double n4_d) {
    const double md = (double)q * (double)k0;  // exact: below 2^26
    const double m2 = md * md;
    const double lo = (l4_d + m2 * m2) - n4_d;
    const double tau = (4.0 * (l2_d * m2) + n4_d) * 0x1p-44;
    return lo > 2.0 * tau;
}

// can an FP64 estimate zz (relative error < 2^-49, zz >= 1) be an exact fourth power?
// One Newton step on r^4 = zz from r ~ zz^(1/4) and t ~ zz^(-1/2), with the
// derivative 1/(4 r^3) taken as r t^2 / 4.  With relative errors e of r and
// e_t of t the result is off by 2.5 e^2 + 2 e e_t + O(e^3) plus a few
// roundings (r'/rho = (1+e) [1 - (1+e_t)^2 ((1+e)^4 - 1)/4]).
HSS_HD inline double root4_refine(double zz, double t, double r) {
    const double w = 0.25 * (t * t);
    const double r2 = r * r;
    const double d = __builtin_fma(r2, r2, -zz);  // r^4 - zz
    return __builtin_fma(-d, r * w, r);
}

// Is zz within tol of the fourth power of n = rint(r)?  n^2 is exact (n <=
// 2^26) and the fma rounds once.  With tol = 2^-47 Z for any Z >= zz, a true
// fourth power n^4 estimated with |zz - n^4| < 2^-49.3 n^4 passes; a random zz
// passes with probability ~2^-48 n Z / zz.  The kernels take Z per pair
// (l^4 + m^4 bounds the first-sign terms zz = l^4 + m^4 - 2|D|, 2 N^4 the
// counted second-sign terms), which saves scaling zz per term.
HSS_HD inline bool root_candidate_ok(double zz, double r, double tol) {
    const double n = __builtin_rint(r);
    const double n2 = n * n;
    const double e = __builtin_fma(n2, n2, -zz);
    return __builtin_fabs(e) <= tol;
}

HSS_HD inline bool root_candidate_ok(double zz, double r) { return root_candidate_ok(zz, r, zz * 0x1p-47); }

// tol as in root_candidate_ok (device); the host form ignores it
HSS_HD inline bool root_near_integer(double zz, double tol) {
#if defined(__HIP_DEVICE_COMPILE__)
    // zz^(1/4) from two FP32 hardware reciprocal square roots (t = zz^-1/2 of
    // zz rounded to FP32, then t^-1/2; zz <= 2^104 is in FP32 range): the
    // rounding (2^-24) and the instruction's 1-ulp bound (2^-23) keep each
    // within 2^-22 relative, so r starts within 1.5 * 2^-22; one root4_refine
    // step in FP64 leaves 2.5 (1.5 * 2^-22)^2 + 3 * 2^-44 < 2^-40.8, i.e.
    // < 4e-5 absolute for n <= 2^26, and rint(r) is the only candidate n.
    // (Even with errors 2^6 times worse the step stays within 0.2 absolute.)
    // The FP32 forms issue faster than the FP64 ones.
    const float t32 = __builtin_amdgcn_rsqf((float)zz);
    return root_candidate_ok(zz, root4_refine(zz, (double)t32, (double)__builtin_amdgcn_rsqf(t32)), tol);
#else
    (void)tol;
    double r = __builtin_sqrt(__builtin_sqrt(zz));
    double f = r - __builtin_rint(r);
    return (f < 0 ? -f : f) <= ROOT_EPS;
#endif
}

HSS_HD inline bool root_near_integer(double zz) { return root_near_integer(zz, zz * 0x1p-47); }

// Tolerances of the root test for a pair: first-sign terms (zz <= l^4 + m^4)
// and counted second-sign terms (zz <= N^4 up to the estimate's error).
HSS_HD inline double root_tol1(double lm) { return lm * 0x1p-47; }
HSS_HD inline double root_tol2(double n4_d) { return n4_d * 0x1p-46; }

// ----------------------------------------------------------- range windows
//
// A leaf z = 2uv has |z| = R = 2 l^2 m^2, so its range terms depend only on
// psi = arg z modulo pi/2 (the four terms are invariant under z -> iz and
// z -> conj z).  With S = [s0, s1] the set of |D| / R that can pass or be
// undecided (the Band widened by tau), a term can count only when
// sin psi or cos psi lies in S:
//     psi in A = [asin s0, asin s1]  or  psi in A' = pi/2 - A.
// Left sets in window mode are sorted by psi_u; since psi(uv) = psi_u + psi_v
// and psi(u conj v) = psi_u - psi_v (mod pi/2), the entries u that can give a
// counted leaf with a partner v form a few contiguous index ranges, found
// through a bucket table.  Every other leaf counts 0 exactly, so skipping it
// changes no count.
//
// Error budget of the window.  s0, s1 carry FP64 errors < 1e-15 and are
// widened by 1e-6 before rounding to FP32 (rounding error < 6e-8), so the
// rounded arguments still enclose [s0, s1] and asinf is evaluated on an
// enclosing interval; asinf (<= 4 ulp, < 5e-7 rad) and the partner angle from
// atan2f on FP32 coordinates (coordinate rounding < 1.2e-7 rad, atan2f <= 3 ulp
// < 3.6e-7 rad) add < 1e-6 rad together; the sort keys psi_u are computed in
// FP64 and stored rounded to FP32 (< 1e-7 rad, psi < pi/2).  The angle window
// is widened by WIN_EPS = 2e-6 rad.  FP32 keeps the register footprint of the
// hot kernel and the sort's LDS low; the bounds only have to enclose.

static constexpr double HALF_PI = 1.5707963267948966;
static constexpr double WIN_EPS = 2e-6;

// Angle of (x, y) != (0, 0) modulo pi/2, in [0, pi/2), FP64 (sort keys).
HSS_HD inline double psi_of(double x, double y) {
    double a, b;  // rotate by a power of i into a > 0, b >= 0
    if (x > 0 && y >= 0) {
        a = x;
        b = y;
    } else if (x <= 0 && y > 0) {
        a = y;
        b = -x;
    } else if (x < 0 && y <= 0) {
        a = -x;
        b = -y;
    } else {
        a = -y;
        b = x;
    }
    double r = atan2(b, a);
    return r < HALF_PI ? r : 0.0;  // an angle that rounds to pi/2 is 0 modulo pi/2
}

// The same angle from FP32 coordinates and atan2f (error < 5e-7 rad).
HSS_HD inline double psi_of_f(double x, double y) {
    float a, b;
    if (x > 0 && y >= 0) {
        a = (float)x;
        b = (float)y;
    } else if (x <= 0 && y > 0) {
        a = (float)y;
        b = (float)-x;
    } else if (x < 0 && y <= 0) {
        a = (float)-x;
        b = (float)-y;
    } else {
        a = (float)-y;
        b = (float)x;
    }
    double r = (double)atan2f(b, a);
    return r < HALF_PI ? r : 0.0;
}

// Bucket of an angle in [0, pi/2] (WIN_NB buckets).
template <uint32_t NB>
HSS_HD inline uint32_t psi_bucket(double t) {
    double b = t * ((double)NB / HALF_PI);
    uint32_t k = b > 0 ? (uint32_t)b : 0u;
    return k < NB ? k : NB - 1;
}

// psi window of the pair (l, m): n = 1 or 2 intervals [a_i, b_i], or n = -1
// for the whole circle.  lo, hi, tau as in make_band; R = 2 l^2 m^2.
struct Window {
    double a0, b0, a1, b1;
    int n;
};

// 1/x for a positive normal x: on the device the hardware reciprocal (2^29 ulp,
// i.e. within 2^-23 relative) refined by one Newton step (error ~2^-46 plus a
// rounding); on the host the division.
HSS_HD inline double recip(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double r = __builtin_amdgcn_rcp(x);
    return __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
#else
    return 1.0 / x;
#endif
}

HSS_HD inline void make_window(double lo, double hi, double tau, double R, Window& w) {
    // one reciprocal instead of a division per bound (relative error < 1e-13)
    const double inv = recip(R);
    const bool low = lo <= tau;
    double s0 = low ? 0.0 : (lo - tau) * inv;
    double s1 = ((low ? (hi > -lo ? hi : -lo) : hi) + tau) * inv;
    s0 -= 1e-6;
    s1 += 1e-6;
    s0 = s0 < 0.0 ? 0.0 : (s0 > 1.0 ? 1.0 : s0);
    s1 = s1 < 0.0 ? 0.0 : (s1 > 1.0 ? 1.0 : s1);
    const double a0 = (double)asinf((float)s0) - WIN_EPS, a1 = (double)asinf((float)s1) + WIN_EPS;
    const double q = 0.5 * HALF_PI;
    if (a0 <= q && a1 >= q) {  // A and A' overlap: one interval
        w.a0 = a0 < HALF_PI - a1 ? a0 : HALF_PI - a1;
        w.b0 = a1 > HALF_PI - a0 ? a1 : HALF_PI - a0;
        w.a1 = w.b1 = 0.0;
        w.n = (w.b0 - w.a0 >= HALF_PI) ? -1 : 1;
    } else {
        w.a0 = a0;
        w.b0 = a1;
        w.a1 = HALF_PI - a1;
        w.b1 = HALF_PI - a0;
        w.n = 2;
    }
}

// Fraction of the psi circle a window covers (planning: work of window mode).
HSS_HD inline double window_fraction(const Window& w) {
    if (w.n < 0) return 1.0;
    double f = (w.b0 - w.a0) + (w.n == 2 ? w.b1 - w.a1 : 0.0);
    f /= HALF_PI;
    return f > 1.0 ? 1.0 : f;
}

// Index range [s, e) of the sorted entries whose bucket lies in
// [bucket(x), bucket(y)], 0 <= x <= y <= pi/2.  bt[b] = number of entries with
// bucket < b (bt[NB] = all sorted entries).
template <uint32_t NB>
HSS_HD inline void bucket_range(const uint32_t* bt, double x, double y, uint32_t& s, uint32_t& e) {
    s = bt[psi_bucket<NB>(x)];
    e = bt[psi_bucket<NB>(y) + 1];
}

// Ranges of entries u with psi_u in the cyclic interval [a + shift, b + shift]
// (b - a < pi/2): one or two index ranges; the second is empty when unused.
template <uint32_t NB>
HSS_HD inline void cyclic_ranges(const uint32_t* bt, double a, double b, double shift, uint32_t& s0, uint32_t& e0,
                                 uint32_t& s1, uint32_t& e1) {
    // x = a + shift reduced modulo pi/2 (floor through the reciprocal: an
    // off-by-one floor next to a multiple of pi/2 lands within an ulp of 0 or
    // of pi/2, which the clamp maps to 0, the same point of the circle)
    constexpr double INV_HALF_PI = 0.63661977236758134;  // 2 / pi
    double x = a + shift;
    x -= HALF_PI * floor(x * INV_HALF_PI);
    x = x < 0.0 ? 0.0 : (x >= HALF_PI ? 0.0 : x);
    const double y = x + (b - a);
    // [x, y] when y < pi/2, else [x, pi/2] and [0, y - pi/2]: bt[bucket(0)] =
    // bt[0] = 0 and bt[bucket(pi/2) + 1] = bt[NB] = all sorted entries, so two
    // table reads cover both cases
    const bool wrap = y >= HALF_PI;
    const uint32_t t = bt[psi_bucket<NB>(wrap ? y - HALF_PI : y) + 1];
    s0 = bt[psi_bucket<NB>(x)];
    e0 = wrap ? bt[NB] : t;
    s1 = 0;
    e1 = wrap ? t : 0u;
}

// Sort four ranges by start and merge overlapping ones in place; merged-away
// and empty ranges end as [0, 0).  Returns the number of entries covered.
// The ranges are held in scalar locals (array elements picked by a select
// would become indexed register reads).
HSS_HD inline void range_cswap(uint32_t& sa, uint32_t& ea, uint32_t& sb, uint32_t& eb) {
    const bool sw = sb < sa;
    const uint32_t s_lo = sw ? sb : sa, e_lo = sw ? eb : ea, s_hi = sw ? sa : sb, e_hi = sw ? ea : eb;
    sa = s_lo;
    ea = e_lo;
    sb = s_hi;
    eb = e_hi;
}

// with sa <= sb: fold range b into range a when they touch (b is cleared)
HSS_HD inline void range_fold(uint32_t sa, uint32_t& ea, uint32_t& sb, uint32_t& eb) {
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    const bool ov = sa != NONE && sb != NONE && sb <= ea;
    ea = ov ? (eb > ea ? eb : ea) : ea;
    sb = ov ? NONE : sb;
    eb = ov ? NONE : eb;
}

HSS_HD inline uint32_t merge4(uint32_t* s, uint32_t* e) {
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    uint32_t s0 = s[0], s1 = s[1], s2 = s[2], s3 = s[3], e0 = e[0], e1 = e[1], e2 = e[2], e3 = e[3];
    // empties sort last
    const bool z0 = s0 >= e0, z1 = s1 >= e1, z2 = s2 >= e2, z3 = s3 >= e3;
    s0 = z0 ? NONE : s0;
    e0 = z0 ? NONE : e0;
    s1 = z1 ? NONE : s1;
    e1 = z1 ? NONE : e1;
    s2 = z2 ? NONE : s2;
    e2 = z2 ? NONE : e2;
    s3 = z3 ? NONE : s3;
    e3 = z3 ? NONE : e3;
    range_cswap(s0, e0, s1, e1);
    range_cswap(s2, e2, s3, e3);
    range_cswap(s0, e0, s2, e2);
    range_cswap(s1, e1, s3, e3);
    range_cswap(s1, e1, s2, e2);
    // starts are sorted: a range touching an earlier live one joins the first such
    range_fold(s0, e0, s1, e1);
    range_fold(s0, e0, s2, e2);
    range_fold(s1, e1, s2, e2);
    range_fold(s0, e0, s3, e3);
    range_fold(s1, e1, s3, e3);
    range_fold(s2, e2, s3, e3);
    const bool d0 = s0 == NONE, d1 = s1 == NONE, d2 = s2 == NONE, d3 = s3 == NONE;
    s[0] = d0 ? 0u : s0;
    e[0] = d0 ? 0u : e0;
    s[1] = d1 ? 0u : s1;
    e[1] = d1 ? 0u : e1;
    s[2] = d2 ? 0u : s2;
    e[2] = d2 ? 0u : e2;
    s[3] = d3 ? 0u : s3;
    e[3] = d3 ? 0u : e3;
    return (d0 ? 0u : e0 - s0) + (d1 ? 0u : e1 - s1) + (d2 ? 0u : e2 - s2) + (d3 ? 0u : e3 - s3);
}

// Entry ranges (up to four, merged) of a left set in window mode for the
// leaves u*v (conj = false: psi_u in W - psi_v) or u*conj(v) (W + psi_v).
// nsort = entries in sorted order (the central entry is excluded).
template <uint32_t NB>
HSS_HD inline uint32_t window_ranges(const Window& w, const uint32_t* bt, uint32_t nsort, double psi_v, bool conj,
                                     uint32_t* s, uint32_t* e) {
    // branch-free (lanes of a wave hold different pairs): both arcs are always
    // mapped, then the whole circle (n < 0) and a single arc (n == 1) select
    const double shift = conj ? psi_v : -psi_v;
    uint32_t s0, e0, s1, e1, s2, e2, s3, e3;
    cyclic_ranges<NB>(bt, w.a0, w.b0, shift, s0, e0, s1, e1);
    cyclic_ranges<NB>(bt, w.a1, w.b1, shift, s2, e2, s3, e3);
    const bool all = w.n < 0, two = w.n == 2;
    s[0] = all ? 0u : s0;
    e[0] = all ? nsort : e0;
    s[1] = all ? 0u : s1;
    e[1] = all ? 0u : e1;
    s[2] = two ? s2 : 0u;
    e[2] = two ? e2 : 0u;
    s[3] = two ? s3 : 0u;
    e[3] = two ? e3 : 0u;
    return merge4(s, e);
}

// ------------------------------------------------- area classes, orientation
//
// The class rule.  Every leaf is z = 2 F z' with F = fixed(l) fixed(m) (an
// integer: the even powers of 2 and of the primes = 3 mod 4) and z' = u' v'
// (or u' conj(v')) a product of Gaussian values pi^r conj(pi)^(4E - r) of
// split primes p.  Such a value is 1 (mod 2) in its real part and 0 (mod 4) in
// its imaginary part, and so is every product of them; hence Re z' is odd,
// |Re z'| != |Im z'|, and Im z' = 0 only for the central-by-central leaf.
//   mod 3: in Z[i]/3 = F_9 conjugation is the Frobenius, conj(pi) = pi^3, and
//     pi^4 = p; a value is p^(3E) pi^(-2r): for p = 1 (mod 3) pi is a unit
//     (+-1, +-i), for p = 2 (mod 3) pi^2 = +-i.  So a value is = +-1 (Im = 0
//     mod 3) or = +-i (Re = 0 mod 3), the latter iff p = 2 (mod 3) and r is odd.
//   mod 5: Z[i]/5 = F_5 x F_5 with conjugation swapping the factors; a value
//     maps to (a^r b^(4E-r), b^r a^(4E-r)), a b = p, whose components differ by
//     the factor (a/b)^(2(r-2E)) = p^(-2(r-2E)): equal (Im = 0 mod 5) unless
//     p = +-2 (mod 5) and r is odd, then opposite (Re = 0 mod 5).
// With value_class (hss_build.h) the class c = (c3, c5) of z' is the xor over
// its factors (conjugating v keeps it), and exactly one of Re z', Im z' is
// divisible by 3, Im iff c3 = 0 (by 5 when 5 does not divide l m, Im iff
// c5 = 0).  The area tests of the reference (src/search_v18.cpp:441-453) then
// reduce to classes and parities, for pairs whose area modulus h = (3 or 9) *
// (1 or 5):
//   orientation 1 (d = Re z, area4 = Im z, 2^3 | 2 F Im z' always):
//     counts only if (c3 = 0 or 3 | l m) and (c5 = 0 or 5 | l m);
//   orientation 2 (d = Im z, area4 = Re z = 2 F Re z', Re z' odd): counts only
//     if 4 | F, i.e. l or m is even, and (c3 = 1 or 3 | l m) and (c5 = 1 or
//     5 | l m).
// (3 | l m puts 9 | F, so 9 divides both parts; 5 never divides F.)  The
// constrained bits of a pair are AREA_C3 unless 3 | l m and AREA_C5 unless
// 5 | l m; orientation 1 needs class 0 on them, orientation 2 class 3.
static constexpr uint32_t AREA_C3 = 1, AREA_C5 = 2;

HSS_HD inline uint32_t area_constrained(bool need9, bool need5) {
    return (need9 ? 0u : AREA_C3) | (need5 ? AREA_C5 : 0u);
}

// The leaf angle theta = arg z modulo pi.  Orientation 1's range test depends
// only on |cos theta| (|Re z| = 2 l^2 m^2 |cos theta|); orientation 2 of z is
// orientation 1 of i z (theta + pi/2), whose class is c xor 3.  With
// theta(u v) = theta_u + theta_v and theta(u conj v) = theta_u - theta_v, left
// sets sorted by (area group, theta_u) give, per partner, index ranges that
// hold exactly the leaves of one orientation and one class that can count.

constexpr double PI_D = 3.141592653589793;

// theta of (x, y) != (0, 0) modulo pi, in [0, pi): psi (the angle modulo pi/2
// of psi_of) plus pi/2 in the second and fourth quadrants.  FP64: sort keys.
HSS_HD inline double theta_of(double x, double y) {
    const bool odd = (x <= 0 && y > 0) || (x >= 0 && y < 0);  // psi_of's second and fourth cases
    const double t = psi_of(x, y) + (odd ? HALF_PI : 0.0);
    return t < PI_D ? t : 0.0;
}

// The same from FP32 coordinates and atan2f (error < 5e-7 rad, as psi_of_f).
HSS_HD inline double theta_of_f(double x, double y) {
    const bool odd = (x <= 0 && y > 0) || (x >= 0 && y < 0);
    const double t = psi_of_f(x, y) + (odd ? HALF_PI : 0.0);
    return t < PI_D ? t : 0.0;
}

// Bucket of an angle in [0, pi] among nb buckets (nb a power of two).
HSS_HD inline uint32_t theta_bucket(double t, uint32_t nb) {
    const double b = t * ((double)nb * (1.0 / PI_D));
    const uint32_t k = b > 0 ? (uint32_t)b : 0u;
    return k < nb ? k : nb - 1;
}

// Group of an entry of class c in a left set whose sort key keeps the class
// bits kmask (the bits any of its pairs constrains): the kept bits, packed.
HSS_HD inline uint32_t area_group(uint32_t c, uint32_t kmask) {
    return kmask == 3u ? c : (kmask == 1u ? (c & 1u) : (kmask == 2u ? (c >> 1) & 1u : 0u));
}

// Class of the j-th entry class allowed for a target class tc on the
// constrained bits cm, the other kept bits (kmask & ~cm) taken from j.
HSS_HD inline uint32_t allowed_class(uint32_t tc, uint32_t cm, uint32_t kmask, uint32_t j) {
    const uint32_t free = kmask & ~cm;
    uint32_t c = tc & cm;
    if (free == 3u) c |= j & 3u;
    else if (free) c |= (j & 1u) ? free : 0u;
    return c;
}

// Number of entry classes allowed: 2^(free bits).
HSS_HD inline uint32_t allowed_count(uint32_t cm, uint32_t kmask) {
    const uint32_t free = kmask & ~cm;
    return free == 3u ? 4u : (free ? 2u : 1u);
}

// Orientation-1 window of the pair in theta (mod pi): |cos theta| in S = [s0,
// s1] (make_window's S), i.e. |theta - pi/2| in [asin s0, asin s1]: up to two
// arcs [a, a + la], [b, b + lb] (cyclic; two arcs have la = lb), n = 1 or 2,
// or n = -1 for all of it.
// The widening is make_window's (the error budget there, with theta_u keys of
// the same precision as psi keys and theta_v from psi_of_f plus a multiple of
// pi/2).
struct Window1 {
    double a, la, b, lb;
    int n;
};

// The arcs are kept at least WIN_GAP apart (two buckets of the coarsest group
// slice, 256 buckets over pi: hss_types.h checks WIN_NB / 4 == 256), on the
// circle and across its wrap point: a closer gap is filled (its leaves are
// evaluated and count 0).  So the index ranges of the arcs' bucket-rounded
// pieces never overlap and need no merging.
constexpr double WIN_GAP = 2.0 * PI_D / 256.0;

HSS_HD inline void make_window_o1(double lo, double hi, double tau, double R, Window1& w) {
    const double inv = recip(R);
    const bool low = lo <= tau;
    double s0 = low ? 0.0 : (lo - tau) * inv;
    double s1 = ((low ? (hi > -lo ? hi : -lo) : hi) + tau) * inv;
    s0 -= 1e-6;
    s1 += 1e-6;
    s0 = s0 < 0.0 ? 0.0 : (s0 > 1.0 ? 1.0 : s0);
    s1 = s1 < 0.0 ? 0.0 : (s1 > 1.0 ? 1.0 : s1);
    const double a0 = (double)asinf((float)s0) - WIN_EPS, a1 = (double)asinf((float)s1) + WIN_EPS;
    // gaps 2 a0 (between the arcs) and pi - 2 a1 (across pi/2 + pi/2)
    const bool inner = 2.0 * a0 < WIN_GAP, outer = PI_D - 2.0 * a1 < WIN_GAP;
    w.n = (inner && outer) ? -1 : ((inner || outer) ? 1 : 2);
    // inner: [pi/2 - a1, pi/2 + a1]; outer: [pi/2 + a0, 3 pi/2 - a0]; else both halves
    const double b0 = a0 > 0.0 ? a0 : 0.0;
    w.a = inner ? HALF_PI - a1 : (outer ? HALF_PI + b0 : HALF_PI - a1);
    w.la = inner ? 2.0 * a1 : (outer ? PI_D - 2.0 * b0 : a1 - a0);
    w.b = HALF_PI + a0;
    w.lb = a1 - a0;
}

HSS_HD inline double window_o1_fraction(const Window1& w) {
    if (w.n < 0) return 1.0;
    const double f = (w.la + (w.n == 2 ? w.lb : 0.0)) * (1.0 / PI_D);
    return f > 1.0 ? 1.0 : f;
}

// Angles modulo pi in 32-bit fixed point for the range lookups of the window
// kernel: 2^32 units per pi (one unit is 7.3e-10 rad), so reduction modulo pi
// is the wraparound of unsigned arithmetic, pi/2 is 2^31, and the bucket of an
// angle among nb = 2^b buckets is its top b bits (the bucket theta_bucket
// gives the same angle, up to the conversion below).  The conversions floor
// (FP64 products: within one unit, plus one for the partner's angle in the
// shift), and the arcs are widened outward by two units at the start and four
// in length, so a fixed-point arc encloses the FP64 one.  These units add
// < 3e-9 rad to the window's error budget, inside the 2e-6 widening.
constexpr double ANGLE_UNITS_PER_RAD = 4294967296.0 / PI_D;  // 2^32 / pi
constexpr uint32_t ANGLE_HALF_PI = 0x80000000u;

// floor(t * 2^32 / pi) for t in (-pi, 2 pi) reduced modulo pi; an angle that
// rounds to pi is 0 (the same point of the circle).
HSS_HD inline uint32_t angle_fx(double t) {
    t = t < 0.0 ? t + PI_D : (t >= PI_D ? t - PI_D : t);
    const double u = t * ANGLE_UNITS_PER_RAD;
    return u < 4294967296.0 ? (uint32_t)u : 0u;
}

// theta of (x, y) != (0, 0) modulo pi from FP32 coordinates (theta_of_f), in
// angle units.
HSS_HD inline uint32_t theta_fx_of_f(double x, double y) {
    const bool odd = (x <= 0 && y > 0) || (x >= 0 && y < 0);
    return angle_fx(psi_of_f(x, y)) + (odd ? ANGLE_HALF_PI : 0u);
}

// Window1 in angle units: arcs [a, a + la] and [b, b + la] (the two arcs of a
// two-arc window have equal lengths), n as in Window1.
struct Window1F {
    uint32_t a, la, b;
    int n;
};

HSS_HD inline void window_o1_fx(const Window1& w, Window1F& f) {
    f.a = angle_fx(w.a) - 2u;
    f.b = angle_fx(w.b) - 2u;
    // la < pi - WIN_GAP whenever it is used (n >= 1); the whole circle ignores it
    const double la = w.la * ANGLE_UNITS_PER_RAD;
    f.la = la < 4294967000.0 ? (uint32_t)la + 4u : 4294967000u;
    f.n = w.n;
}

// Element i of a bucket table through a 32-bit byte offset from its base (on
// the device a scalar base plus a per-lane offset, no 64-bit address
// arithmetic per lane).
HSS_HD inline uint32_t bucket_at(const uint32_t* bt, uint32_t i) {
    return *(const uint32_t*)((const char*)bt + (i << 2));
}

// One group of a bucket table: its slice starts at entry g0 of bt and has nbg
// = 2^(32 - sh) buckets; lo = bt[g0] and hi = bt[g0 + nbg] bound the group's
// entries.
struct BucketGroup {
    const uint32_t* bt;
    uint32_t g0, sh, lo, hi;
};

// arc_ranges in angle units: entries of one group whose theta lies in the
// cyclic arc [a + shift, a + shift + len]: the arc wraps past pi when the
// unsigned sum does.  One or two index ranges.
HSS_HD inline void arc_ranges_fx(const BucketGroup& G, uint32_t a, uint32_t len, uint32_t shift, uint32_t& s0,
                                 uint32_t& e0, uint32_t& s1, uint32_t& e1) {
    const uint32_t x = a + shift, y = x + len;
    const bool wrap = y < x;
    const uint32_t t = bucket_at(G.bt, G.g0 + (y >> G.sh) + 1u);
    s0 = bucket_at(G.bt, G.g0 + (x >> G.sh));
    e0 = wrap ? G.hi : t;
    s1 = G.lo;
    e1 = wrap ? t : G.lo;
}

// Up to four disjoint index ranges of one group for the leaves u * w, w the
// partner (rotated by i for orientation 2) of angle theta_w (angle units), or
// u * conj(w) (conj): theta_u in W - theta_w, or W + theta_w.  The arcs and
// their wrapped pieces are at least WIN_GAP (two buckets) apart, so the
// bucket-rounded ranges are disjoint.  Returns the entries covered.
HSS_HD inline uint32_t window_o1_ranges(const Window1F& w, const BucketGroup& G, uint32_t theta_w, bool conj,
                                        uint32_t* s, uint32_t* e) {
    const uint32_t shift = conj ? theta_w : 0u - theta_w;
    uint32_t s0, e0, s1, e1, s2, e2, s3, e3;
    arc_ranges_fx(G, w.a, w.la, shift, s0, e0, s1, e1);
    arc_ranges_fx(G, w.b, w.la, shift, s2, e2, s3, e3);
    const bool all = w.n < 0, two = w.n == 2;
    s[0] = all ? G.lo : s0;
    e[0] = all ? G.hi : e0;
    s[1] = all ? G.lo : s1;
    e[1] = all ? G.lo : e1;
    s[2] = two ? s2 : 0u;
    e[2] = two ? e2 : 0u;
    s[3] = two ? s3 : 0u;
    e[3] = two ? e3 : 0u;
    return (e[0] - s[0]) + (e[1] - s[1]) + (e[2] - s[2]) + (e[3] - s[3]);
}

// Orientation-1 terms of a leaf with X = Re(u w) (the class rule already
// holds): bit 0 term 1 (d = +|2X|), bit 1 term 2 (d = -|2X|, not on ONE tiles);
// und when the FP64 band leaves the leaf undecided (then no term is returned).
template <bool ONE>
HSS_HD inline uint32_t o1_decide(bool live, double X, const Band& b, bool& und) {
    // fabs (not a select): an absolute-value source modifier on the device
    const double ax = __builtin_fabs(X);
    const double adx = __builtin_fabs(ax - b.mid);
    const bool t1w = adx <= b.half_w, t1n = adx <= b.half_w - b.tau;
    const bool t2w = !ONE && ax <= b.h2_w, t2n = !ONE && ax <= b.h2_w - b.tau;
    const bool u = live & ((t1w != t1n) | (t2w != t2n));
    und = u;
    return ((live & t1n & !u) ? 1u : 0u) | ((live & t2n & !u) ? 2u : 0u);
}

// ------------------------------------------------------------ exact stage

struct ExactCtx {
    uint64_t l, m, limit;
    u128 l4, m4, n4;
    i128 lo, hi;
    bool need9, need5;
};

HSS_HD inline void make_exact_ctx(uint64_t l, uint64_t m, uint64_t limit, ExactCtx& e) {
    e.l = l;
    e.m = m;
    e.limit = limit;
    e.l4 = (u128)l * l * l * l;
    e.m4 = (u128)m * m * m * m;
    e.n4 = (u128)limit * limit * limit * limit;
    e.lo = (i128)e.l4 + (i128)e.m4 - (i128)e.n4;
    e.hi = (i128)e.l4;
    e.need9 = (l % 3 == 0) || (m % 3 == 0);
    e.need5 = (l % 5 != 0) && (m % 5 != 0);
}

HSS_HD inline uint64_t gcd_u64(uint64_t a, uint64_t b) {
    while (b) {
        uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

// 256-bit limb product a (na limbs) * b (nb limbs) -> out (8 limbs, truncated)
HSS_HD inline void mul_limbs(const uint32_t* a, int na, const uint32_t* b, int nb, uint32_t* out) {
    uint64_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < na; ++i) {
        uint64_t carry = 0;
        for (int j = 0; j < nb && i + j < 8; ++j) {
            uint64_t v = (uint64_t)a[i] * b[j] + acc[i + j] + carry;
            acc[i + j] = v & 0xFFFFFFFFu;
            carry = v >> 32;
        }
        for (int j = i + nb; j < 8 && carry; ++j) {
            uint64_t v = acc[j] + carry;
            acc[j] = v & 0xFFFFFFFFu;
            carry = v >> 32;
        }
    }
    for (int i = 0; i < 8; ++i) out[i] = (uint32_t)acc[i];
}

// Heron product (x+y+z)(-x+y+z)(x-y+z)(x+y-z) == area4^2 for sides l^2, m^2,
// n^2, all factors positive and 4 | area4 (src/search_v18.cpp:485-501)
HSS_HD inline bool heron_matches(uint64_t l, uint64_t m, uint64_t n, u128 area4) {
    i128 x = (i128)l * l, y = (i128)m * m, z = (i128)n * n;
    i128 f[4] = {x + y + z, -x + y + z, x - y + z, x + y - z};
    for (int i = 0; i < 4; ++i)
        if (f[i] <= 0) return false;
    if (area4 & 3) return false;
    uint32_t p[8], b[4], t[8];
    for (int i = 0; i < 8; ++i) p[i] = i < 4 ? (uint32_t)((u128)f[0] >> (32 * i)) : 0u;
    for (int k = 1; k < 4; ++k) {
        for (int i = 0; i < 4; ++i) b[i] = (uint32_t)((u128)f[k] >> (32 * i));
        mul_limbs(p, 8, b, 4, t);
        for (int i = 0; i < 8; ++i) p[i] = t[i];
    }
    uint32_t a[4];
    for (int i = 0; i < 4; ++i) a[i] = (uint32_t)(area4 >> (32 * i));
    mul_limbs(a, 4, a, 4, t);
    for (int i = 0; i < 8; ++i)
        if (p[i] != t[i]) return false;
    return true;
}

// Fourth-power test of zz = l^4 + m^4 - d for one candidate: residues modulo
// 16, 625, 9, 7, 13, 17 and the exact integer fourth root
// (src/search_v18.cpp:457-475).  Returns n when zz = n^4, else 0.
HSS_HD inline uint64_t fourth_root_of(const ExactCtx& e, i128 d, const QR4& q) {
    i128 zs = (i128)e.l4 + (i128)e.m4 - d;
    if (zs < 0) return 0;
    u128 z = (u128)zs;
    if (z < e.m4 || z > e.n4) return 0;
    if ((uint32_t)(z & 15u) > 1u) return 0;
    uint32_t r625 = umod128(z, 625);
    if (e.need5 ? r625 != 0 : !bit_of(q.m625, r625)) return 0;
    if (!((q.m9 >> umod128(z, 9)) & 1u) || !((q.m7 >> umod128(z, 7)) & 1u) ||
        !((q.m13 >> umod128(z, 13)) & 1u) || !((q.m17 >> umod128(z, 17)) & 1u))
        return 0;
    double zd = (double)(uint64_t)(z >> 64) * 18446744073709551616.0 + (double)(uint64_t)z;
    uint64_t r = (uint64_t)__builtin_sqrt(__builtin_sqrt(zd));
    while (r > 0 && (u128)r * r * r * r > z) --r;
    while ((u128)(r + 1) * (r + 1) * (r + 1) * (r + 1) <= z) ++r;
    return (u128)r * r * r * r == z ? r : 0;
}

// Acceptance after an exact fourth power (src/search_v18.cpp:476-477 and the
// recheck of :485-501): gcd(l, m, n) = 1, 5 | l*m*n, m <= n <= LIMIT, Heron.
HSS_HD inline bool accept_hit(uint64_t l, uint64_t m, uint64_t n, uint64_t limit, u128 area4) {
    if (gcd_u64(gcd_u64(l, m), n) != 1) return false;
    if (l % 5 && m % 5 && n % 5) return false;
    if (n < m || n > limit) return false;
    return heron_matches(l, m, n, area4);
}

struct HitOut {
    uint64_t l, m, n;
    uint64_t area4_lo, area4_hi;  // |4A|
    uint32_t pos;                 // position index in the batch that found it
    uint32_t pad;
};

// Full reference filter chain of one leaf z (exact).  Returns the candidates
// of the orientations in omask (bit 0: d = Re z, bit 1: d = Im z); every exact
// fourth power of those orientations is passed to sink(n, |area4|).
template <typename Sink>
HSS_HD inline uint32_t exact_leaf(i128 zre, i128 zim, const ExactCtx& e, const QR4& q, Sink&& sink,
                                  uint32_t omask = 3u) {
    uint32_t cand = 0;
    for (int o = 0; o < 2; ++o) {
        if (o == 1 && (zre == zim || zre == -zim)) break;
        if (!((omask >> o) & 1u)) continue;
        i128 d = o == 0 ? zre : zim;
        i128 a4 = o == 0 ? zim : zre;
        if (a4 == 0 || ((uint64_t)a4 & 7u)) continue;
        bool pos = d >= e.lo && d <= e.hi;
        bool neg = -d >= e.lo && -d <= e.hi;
        if (!pos && !neg) continue;
        u128 aa = (u128)(a4 < 0 ? -a4 : a4);
        if (umod128(aa, 3)) continue;
        if (e.need9 && umod128(aa, 9)) continue;
        if (e.need5 && umod128(aa, 5)) continue;
        for (int s = 0; s < 2; ++s) {
            if (s == 0 ? !pos : (!neg || d == 0)) continue;
            ++cand;
            uint64_t r = fourth_root_of(e, s == 0 ? d : -d, q);
            if (r) sink(r, aa);
        }
    }
    return cand;
}

// exact Gaussian product z = 2 * u * v (conj: 2 * u * conj(v))
HSS_HD inline void exact_product(i128 ure, i128 uim, int64_t vre, int64_t vim, bool conj, i128& zre, i128& zim) {
    i128 c = vre, d = conj ? -(i128)vim : (i128)vim;
    zre = 2 * (ure * c - uim * d);
    zim = 2 * (ure * d + uim * c);
}

}  // namespace hss
