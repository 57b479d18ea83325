// hss_build.h -- table-building arithmetic shared by the kernels and the
// host-side checker: Gaussian splits of primes p = 1 (mod 4), the values
// pi^r conj(pi)^(4E-r), and the mixed-radix decode of a half-set rank.
#pragma once

#include "hss_math.h"

namespace hss {

struct GInt {
    i128 re, im;
};

HSS_HD inline GInt gmul(GInt a, GInt b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

HSS_HD inline GInt gpow(GInt b, uint32_t e) {
    GInt r = {1, 0};
    while (e) {
        if (e & 1) r = gmul(r, b);
        b = gmul(b, b);
        e >>= 1;
    }
    return r;
}

HSS_HD inline uint64_t powmod_u64(uint64_t b, uint64_t e, uint64_t p) {
    uint64_t r = 1, x = b % p;  // p < 2^26: products stay below 2^52
    while (e) {
        if (e & 1) r = r * x % p;
        x = x * x % p;
        e >>= 1;
    }
    return r;
}

// a + bi with a^2 + b^2 = p, a odd, b even, both positive (Cornacchia: a square
// root x of -1 modulo p from a non-residue, then Euclid on (p, x) until the
// remainder drops below sqrt(p)).  Any choice of pi works: a unit multiple or
// the conjugate only negates or conjugates the enumerated values.
HSS_HD inline void prime_split(uint64_t p, int64_t& a, int64_t& b) {
    uint64_t x = 0;
    for (uint64_t c = 2; c < p; ++c) {
        uint64_t r = powmod_u64(c, (p - 1) / 4, p);
        if (r * r % p == p - 1) {
            x = r;
            break;
        }
    }
    uint64_t r0 = p, r1 = x;
    while (r1 * r1 > p) {
        uint64_t t = r0 % r1;
        r0 = r1;
        r1 = t;
    }
    uint64_t rem = p - r1 * r1;
    uint64_t s = (uint64_t)__builtin_sqrt((double)rem);
    while (s * s > rem) --s;
    while ((s + 1) * (s + 1) <= rem) ++s;
    a = (int64_t)((r1 & 1) ? r1 : s);
    b = (int64_t)((r1 & 1) ? s : r1);
}

// pi^r conj(pi)^(4E - r) for pi = a + bi
HSS_HD inline GInt gauss_value(int64_t a, int64_t b, uint32_t E, uint32_t r) {
    GInt pi = {(i128)a, (i128)b}, pc = {(i128)a, -(i128)b};
    return gmul(gpow(pi, r), gpow(pc, 4 * E - r));
}

// Area class of a Gaussian value pi^r conj(pi)^(4E - r) of the split prime p
// (the class rule of hss_math.h): bit 0 is set when p = 2 (mod 3) and r is
// odd, bit 1 when p = 2 or 3 (mod 5) and r is odd.  The class of a product of
// values is the xor of its factors' classes; the central value (r = 2E) has
// class 0.
HSS_HD inline uint32_t value_class(uint64_t p, uint32_t r) {
    const uint32_t odd = r & 1u;
    const uint32_t p3 = (uint32_t)(p % 3), p5 = (uint32_t)(p % 5);
    return (p3 == 2 ? odd : 0u) | ((p5 == 2 || p5 == 3) ? odd << 1 : 0u);
}

// class of the product of values with digits dig over primes p[0..k)
HSS_HD inline uint32_t digits_class(uint32_t k, const uint32_t* p, const uint32_t* dig) {
    uint32_t c = 0;
    for (uint32_t j = 0; j < k; ++j) c ^= value_class(p[j], dig[j]);
    return c;
}

// offset of exponent E inside a prime's value block: sum_{e<E} (4e+1)
HSS_HD inline uint32_t exp_offset(uint32_t E) { return (E - 1) * (2 * E + 1); }

// digits of half-set rank t over radices 4E_j + 1 (first prime most significant)
HSS_HD inline void decode_rank(uint32_t t, uint32_t k, const uint32_t* E, uint32_t* dig) {
    for (int j = (int)k - 1; j >= 0; --j) {
        uint32_t R = 4 * E[j] + 1;
        dig[j] = t % R;
        t /= R;
    }
}

// nearest-ish FP64 value of v (relative error below 2^-52): the magnitude is
// converted as two non-negative halves, so nothing cancels
HSS_HD inline double i128_to_double(i128 v) {
    u128 a = v < 0 ? (u128)(-v) : (u128)v;
    double d = (double)(uint64_t)(a >> 64) * 18446744073709551616.0 + (double)(uint64_t)a;
    return v < 0 ? -d : d;
}

}  // namespace hss
