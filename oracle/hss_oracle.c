/*
 * hss_oracle.c -- CPU oracle for the Gaussian-integer enumeration core.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker.
 * The product never links, loads or calls it.
 *
 * It restates the CPU reference (src/search_v18.cpp) from its mathematics:
 *
 *   - tables per integer i: split primes p = 1 (mod 4) with exponents, the fixed
 *     part prod p^(2a) over p = 2 and p = 3 (mod 4), the product of distinct inert
 *     primes (search_v18.cpp:159-197);
 *   - for a split prime p = pi * conj(pi) and merged exponent E the 4E+1 Gaussian
 *     integers pi^r conj(pi)^(4E-r) of norm p^(4E) (search_v18.cpp:285-309);
 *   - for each pair (l, m) with m in [l, LIMIT] the skip rules and context of
 *     search_v18.cpp:329-371, then every mixed-radix digit tuple that is
 *     lexicographically <= the central tuple (search_v18.cpp:373-434), here
 *     walked by an odometer with cached prefix products instead of recursion;
 *   - the filter chain test_all / test_pair / test_d with the candidate counter
 *     at entry of test_d (search_v18.cpp:436-479);
 *   - the exact 256-bit Heron recheck before a hit is recorded
 *     (search_v18.cpp:481-507), here with 32-bit limbs.
 *
 * The Gaussian split of p is found by Cornacchia's method (a different but
 * equivalent choice of pi: a unit multiple or conjugate of pi only negates or
 * conjugates the enumerated values, which every filter treats identically).
 */
#define _GNU_SOURCE
#include "hss_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef __int128 i128;

typedef struct {
    i128 re, im;
} gint;

typedef struct {
    uint32_t p[8];
    uint8_t e[8];
    uint8_t k;
    uint64_t fixed;
    uint32_t inert;
} number_t;

struct hsso_ctx {
    uint64_t nmax;
    uint32_t* spf;
    number_t* num;
    uint32_t* vbase; /* per prime p = 1 (mod 4): offset of its value block in vals */
    gint* vals;
    size_t nvals;
};

/* ------------------------------------------------------------------ helpers */

static gint gmul(gint a, gint b) {
    gint r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}

static uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) {
        uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

static uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
    u128 r = 1, x = b % m;
    while (e) {
        if (e & 1) r = r * x % m;
        x = x * x % m;
        e >>= 1;
    }
    return (uint64_t)r;
}

static uint64_t isqrt64(uint64_t v) {
    uint64_t r = (uint64_t)sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

int hsso_prime_split(uint64_t p, uint64_t* a, uint64_t* b) {
    if (p % 4 != 1) return 0;
    /* x^2 = -1 (mod p) from a quadratic non-residue c: x = c^((p-1)/4) */
    uint64_t x = 0;
    for (uint64_t c = 2; c < p; ++c) {
        uint64_t t = powmod(c, (p - 1) / 4, p);
        if ((u128)t * t % p == p - 1) {
            x = t;
            break;
        }
    }
    if (!x) return 0;
    /* Euclid on (p, x) until the remainder drops below sqrt(p) */
    uint64_t r0 = p, r1 = x;
    while (r1 * r1 > p) {
        uint64_t t = r0 % r1;
        r0 = r1;
        r1 = t;
    }
    uint64_t s = isqrt64(p - r1 * r1);
    if (r1 * r1 + s * s != p) return 0;
    if (r1 & 1) {
        *a = r1;
        *b = s;
    } else {
        *a = s;
        *b = r1;
    }
    return 1;
}

/* offset of exponent E inside one prime's block: sum_{e<E} (4e+1) */
static inline uint32_t exp_offset(uint32_t E) { return (E - 1) * (2 * E + 1); }

static uint32_t max_merged_exponent(uint64_t p, uint64_t n) {
    uint32_t a = 0;
    for (u128 q = p; q <= n; q *= p) ++a;
    return 2 * a;
}

/* ------------------------------------------------------------------ tables */

hsso_ctx* hsso_create(uint64_t nmax) {
    /* coordinates reach 2 * nmax^4, which must stay below 2^127 */
    if (nmax < 1 || nmax > (1ull << 27)) return NULL;
    hsso_ctx* c = calloc(1, sizeof(*c));
    if (!c) return NULL;
    c->nmax = nmax;
    c->spf = calloc(nmax + 1, sizeof(uint32_t));
    c->num = calloc(nmax + 1, sizeof(number_t));
    c->vbase = malloc((nmax + 1) * sizeof(uint32_t));
    if (!c->spf || !c->num || !c->vbase) {
        hsso_destroy(c);
        return NULL;
    }
    /* smallest prime factor sieve */
    for (uint64_t i = 2; i <= nmax; ++i) {
        if (c->spf[i]) continue;
        c->spf[i] = (uint32_t)i;
        if (i > nmax / i) continue;
        for (uint64_t j = i * i; j <= nmax; j += i)
            if (!c->spf[j]) c->spf[j] = (uint32_t)i;
    }
    c->spf[1] = 1;
    /* per-integer factor facts */
    for (uint64_t i = 1; i <= nmax; ++i) {
        number_t* nu = &c->num[i];
        nu->fixed = 1;
        nu->inert = 1;
        uint64_t x = i;
        while (x > 1) {
            uint32_t p = c->spf[x];
            uint32_t a = 0;
            while (x % p == 0) {
                x /= p;
                ++a;
            }
            if (p % 4 == 1) {
                nu->p[nu->k] = p;
                nu->e[nu->k] = (uint8_t)a;
                nu->k++;
            } else {
                for (uint32_t t = 0; t < 2 * a; ++t) nu->fixed *= p;
                if (p % 4 == 3) nu->inert *= p;
            }
        }
    }
    /* Gaussian value blocks of the split primes */
    size_t total = 0;
    for (uint64_t p = 5; p <= nmax; p += 4) {
        c->vbase[p] = UINT32_MAX;
        if (c->spf[p] != p) continue;
        uint32_t E = max_merged_exponent(p, nmax);
        total += exp_offset(E + 1);
    }
    c->vals = malloc((total ? total : 1) * sizeof(gint));
    if (!c->vals) {
        hsso_destroy(c);
        return NULL;
    }
    size_t at = 0;
    gint pw[4 * 64 + 1], pc[4 * 64 + 1];
    for (uint64_t p = 5; p <= nmax; p += 4) {
        if (c->spf[p] != p) continue;
        uint64_t a, b;
        if (!hsso_prime_split(p, &a, &b)) {
            hsso_destroy(c);
            return NULL;
        }
        uint32_t Emax = max_merged_exponent(p, nmax);
        c->vbase[p] = (uint32_t)at;
        gint pi = {(i128)a, (i128)b}, pib = {(i128)a, -(i128)b};
        pw[0] = pc[0] = (gint){1, 0};
        for (uint32_t r = 1; r <= 4 * Emax; ++r) {
            pw[r] = gmul(pw[r - 1], pi);
            pc[r] = gmul(pc[r - 1], pib);
        }
        for (uint32_t E = 1; E <= Emax; ++E)
            for (uint32_t r = 0; r <= 4 * E; ++r) c->vals[at++] = gmul(pw[r], pc[4 * E - r]);
    }
    c->nvals = at;
    return c;
}

void hsso_destroy(hsso_ctx* c) {
    if (!c) return;
    free(c->spf);
    free(c->num);
    free(c->vbase);
    free(c->vals);
    free(c);
}

uint32_t hsso_weight(const hsso_ctx* c, uint64_t i) {
    const number_t* nu = &c->num[i];
    uint32_t w = 1;
    for (int t = 0; t < nu->k; ++t) w *= 4u * nu->e[t] + 1u;
    return w;
}

uint64_t hsso_fixed(const hsso_ctx* c, uint64_t i) { return c->num[i].fixed; }
int hsso_has_inert(const hsso_ctx* c, uint64_t i) { return c->num[i].inert > 1; }

/* ------------------------------------------------------------ 256-bit Heron */

typedef struct {
    uint32_t w[8];
} u256;

static u256 u256_of(u128 v) {
    u256 r;
    memset(&r, 0, sizeof r);
    for (int i = 0; i < 4; ++i) r.w[i] = (uint32_t)(v >> (32 * i));
    return r;
}

static u256 u256_mul(u256 a, u256 b) {
    uint64_t acc[8] = {0};
    u256 r;
    for (int i = 0; i < 8; ++i) {
        uint64_t carry = 0;
        for (int j = 0; i + j < 8; ++j) {
            uint64_t t = (uint64_t)a.w[i] * b.w[j] + acc[i + j] + carry;
            acc[i + j] = t & 0xFFFFFFFFull;
            carry = t >> 32;
        }
    }
    for (int i = 0; i < 8; ++i) r.w[i] = (uint32_t)acc[i];
    return r;
}

static int u256_cmp(u256 a, u256 b) {
    for (int i = 7; i >= 0; --i)
        if (a.w[i] != b.w[i]) return a.w[i] < b.w[i] ? -1 : 1;
    return 0;
}

/* Heron product (x+y+z)(-x+y+z)(x-y+z)(x+y-z) for sides x=l^2, y=m^2, z=n^2;
 * returns 0 when some factor is not positive. */
static int heron_product(uint64_t l, uint64_t m, uint64_t n, u256* out) {
    i128 x = (i128)l * l, y = (i128)m * m, z = (i128)n * n;
    i128 f[4] = {x + y + z, -x + y + z, x - y + z, x + y - z};
    for (int i = 0; i < 4; ++i)
        if (f[i] <= 0) return 0;
    u256 p = u256_of((u128)f[0]);
    for (int i = 1; i < 4; ++i) p = u256_mul(p, u256_of((u128)f[i]));
    *out = p;
    return 1;
}

int hsso_heron(uint64_t l, uint64_t m, uint64_t n, uint64_t* lo, uint64_t* hi) {
    u256 p;
    if (!heron_product(l, m, n, &p)) return 0;
    /* bitwise integer square root over u128 (the root is below 2^100) */
    u128 s = 0;
    for (int bit = 110; bit >= 0; --bit) {
        u128 t = s | ((u128)1 << bit);
        if (u256_cmp(u256_mul(u256_of(t), u256_of(t)), p) <= 0) s = t;
    }
    if (u256_cmp(u256_mul(u256_of(s), u256_of(s)), p) != 0) return 0;
    if (s & 3) return 0;
    *lo = (uint64_t)s;
    *hi = (uint64_t)(s >> 64);
    return 1;
}

/* --------------------------------------------------------- modular filters */

/* residue of a value below 2^128 modulo a small odd modulus */
static inline uint32_t mod_small(u128 v, uint32_t q) {
    uint64_t hi = (uint64_t)(v >> 64), lo = (uint64_t)v;
    uint64_t two64 = (uint64_t)(((u128)1 << 64) % q);
    return (uint32_t)(((hi % q) * two64 + lo % q) % q);
}

/* fourth-power residue sets, built once */
static uint8_t qr625[625], qr9[9], qr7[7], qr13[13], qr17[17];
static pthread_once_t qr_once = PTHREAD_ONCE_INIT;

static void fill_qr(uint8_t* t, uint32_t q) {
    memset(t, 0, q);
    for (uint64_t x = 0; x < q; ++x) t[(x * x % q) * (x * x % q) % q] = 1;
}

static void build_qr(void) {
    fill_qr(qr625, 625);
    fill_qr(qr9, 9);
    fill_qr(qr7, 7);
    fill_qr(qr13, 13);
    fill_qr(qr17, 17);
}

/* --------------------------------------------------------------- one pair */

typedef struct {
    uint64_t l, m, limit;
    u128 l4, m4, n4;
    i128 lower, upper;
    int need9, need5; /* 3 | l*m, and 5 divides neither */
    uint64_t cand;
    hsso_hit* hits;
    uint32_t cap, nhits;
} pairctx;

static void record_hit(pairctx* x, uint64_t root, i128 area4) {
    u256 p;
    u128 a = (u128)(area4 < 0 ? -area4 : area4);
    if (!heron_product(x->l, x->m, root, &p) || (a & 3)) return;
    if (u256_cmp(p, u256_mul(u256_of(a), u256_of(a))) != 0) return;
    for (uint32_t i = 0; i < x->nhits && i < x->cap; ++i)
        if (x->hits[i].l == x->l && x->hits[i].m == x->m && x->hits[i].n == root) return;
    if (x->nhits < x->cap) {
        hsso_hit* h = &x->hits[x->nhits];
        h->l = x->l;
        h->m = x->m;
        h->n = root;
        h->area4_lo = (uint64_t)a;
        h->area4_hi = (uint64_t)(a >> 64);
    }
    x->nhits++;
}

static void fourth_power_test(pairctx* x, i128 d, i128 area4) {
    x->cand++;
    i128 zs = (i128)x->l4 + (i128)x->m4 - d;
    if (zs < 0) return;
    u128 z = (u128)zs;
    if (z < x->m4 || z > x->n4) return;
    if ((uint32_t)(z & 15) > 1) return;
    uint32_t r625 = mod_small(z, 625);
    if (x->need5 ? r625 != 0 : !qr625[r625]) return;
    if (!qr9[mod_small(z, 9)] || !qr7[mod_small(z, 7)] || !qr13[mod_small(z, 13)] ||
        !qr17[mod_small(z, 17)])
        return;
    uint64_t r = (uint64_t)sqrtl(sqrtl((long double)z));
    while (r > 0 && (u128)r * r * r * r > z) --r;
    while ((u128)(r + 1) * (r + 1) * (r + 1) * (r + 1) <= z) ++r;
    if ((u128)r * r * r * r != z) return;
    if (gcd64(gcd64(x->l, x->m), r) != 1) return;
    if (x->l % 5 && x->m % 5 && r % 5) return;
    if (r < x->m || r > x->limit) return;
    record_hit(x, r, area4);
}

static void orientation(pairctx* x, i128 d, i128 area4) {
    if (area4 == 0 || ((uint64_t)area4 & 7)) return;
    int pos = d >= x->lower && d <= x->upper;
    int neg = -d >= x->lower && -d <= x->upper;
    if (!pos && !neg) return;
    u128 a = (u128)(area4 < 0 ? -area4 : area4);
    if (mod_small(a, 3)) return;
    if (x->need9 && mod_small(a, 9)) return;
    if (x->need5 && mod_small(a, 5)) return;
    if (pos) fourth_power_test(x, d, area4);
    if (neg && d != 0) fourth_power_test(x, -d, area4);
}

static void leaf(pairctx* x, gint z) {
    orientation(x, z.re, z.im);
    if (z.re != z.im && z.re != -z.im) orientation(x, z.im, z.re);
}

typedef struct {
    const gint* v; /* 4E+1 values */
    uint32_t radix;
} term_t;

/* Merge the split primes of l and m into terms; returns the term count. */
static int merge_terms(const hsso_ctx* c, const number_t* a, const number_t* b, term_t* t) {
    int i = 0, j = 0, k = 0;
    while (i < a->k || j < b->k) {
        uint32_t p, E;
        if (j == b->k || (i < a->k && a->p[i] < b->p[j])) {
            p = a->p[i];
            E = a->e[i++];
        } else if (i == a->k || b->p[j] < a->p[i]) {
            p = b->p[j];
            E = b->e[j++];
        } else {
            p = a->p[i];
            E = (uint32_t)a->e[i++] + b->e[j++];
        }
        t[k].v = c->vals + c->vbase[p] + exp_offset(E);
        t[k].radix = 4 * E + 1;
        ++k;
    }
    return k;
}

/* pair skip rules of the reference: both even; a shared prime = 3 (mod 4) */
static int pair_skipped(const hsso_ctx* c, uint64_t l, uint64_t m) {
    if (!(l & 1) && !(m & 1)) return 1;
    uint32_t in = c->num[l].inert;
    return in > 1 && gcd64(in, m) > 1;
}

static void walk_pair(const hsso_ctx* c, pairctx* x) {
    const number_t* a = &c->num[x->l];
    const number_t* b = &c->num[x->m];
    term_t t[16];
    int k = merge_terms(c, a, b, t);
    if (k == 0) return;
    uint64_t W = 1;
    for (int i = 0; i < k; ++i) W *= t[i].radix;
    uint64_t leaves = (W + 1) / 2; /* tuples lexicographically <= the central one */
    uint32_t dig[16] = {0};
    gint pre[17];
    pre[0] = (gint){(i128)2 * a->fixed * b->fixed, 0};
    for (int i = 0; i < k; ++i) pre[i + 1] = gmul(pre[i], t[i].v[0]);
    for (uint64_t s = 0;; ++s) {
        leaf(x, pre[k]);
        if (s + 1 == leaves) break;
        int i = k - 1;
        while (++dig[i] == t[i].radix) dig[i--] = 0;
        for (int j = i; j < k; ++j) pre[j + 1] = gmul(pre[j], t[j].v[dig[j]]);
    }
}

uint64_t hsso_position(const hsso_ctx* c, uint64_t limit, uint64_t l, hsso_hit* hits,
                       uint32_t cap, uint32_t* nhits) {
    pthread_once(&qr_once, build_qr);
    pairctx x;
    memset(&x, 0, sizeof x);
    x.l = l;
    x.limit = limit;
    x.hits = hits;
    x.cap = cap;
    if (l < 1 || limit > c->nmax || l > limit) {
        if (nhits) *nhits = 0;
        return 0;
    }
    x.l4 = (u128)l * l * l * l;
    x.n4 = (u128)limit * limit * limit * limit;
    x.upper = (i128)x.l4;
    for (uint64_t m = l; m <= limit; ++m) {
        if (pair_skipped(c, l, m)) continue;
        x.m = m;
        x.m4 = (u128)m * m * m * m;
        x.lower = (i128)x.l4 + (i128)x.m4 - (i128)x.n4;
        x.need9 = (l % 3 == 0 || m % 3 == 0);
        x.need5 = (l % 5 != 0 && m % 5 != 0);
        walk_pair(c, &x);
    }
    if (nhits) *nhits = x.nhits;
    return x.cand;
}

void hsso_work(const hsso_ctx* c, uint64_t limit, uint64_t l, uint64_t* evaluated,
               uint64_t* leaves) {
    uint64_t ev = 0, lv = 0;
    const number_t* a = &c->num[l];
    for (uint64_t m = l; m <= limit && l >= 1 && limit <= c->nmax; ++m) {
        if (pair_skipped(c, l, m)) continue;
        term_t t[16];
        int k = merge_terms(c, a, &c->num[m], t);
        if (k == 0) continue;
        uint64_t W = 1;
        for (int i = 0; i < k; ++i) W *= t[i].radix;
        ++ev;
        lv += (W + 1) / 2;
    }
    if (evaluated) *evaluated = ev;
    if (leaves) *leaves = lv;
}

/* ------------------------------------------------------- threaded interval */

typedef struct {
    const hsso_ctx* c;
    uint64_t limit, first, last;
    _Atomic uint64_t next;
    uint64_t* per_l;
    _Atomic uint64_t total;
    pthread_mutex_t mu;
    hsso_hit* hits;
    uint32_t cap, nhits;
} job_t;

static void add_hits(job_t* j, const hsso_hit* h, uint32_t n) {
    pthread_mutex_lock(&j->mu);
    for (uint32_t i = 0; i < n; ++i) {
        int dup = 0;
        for (uint32_t k = 0; k < j->nhits && k < j->cap; ++k)
            if (j->hits[k].l == h[i].l && j->hits[k].m == h[i].m && j->hits[k].n == h[i].n) dup = 1;
        if (dup) continue;
        if (j->nhits < j->cap) j->hits[j->nhits] = h[i];
        j->nhits++;
    }
    pthread_mutex_unlock(&j->mu);
}

static void* worker(void* arg) {
    job_t* j = arg;
    hsso_hit local[64];
    for (;;) {
        uint64_t l = atomic_fetch_add(&j->next, 1);
        if (l > j->last) break;
        uint32_t n = 0;
        uint64_t cnd = hsso_position(j->c, j->limit, l, local, 64, &n);
        if (j->per_l) j->per_l[l - j->first] = cnd;
        atomic_fetch_add(&j->total, cnd);
        if (n) add_hits(j, local, n < 64 ? n : 64);
    }
    return NULL;
}

static int hit_less(const void* pa, const void* pb) {
    const hsso_hit *a = pa, *b = pb;
    if (a->l != b->l) return a->l < b->l ? -1 : 1;
    if (a->m != b->m) return a->m < b->m ? -1 : 1;
    if (a->n != b->n) return a->n < b->n ? -1 : 1;
    return 0;
}

uint64_t hsso_interval(const hsso_ctx* c, uint64_t limit, uint64_t first, uint64_t last,
                       uint32_t threads, uint64_t* per_l, hsso_hit* hits, uint32_t cap,
                       uint32_t* nhits, uint64_t* pairs) {
    job_t j;
    memset(&j, 0, sizeof j);
    j.c = c;
    j.limit = limit;
    j.first = first;
    j.last = last;
    atomic_init(&j.next, first);
    atomic_init(&j.total, 0);
    j.per_l = per_l;
    j.hits = hits;
    j.cap = cap;
    pthread_mutex_init(&j.mu, NULL);
    if (threads < 1) threads = 1;
    pthread_t* tid = calloc(threads, sizeof(pthread_t));
    for (uint32_t t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, worker, &j);
    for (uint32_t t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
    free(tid);
    pthread_mutex_destroy(&j.mu);
    uint32_t kept = j.nhits < cap ? j.nhits : cap;
    if (kept > 1) qsort(hits, kept, sizeof(hsso_hit), hit_less);
    if (nhits) *nhits = j.nhits;
    if (pairs) {
        uint64_t s = 0;
        for (uint64_t l = first; l <= last; ++l) s += limit - l + 1;
        *pairs = s;
    }
    return atomic_load(&j.total);
}
