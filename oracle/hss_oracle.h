/*
 * hss_oracle.h -- CPU oracle for the Gaussian-integer enumeration core.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker;
 * the product (libhss.so and the gpu_search / search executables) never links,
 * loads or calls it.
 */
#ifndef HSS_ORACLE_H
#define HSS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hsso_ctx hsso_ctx;

/* One recorded solution: roots l <= m <= n and |4A| as a 128-bit integer. */
typedef struct {
    uint64_t l, m, n;
    uint64_t area4_lo, area4_hi;
} hsso_hit;

/* Build the number-theory tables for every integer <= nmax. Returns NULL on OOM. */
hsso_ctx* hsso_create(uint64_t nmax);
void hsso_destroy(hsso_ctx* ctx);

/*
 * Search one root position l against every m in [l, limit] (limit <= nmax).
 * Returns the candidate count (entries of the fourth-power test); writes at most
 * `cap` hits and stores the number found (may exceed cap) in *nhits.
 */
uint64_t hsso_position(const hsso_ctx* ctx, uint64_t limit, uint64_t l,
                       hsso_hit* hits, uint32_t cap, uint32_t* nhits);

/*
 * Search root positions [first, last] on `threads` threads.  per_l (nullable,
 * length last-first+1) receives each position's candidate count.  Hits are
 * de-duplicated on (l, m, n) and sorted.  Returns total candidates; *pairs gets
 * the reference's pair counter sum(limit - l + 1).
 */
uint64_t hsso_interval(const hsso_ctx* ctx, uint64_t limit, uint64_t first, uint64_t last,
                       uint32_t threads, uint64_t* per_l, hsso_hit* hits, uint32_t cap,
                       uint32_t* nhits, uint64_t* pairs);

/*
 * Closed-form work of one position: number of (l, m) pairs that reach the
 * enumeration, and the number of enumerated Gaussian representations
 * sum((W+1)/2) over them.
 */
void hsso_work(const hsso_ctx* ctx, uint64_t limit, uint64_t l,
               uint64_t* evaluated_pairs, uint64_t* leaves);

/* Per-integer facts used by the tests: Gaussian weight, fixed part, inert flag. */
uint32_t hsso_weight(const hsso_ctx* ctx, uint64_t i);
uint64_t hsso_fixed(const hsso_ctx* ctx, uint64_t i);
int hsso_has_inert(const hsso_ctx* ctx, uint64_t i);

/* Gaussian split of a prime p = 1 (mod 4): a^2 + b^2 = p, a odd, b even, both > 0. */
int hsso_prime_split(uint64_t p, uint64_t* a, uint64_t* b);

/*
 * Exact Heron check of a root triple: returns 1 and |4A| when (l^2, m^2, n^2) is a
 * nondegenerate triangle with integer area, else 0.
 */
int hsso_heron(uint64_t l, uint64_t m, uint64_t n, uint64_t* area4_lo, uint64_t* area4_hi);

#ifdef __cplusplus
}
#endif
#endif
