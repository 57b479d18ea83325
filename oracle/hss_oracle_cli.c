/*
 * hss_oracle_cli.c -- command-line driver of the CPU oracle (test infrastructure only).
 *
 *   hss_oracle_cli LIMIT FIRST_L LAST_L THREADS [--per-l]
 *
 * Prints "pairs P candidates C seconds S" and one "HIT l m n area A" line per hit;
 * with --per-l also "L l candidates C" for every position.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "hss_oracle.h"

static void print_u128(unsigned __int128 v) {
    char buf[48];
    int i = 47;
    buf[i] = 0;
    if (!v) buf[--i] = '0';
    while (v) {
        buf[--i] = (char)('0' + (int)(v % 10));
        v /= 10;
    }
    fputs(buf + i, stdout);
}

int main(int argc, char** argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: hss_oracle_cli LIMIT FIRST_L LAST_L THREADS [--per-l]\n");
        return 2;
    }
    uint64_t limit = strtoull(argv[1], 0, 10), first = strtoull(argv[2], 0, 10),
             last = strtoull(argv[3], 0, 10);
    uint32_t threads = (uint32_t)strtoul(argv[4], 0, 10);
    int per = argc > 5 && !strcmp(argv[5], "--per-l");
    if (!limit || !first || first > last || last > limit || !threads) return 2;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    hsso_ctx* c = hsso_create(limit);
    if (!c) return 1;
    uint64_t* pl = per ? calloc(last - first + 1, 8) : NULL;
    hsso_hit hits[256];
    uint32_t nh = 0;
    uint64_t pairs = 0;
    uint64_t cand = hsso_interval(c, limit, first, last, threads, pl, hits, 256, &nh, &pairs);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    double s = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    printf("pairs %llu candidates %llu seconds %.3f\n", (unsigned long long)pairs,
           (unsigned long long)cand, s);
    for (uint32_t i = 0; i < nh && i < 256; ++i) {
        unsigned __int128 a = ((unsigned __int128)hits[i].area4_hi << 64) | hits[i].area4_lo;
        printf("HIT %llu %llu %llu area ", (unsigned long long)hits[i].l,
               (unsigned long long)hits[i].m, (unsigned long long)hits[i].n);
        print_u128(a / 4);
        printf("\n");
    }
    if (pl)
        for (uint64_t l = first; l <= last; ++l)
            printf("L %llu candidates %llu\n", (unsigned long long)l,
                   (unsigned long long)pl[l - first]);
    hsso_destroy(c);
    free(pl);
    return 0;
}
