"""Named workloads for the Gaussian enumeration core (root positions at a LIMIT).

Each workload is a list of root positions ``l`` searched against every partner
``m`` in ``[l, LIMIT]``.  The sets mirror the reference's regression and
benchmark intervals:

* ``full5000``   LIMIT 5000, l in [1, 5000]        (scripts/smoke_test.py:9-16)
* ``known2``     LIMIT 70000, l = 11789             (scripts/smoke_test.py:17-24)
* ``dense1e5``   LIMIT 100000, l in [1, 100000]
* ``slice512``   LIMIT 10^7, l in [7500000, 7500511] (evidence/benchmarks/
                 gpu-opt-benchmark.json:15-20, the reference interval)
* ``uneven1536`` LIMIT 10^7: slice512 + 512 heavy roots + 512 light roots drawn
                 from [9*10^6, 10^7] with a fixed seed.

The served stream (``served_stream``) is a seeded sequence of interval
requests: LIMIT drawn from {10^5, 10^6, 10^7}, length log-uniform in
[1, 4096], FIRST_L uniform in [1, LIMIT - length + 1], Poisson arrivals.

Heavy roots have at least four distinct prime factors p = 1 (mod 4), so their
Gaussian weight w(l) = prod(4e+1) over those primes is at least 625.  Light roots
are prime powers with w in {1, 5, 9}.  Weights follow
experiments/heavy/README.md:9.
"""
from __future__ import annotations

import numpy as np

SEED = 20261203
HEAVY_LO, HEAVY_HI = 9_000_000, 10_000_000


def spf_table(n: int) -> np.ndarray:
    """Smallest prime factor of every i <= n (spf[0] = 0, spf[1] = 1)."""
    spf = np.zeros(n + 1, dtype=np.int64)
    spf[1] = 1
    for p in range(2, int(n ** 0.5) + 1):
        if spf[p] == 0:
            block = spf[p * p:: p]
            block[block == 0] = p
            spf[p * p:: p] = block
    rest = np.nonzero(spf == 0)[0]
    spf[rest[rest >= 2]] = rest[rest >= 2]
    return spf


def factor(i: int, spf: np.ndarray) -> list[tuple[int, int]]:
    out = []
    while i > 1:
        p = int(spf[i])
        e = 0
        while i % p == 0:
            i //= p
            e += 1
        out.append((p, e))
    return out


def weight(i: int, spf: np.ndarray) -> int:
    w = 1
    for p, e in factor(i, spf):
        if p % 4 == 1:
            w *= 4 * e + 1
    return w


def uneven_roots(spf: np.ndarray | None = None, seed: int = SEED):
    """Return (heavy, light) lists of 512 roots each, sorted ascending."""
    if spf is None:
        spf = spf_table(HEAVY_HI)
    heavy_pool, light_pool = [], []
    for i in range(HEAVY_LO, HEAVY_HI + 1):
        f = factor(i, spf)
        split = [(p, e) for p, e in f if p % 4 == 1]
        if len(split) >= 4:
            heavy_pool.append(i)
        elif len(f) == 1:
            p, e = f[0]
            w = 4 * e + 1 if p % 4 == 1 else 1
            if w in (1, 5, 9):
                light_pool.append(i)
    rng = np.random.default_rng(seed)
    heavy = sorted(int(x) for x in rng.choice(heavy_pool, 512, replace=False))
    light = sorted(int(x) for x in rng.choice(light_pool, 512, replace=False))
    return heavy, light


def workload(name: str):
    """Return (limit, list_of_roots) for a named workload."""
    if name == "full5000":
        return 5000, list(range(1, 5001))
    if name == "known2":
        return 70000, [11789]
    if name == "dense1e5":
        return 100000, list(range(1, 100001))
    if name == "slice512":
        return 10_000_000, list(range(7_500_000, 7_500_512))
    if name == "uneven1536":
        heavy, light = uneven_roots()
        return 10_000_000, list(range(7_500_000, 7_500_512)) + heavy + light
    raise KeyError(name)


STREAM_LIMITS = (100_000, 1_000_000, 10_000_000)


def served_stream(n: int = 40, rate_hz: float = 5.0, seed: int = SEED):
    """The first n requests of the seeded served stream.

    Returns a list of (arrival_seconds, limit, first, last).  Request i draws
    its LIMIT, length and FIRST_L in that order from one generator; arrival
    gaps are exponential with mean 1/rate_hz, drawn from a second generator so
    that the request set does not depend on the rate.  With the default seed
    the first 40 requests hold about 1.2e12 planner cells of work.
    """
    rng = np.random.default_rng(seed)
    gaps = np.random.default_rng(seed + 1)
    out, t = [], 0.0
    for _ in range(n):
        limit = int(rng.choice(STREAM_LIMITS))
        length = int(round(2 ** rng.uniform(0, 12)))
        first = int(rng.integers(1, limit - length + 2))
        t += float(gaps.exponential(1.0 / rate_hz))
        out.append((t, limit, first, first + length - 1))
    return out
