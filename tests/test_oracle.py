"""Pin the CPU oracle (oracle/hss_oracle.c) before trusting it.

* Golden fixtures written by the compiled reference search (tests/golden):
  regression cases, per-position counts at several limits, short intervals.
* Known-answer tests of scripts/verify.py:26-28 for the Heron check.
* Closed forms: leaf totals (sum of (W+1)/2) quoted by the survey for full5000
  and known2; Gaussian weights and fixed parts from brute-force factoring;
  prime splits a^2 + b^2 = p.
* Brute force at tiny limits: every Gaussian integer of norm (2 l^2 m^2)^2 is
  found by scanning re, and the class-weighted count over all of them must
  equal the oracle's enumeration of one representative per class.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import result_lines


def test_full5000_and_known2(oracle, golden):
    for name in ("full5000", "known2"):
        g = golden[name]
        tot, per, hits, pairs = oracle.interval(g["limit"], g["first"], g["last"], threads=8)
        assert tot == g["candidates"], name
        assert pairs == g["pairs"], name
        assert hits == result_lines(g["results"]), name


@pytest.mark.parametrize("limit", [5000, 70000])
def test_per_position_small_limits(oracle, golden, limit):
    cases = [g for g in golden.values() if g["name"].startswith(f"l{limit}_")]
    assert len(cases) >= 20
    for g in cases:
        c, hits = oracle.position(limit, g["first"])
        assert c == g["candidates"], g["name"]
        assert hits == result_lines(g["results"]), g["name"]


def test_intervals_limit_1e5(oracle, golden):
    for g in golden.values():
        if g["name"].startswith("i100000_"):
            tot, per, hits, pairs = oracle.interval(100000, g["first"], g["last"], threads=8)
            assert tot == g["candidates"], g["name"]
            assert pairs == g["pairs"], g["name"]


def test_per_position_limit_1e6(oracle, golden):
    for g in golden.values():
        if g["name"].startswith("l1000000_"):
            c, hits = oracle.position(1_000_000, g["first"])
            assert c == g["candidates"], g["name"]


def test_light_and_top_positions_limit_1e7(oracle, golden):
    names = sorted(n for n in golden if n.startswith("l1e7_light_"))[:3] + ["l1e7_top_9999999", "l1e7_top_10000000"]
    for name in names:
        g = golden[name]
        c, hits = oracle.position(10_000_000, g["first"])
        assert c == g["candidates"], name


def test_verify_known_answers(oracle):
    lo, hi = (ctypes_u64(), ctypes_u64())
    for (l, m, n, area) in [(1853, 4380, 4427, 32918611718880), (11789, 68104, 68595, 284239560530875680)]:
        assert oracle.lib.hsso_heron(l, m, n, lo, hi) == 1
        assert ((hi.value << 64) | lo.value) == 4 * area
    # degenerate, non-integral and non-triangle triples are rejected
    for (l, m, n) in [(1, 1, 2), (3, 4, 5), (1853, 4380, 4428), (2, 3, 100)]:
        assert oracle.lib.hsso_heron(l, m, n, lo, hi) == 0


def ctypes_u64():
    import ctypes
    return ctypes.c_uint64()


def test_leaf_closed_forms(oracle):
    # survey: full5000 enumerates 170,863,629 leaves, known2 1,580,864
    total = sum(oracle.work(5000, l)[1] for l in range(1, 5001))
    assert total == 170_863_629
    assert oracle.work(70000, 11789)[1] == 1_580_864


def factor(i: int) -> dict[int, int]:
    out, p = {}, 2
    while p * p <= i:
        while i % p == 0:
            out[p] = out.get(p, 0) + 1
            i //= p
        p += 1
    if i > 1:
        out[i] = out.get(i, 0) + 1
    return out


def test_weights_fixed_and_inert(oracle):
    ctx = oracle.ctx(5000)
    for i in range(1, 3001):
        f = factor(i)
        w = math.prod(4 * e + 1 for p, e in f.items() if p % 4 == 1)
        fixed = math.prod(p ** (2 * e) for p, e in f.items() if p % 4 != 1)
        assert oracle.lib.hsso_weight(ctx, i) == w
        assert oracle.lib.hsso_fixed(ctx, i) == fixed
        assert bool(oracle.lib.hsso_has_inert(ctx, i)) == any(p % 4 == 3 for p in f)


def test_prime_splits(oracle):
    a, b = ctypes_u64(), ctypes_u64()
    primes = [p for p in range(5, 200000, 4) if all(p % q for q in range(2, int(p ** 0.5) + 1))]
    for p in primes:
        assert oracle.lib.hsso_prime_split(p, a, b) == 1
        assert a.value ** 2 + b.value ** 2 == p and a.value % 2 == 1 and b.value % 2 == 0


def _filters_count(zre: int, zim: int, l: int, m: int, N: int) -> int:
    """Candidates of one Gaussian integer z under the reference filter chain
    (scripts-level restatement of search_v18.cpp:436-456, counting only)."""
    lo, hi = l ** 4 + m ** 4 - N ** 4, l ** 4
    need9 = l % 3 == 0 or m % 3 == 0
    need5 = l % 5 != 0 and m % 5 != 0
    cnt = 0
    orients = [(zre, zim)] + ([(zim, zre)] if zre != zim and zre != -zim else [])
    for d, a4 in orients:
        if a4 == 0 or a4 % 8:
            continue
        pos, neg = lo <= d <= hi, lo <= -d <= hi
        if not (pos or neg):
            continue
        if a4 % 3 or (need9 and a4 % 9) or (need5 and a4 % 5):
            continue
        cnt += int(pos) + int(neg and d != 0)
    return cnt


@pytest.mark.parametrize("N", [14, 22])
def test_bruteforce_enumeration_tiny(oracle, N):
    """Every Gaussian integer of norm (2 l^2 m^2)^2 by scanning re; classes under
    units and conjugation have 8 members except the real one (4), so the
    reference count is (sum over all z + 4 * c(real)) / 8."""
    for l in range(1, N + 1):
        expect = 0
        for m in range(l, N + 1):
            if l % 2 == 0 and m % 2 == 0:
                continue
            if any(p % 4 == 3 and l % p == 0 and m % p == 0 for p in factor(math.gcd(l, m))):
                continue
            if not any(p % 4 == 1 for p in factor(l * m)):
                continue  # a single real representation: the reference skips the pair
            R = 2 * l * l * m * m
            re = np.arange(-R, R + 1, dtype=np.int64)
            im2 = R * R - re * re
            im = np.round(np.sqrt(im2.astype(np.float64))).astype(np.int64)
            ok = im * im == im2
            zs = []
            for x, y in zip(re[ok].tolist(), im[ok].tolist()):
                zs.append((x, y))
                if y:
                    zs.append((x, -y))
            total = sum(_filters_count(x, y, l, m, N) for x, y in zs)
            real = _filters_count(R, 0, l, m, N)
            assert (total + 4 * real) % 8 == 0
            expect += (total + 4 * real) // 8
        c, _ = oracle.position(N, l, nmax=N)
        assert c == expect, (N, l)
