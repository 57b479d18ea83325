"""The kernels' arithmetic, run on the CPU (tests/hostcheck), against the
reference fixtures: the plan (shared-prime signatures, tiles), the half-set
cache and left sets, the FP64 range band, area residues, root-proximity filter
and exact stage must reproduce the reference's counts and hits."""
from __future__ import annotations

import ctypes
import random
import subprocess

import pytest

from conftest import PKG, ROOT, result_lines

LIB = ROOT / "tests" / "hostcheck" / "libhss_hostcheck.so"


class HcHit(ctypes.Structure):
    _fields_ = [("l", ctypes.c_uint64), ("m", ctypes.c_uint64), ("n", ctypes.c_uint64),
                ("lo", ctypes.c_uint64), ("hi", ctypes.c_uint64)]


class HcStats(ctypes.Structure):
    _fields_ = [("deferred_full", ctypes.c_uint64), ("deferred_root", ctypes.c_uint64),
                ("raw_roots", ctypes.c_uint64), ("cells", ctypes.c_uint64), ("window_tiles", ctypes.c_uint64),
                ("window_items", ctypes.c_uint64), ("class_mismatch", ctypes.c_uint64),
                ("window_units", ctypes.c_uint64)]


DEFAULT_TUNING = (0.5, 3, 8)  # threshold for weight classes >= 3 (H >= 64), tiles_per_wave (hss_plan.h Tuning)


@pytest.fixture(scope="module")
def hc():
    if not LIB.is_file():
        subprocess.run(["make", "-C", str(PKG), "hostcheck"], check=True, capture_output=True)
    lib = ctypes.CDLL(str(LIB))
    lib.hc_set_tuning.argtypes = [ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32]
    lib.hc_set_tuning.restype = None
    lib.hc_set_tuning(*DEFAULT_TUNING)
    return lib


@pytest.fixture
def window_everywhere(hc):
    """Window mode on every left set that can be sorted, on every tile."""
    hc.hc_set_tuning(1.0, 0, 8)
    yield
    hc.hc_set_tuning(*DEFAULT_TUNING)


@pytest.fixture
def window_off(hc):
    hc.hc_set_tuning(0.0, 0, 8)
    yield
    hc.hc_set_tuning(*DEFAULT_TUNING)


def hc_search(hc, nmax, limit, roots):
    n = len(roots)
    R = (ctypes.c_uint64 * n)(*roots)
    c, lv, pr = (ctypes.c_uint64 * n)(), (ctypes.c_uint64 * n)(), (ctypes.c_uint64 * n)()
    H = (HcHit * 64)()
    nh = ctypes.c_uint32()
    st = HcStats()
    rc = hc.hc_search(ctypes.c_uint64(nmax), ctypes.c_uint64(limit), R, n, c, lv, pr, H, 64, ctypes.byref(nh),
                      ctypes.byref(st))
    assert rc == 0
    # the class rule (hss_math.h) decides every evaluated leaf's area tests as the residues do
    assert st.class_mismatch == 0
    hits = [(h.l, h.m, h.n, ((h.hi << 64) | h.lo) // 4) for h in H[: nh.value]]
    return list(c), list(lv), list(pr), hits, st


def test_known2(hc, golden):
    g = golden["known2"]
    c, lv, pr, hits, st = hc_search(hc, 70000, 70000, [11789])
    assert c == [g["candidates"]]
    assert lv == [1_580_864]
    assert pr == [g["pairs"]]  # every pair of l = 11789 reaches the enumeration
    assert hits == result_lines(g["results"])


def test_full5000(hc, golden):
    g = golden["full5000"]
    c, lv, pr, hits, st = hc_search(hc, 5000, 5000, list(range(1, 5001)))
    assert sum(c) == g["candidates"]
    assert sum(lv) == 170_863_629
    assert sum(pr) == 6_994_365  # evaluated pairs quoted by the survey
    assert hits == result_lines(g["results"])
    # the FP64 band leaves almost nothing undecided; root proximity is rare
    assert st.deferred_full < 100
    assert st.deferred_root < sum(c) // 1000


@pytest.mark.parametrize("limit", [5000, 70000])
def test_per_position(hc, golden, limit):
    cases = [g for g in golden.values() if g["name"].startswith(f"l{limit}_")]
    roots = [g["first"] for g in cases]
    c, lv, pr, hits, st = hc_search(hc, limit, limit, roots)
    for g, got in zip(cases, c):
        assert got == g["candidates"], g["name"]


def test_order_and_duplicates(hc):
    roots = [4380, 17, 1853, 17, 2, 4999]
    a = hc_search(hc, 5000, 5000, roots)[0]
    b = hc_search(hc, 5000, 5000, list(reversed(roots)))[0]
    assert a == list(reversed(b))
    assert a[1] == a[3]


def test_tables_above_limit(hc, golden):
    """Tables built for a larger N_max serve a smaller LIMIT unchanged."""
    g = golden["l5000_1853"]
    c = hc_search(hc, 70000, 5000, [1853])[0]
    assert c == [g["candidates"]]


def test_interval_limit_1e5(hc, golden):
    g = golden["i100000_99901_100000"]
    c = hc_search(hc, 100000, 100000, list(range(g["first"], g["last"] + 1)))[0]
    assert sum(c) == g["candidates"]


def hc_search_mixed(hc, nmax, positions):
    n = len(positions)
    R = (ctypes.c_uint64 * n)(*[l for l, _ in positions])
    L = (ctypes.c_uint64 * n)(*[lim for _, lim in positions])
    c, lv, pr = (ctypes.c_uint64 * n)(), (ctypes.c_uint64 * n)(), (ctypes.c_uint64 * n)()
    H = (HcHit * 64)()
    nh = ctypes.c_uint32()
    rc = hc.hc_search_mixed(ctypes.c_uint64(nmax), R, L, n, c, lv, pr, H, 64, ctypes.byref(nh), None)
    assert rc == 0
    hits = [(h.l, h.m, h.n, ((h.hi << 64) | h.lo) // 4) for h in H[: nh.value]]
    return list(c), list(lv), list(pr), hits


def test_mixed_limits_in_one_batch(hc, golden):
    """Positions at different LIMITs planned into one batch count exactly as
    each LIMIT alone (the served form mixes requests of several LIMITs)."""
    positions = [(1853, 5000), (11789, 70000), (1853, 70000), (4999, 5000), (30000, 70000), (1, 5000)]
    c, lv, pr, hits = hc_search_mixed(hc, 70000, positions)
    for i, (l, lim) in enumerate(positions):
        a = hc_search(hc, 70000, lim, [l])
        assert (c[i], lv[i], pr[i]) == (a[0][0], a[1][0], a[2][0]), (l, lim)
    assert c[1] == golden["known2"]["candidates"]
    # the hit of l=1853 is found at LIMIT 5000 (n = 4427) and again at 70000
    assert hits.count((1853, 4380, 4427, 32918611718880)) == 1
    assert (11789, 68104, 68595, 284239560530875680) in hits


def test_estimate_bounds_leaves(hc):
    """The planner's work estimate (cells) bounds the enumerated leaves:
    every cell yields at most two leaves (u*v and u*conj(v))."""
    for l, lim in [(1, 5000), (1853, 5000), (11789, 70000), (4225, 70000), (65537, 70000)]:
        est = hc.hc_estimate
        est.restype = ctypes.c_double
        e = est(ctypes.c_uint64(70000), ctypes.c_uint64(l), ctypes.c_uint64(lim))
        lv = hc_search(hc, 70000, lim, [l])[1][0]
        assert 0 < lv <= 2 * e, (l, lim, lv, e)
    assert hc.hc_estimate(ctypes.c_uint64(70000), ctypes.c_uint64(5001), ctypes.c_uint64(5000)) == -1.0


def split(hc, nmax, limit, first, last, parts):
    out = (ctypes.c_uint64 * (parts + 1))()
    rc = hc.hc_split_by_work(ctypes.c_uint64(nmax), ctypes.c_uint64(limit), ctypes.c_uint64(first),
                             ctypes.c_uint64(last), ctypes.c_uint32(parts), out)
    return rc, list(out)


@pytest.mark.parametrize("limit,first,last,parts", [(70000, 1, 3000, 8), (70000, 60000, 70000, 5),
                                                    (5000, 1, 5000, 3), (70000, 11789, 11789, 4),
                                                    (70000, 100, 131, 1)])
def test_split_by_work(hc, limit, first, last, parts):
    """Interval split by the planner's work estimate (the sharding of an
    interval over GPUs): contiguous cover of [first, last]; every boundary is
    one past the position whose inclusive work prefix first reaches its share;
    no part exceeds its share by more than one position's work."""
    est = hc.hc_estimate
    est.restype = ctypes.c_double
    work = [est(ctypes.c_uint64(70000), ctypes.c_uint64(l), ctypes.c_uint64(limit)) for l in range(first, last + 1)]
    assert min(work) >= 0
    rc, b = split(hc, 70000, limit, first, last, parts)
    assert rc == 0
    assert b[0] == first and b[-1] == last + 1 and b == sorted(b)
    total = sum(work)
    prefix = [0.0]
    for w in work:
        prefix.append(prefix[-1] + w)  # prefix[i] = work of first .. first + i - 1
    for i in range(1, parts):
        x = b[i] - 1 - first  # index of the crossing position
        target = total * i / parts
        assert prefix[x + 1] >= target * (1 - 1e-12) and (x == 0 or prefix[x] < target * (1 + 1e-12)), (i, b)
    for i in range(parts):
        part = prefix[b[i + 1] - first] - prefix[b[i] - first]
        assert part <= total / parts + max(work) * (1 + 1e-12)


def plan_cover(hc, nmax, limit, roots, waves):
    n = len(roots)
    out = (ctypes.c_uint64 * 5)()
    rc = hc.hc_plan_cover(ctypes.c_uint64(nmax), ctypes.c_uint64(limit), (ctypes.c_uint64 * n)(*roots),
                          ctypes.c_uint32(n), ctypes.c_uint64(waves), out)
    assert rc == 0
    return list(out)


@pytest.mark.parametrize("waves", [64, 4096, 1 << 20])
@pytest.mark.parametrize("tuning", ["default", "window_everywhere", "window_off"])
def test_plan_tiles_cover_each_left_set_once(hc, waves, tuning, request):
    """The planner's tiles (both lists, any order) partition every left set's
    partner range [kmin, kmax]: no gap, no overlap, nothing past the range --
    for coarse and fine tiles (few or many resident waves) and for every
    window-mode setting.  A gap would lose counts and an overlap double them,
    which only the device parity tests would otherwise show."""
    if tuning != "default":
        request.getfixturevalue(tuning)
    roots = list(range(1, 3001, 7)) + [11789, 65000, 69997, 70000] + [2 * 3 * 5 * 13 * 17, 5 ** 4 * 13 * 2]
    plain, win, lefts, bad, rev = plan_cover(hc, 70000, 70000, roots, waves)
    assert lefts > len(roots) and plain + win >= lefts
    assert bad == 0 and rev == 0
    if tuning == "window_off":
        assert win == 0
    if tuning == "window_everywhere":
        assert win > 0


def test_split_by_work_rejects(hc):
    assert split(hc, 70000, 70000, 10, 9, 2)[0] != 0
    assert split(hc, 70000, 70000, 1, 10, 0)[0] != 0
    assert split(hc, 70000, 70001, 1, 10, 2)[0] != 0


@pytest.mark.parametrize("limit", [5000, 70000])
def test_window_mode_per_position_golden(hc, golden, window_everywhere, limit):
    """Every sortable left set in window mode: the per-position counts still
    equal the reference's (only leaves that cannot count are skipped)."""
    cases = [g for g in golden.values() if g["name"].startswith(f"l{limit}_")]
    roots = [g["first"] for g in cases]
    c, lv, pr, hits, st = hc_search(hc, limit, limit, roots)
    for g, got in zip(cases, c):
        assert got == g["candidates"], g["name"]
    assert st.window_tiles > 0 and st.window_items > 0


def test_window_mode_interval_1e5(hc, golden, window_everywhere):
    for name in ("i100000_99901_100000", "i100000_50000_50099", "i100000_4380_4479"):
        g = golden[name]
        c = hc_search(hc, 100000, 100000, list(range(g["first"], g["last"] + 1)))[0]
        assert sum(c) == g["candidates"], name


def test_window_mode_full5000(hc, golden, window_everywhere):
    g = golden["full5000"]
    c, lv, pr, hits, st = hc_search(hc, 5000, 5000, list(range(1, 5001)))
    assert sum(c) == g["candidates"]
    assert sum(lv) == 170_863_629
    assert hits == result_lines(g["results"])
    assert st.window_tiles > 0


def test_window_mode_skips_work_near_limit(hc, window_everywhere):
    """Roots near LIMIT have narrow range windows: far fewer leaves evaluated
    than enumerated, same counts as plain mode."""
    roots = [69300, 69615, 69700, 69905]  # 69615 = 3^2 * 5 * 7 * 13 * 17
    cw, lw, pw, hw, sw = hc_search(hc, 70000, 70000, roots)
    hc.hc_set_tuning(0.0, 0, 8)
    cp, lp, pp, hp, sp = hc_search(hc, 70000, 70000, roots)
    hc.hc_set_tuning(1.0, 0, 8)
    assert (cw, lw, pw, hw) == (cp, lp, pp, hp)
    assert sw.window_tiles > 0
    assert sw.cells < 0.5 * sp.cells, (sw.cells, sp.cells)


def test_class_rule(hc):
    """The class rule the kernels' area tests rest on, on random Gaussian
    products: Re z odd, 4 | Im z, exactly one part divisible by 3 and by 5,
    which one given by the xor of the factors' classes (conjugation keeps a
    factor's class).  A wrong class bit for one prime residue, a wrong parity
    of r or a product that does not xor fails it."""
    hc.hc_class_rule_mismatches.restype = ctypes.c_uint64
    hc.hc_class_rule_mismatches.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    assert hc.hc_class_rule_mismatches(200000, 20261204) == 0


def test_device_root_candidate(hc):
    """The device fourth-root test: two reciprocal square roots, one Newton step
    (root4_refine), then the single candidate n = rint(r) checked by
    |n^4 - zz| <= 2^-47 zz.  With reciprocal-square-root errors of 2^-22 each
    way and zz within 2^-49.3 of n^4 (the fast path's error bound), r stays
    within 1e-3 of n for n up to 2^26 and the check accepts; zz between fourth
    powers, or 2^-45 off n^4, is rejected.  Errors of 2^-16 still leave r
    within 0.25 of n (the refinement is quadratic)."""
    ref = hc.hc_root_refined
    ref.restype = ctypes.c_double
    ref.argtypes = [ctypes.c_double] * 3
    ok = hc.hc_root_candidate_ok
    ok.restype = ctypes.c_int
    ok.argtypes = [ctypes.c_double] * 2
    rng = random.Random(20261203)
    ns = [1, 2, 3, 7, 1000, 65535, 2**20 + 1, 9_999_991, 10**7, 2**26 - 1, 2**26]
    ns += [rng.randrange(1, 2**26 + 1) for _ in range(1500)]
    errs = (-2.0**-22, 0.0, 2.0**-22)
    worst = 0.0
    for n in ns:
        z4 = float(n**4)
        for zz in (z4, z4 * (1 + 2.0**-49.3), z4 * (1 - 2.0**-49.3)):
            for dt in errs:
                for dr in errs:
                    r = ref(zz, dt, dr)
                    worst = max(worst, abs(r - n))
                    assert ok(zz, r), (n, zz, dt, dr, r)
        far = ref(z4, 2.0**-16, -2.0**-16)
        assert abs(far - n) < 0.25, (n, far)
        between = float((n + 0.5)**4)
        assert not ok(between, ref(between, 0.0, 0.0))
        off = z4 * (1 + 2.0**-45)
        assert not ok(off, ref(off, 0.0, 0.0)), n
    assert worst < 1e-3, worst


def test_one_term_tiles(hc):
    """Tiles whose smallest partner already has lo > 2 tau skip the second-sign
    range compares (tile_one_term).  From the switch-over partner on, every
    pair's context has h2_w < 0, so skipping changes no decision; a few
    partners before it h2_w >= 0 (the rule is tight, not vacuous)."""
    first = hc.hc_one_term_first
    first.restype = ctypes.c_uint32
    first.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]
    bad = hc.hc_one_term_violations
    bad.restype = ctypes.c_uint64
    bad.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
    cases = [(7_500_000, 10_000_000, 1), (7_500_511, 10_000_000, 5), (9_000_001, 10_000_000, 1),
             (11789, 70_000, 1), (11789, 70_000, 13), (60_000, 70_000, 1), (2**25, 2**26, 1)]
    for l, limit, q in cases:
        k_max = limit // q
        k = first(l, limit, q, k_max)
        assert k <= k_max, (l, limit, q)
        for k0 in range(max(1, k - 2), k + 3):
            assert bad(l, limit, q, k0, min(k_max, k0 + 2000), 0) == 0, (l, limit, q, k0)
        assert bad(l, limit, q, k, k_max, 1) == 0, (l, limit, q)
        assert k > 3 and bad(l, limit, q, k - 3, k - 1, 1) > 0, (l, limit, q, k)


def test_window_fraction_bounds(hc):
    """The planner's window measure is a fraction of the circle, 1 where the
    range test cannot fail (small m, large LIMIT) and small near LIMIT."""
    f = hc.hc_window_fraction
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]
    assert f(10, 10_000_000, 1, 10, 20) == 1.0
    near = f(9_990_000, 10_000_000, 1, 9_990_000, 10_000_000)
    assert 0 < near < 0.02
    assert 0 < f(7_500_000, 10_000_000, 1, 7_500_000, 10_000_000) < 0.6
