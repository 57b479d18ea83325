"""Window mode on the GPU (cell_win_kernel: only the left entries whose leaves
can pass the range test are evaluated) against the reference's counts.

The scheduling knobs decide which tiles run in window mode; counts must not
depend on them.  Each case runs with window mode forced onto every sortable
left set, and with it switched off, and must give the reference's fixtures
(tests/golden, written by the compiled reference) or aggregates
(evidence/benchmarks/*.json) either way.
"""
from __future__ import annotations

import pytest

from conftest import result_lines

pytestmark = pytest.mark.gpu

MODES = {
    "everywhere": dict(win_max_frac=1.0),
    "off": dict(win_max_frac=0.0),
}


@pytest.fixture(scope="module")
def hss_mod():
    import hss
    return hss


@pytest.fixture(scope="module")
def ctx_small(hss_mod):
    c = hss_mod.Context(100_000)
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx_big(hss_mod):
    c = hss_mod.Context(10_000_000)
    yield c
    c.close()


class tuned:
    def __init__(self, ctx, **kw):
        self.ctx, self.kw = ctx, kw

    def __enter__(self):
        self.old = self.ctx.set_tuning(**self.kw)
        return self.ctx

    def __exit__(self, *exc):
        self.ctx.set_tuning(**self.old)


@pytest.mark.parametrize("mode", list(MODES))
def test_small_limits_golden(ctx_small, golden, mode):
    with tuned(ctx_small, **MODES[mode]) as ctx:
        for limit in (5000, 70000):
            cases = [g for g in golden.values() if g["name"].startswith(f"l{limit}_")]
            r = ctx.search(limit, [g["first"] for g in cases])
            for g, got in zip(cases, r.candidates):
                assert got == g["candidates"], (mode, g["name"])
            assert r.hits == sorted({h for g in cases for h in result_lines(g["results"])})
            st = r.stats
            print(mode, limit, "window tiles", st["window_tiles"], "of", st["tiles"], flush=True)
            assert (st["window_tiles"] > 0) == (mode == "everywhere")
        g = golden["full5000"]
        r = ctx.search(5000, range(1, 5001))
        assert r.total == g["candidates"] and sum(r.leaves) == 170_863_629
        assert r.hits == result_lines(g["results"])


@pytest.mark.parametrize("mode", list(MODES))
def test_intervals_1e5_golden(ctx_small, golden, mode):
    with tuned(ctx_small, **MODES[mode]) as ctx:
        for g in golden.values():
            if g["name"].startswith("i100000_"):
                tot, hits = ctx.search_interval(100000, g["first"], g["last"])
                assert tot == g["candidates"], (mode, g["name"])
                assert hits == result_lines(g["results"])


@pytest.mark.parametrize("mode", list(MODES))
def test_limit_1e6_1e7_golden(ctx_big, golden, mode):
    with tuned(ctx_big, **MODES[mode]) as ctx:
        for limit, prefix in ((1_000_000, "l1000000_"), (10_000_000, "l1e7_")):
            cases = [g for g in golden.values() if g["name"].startswith(prefix)]
            r = ctx.search(limit, [g["first"] for g in cases])
            for g, got in zip(cases, r.candidates):
                assert got == g["candidates"], (mode, g["name"])
            assert r.hits == []


AGGREGATES = [
    ((7_500_000, 7_500_127), 10_369_431_230),   # evidence/benchmarks/v45-ab.json:11
    ((1_116_389, 1_116_389), 7_175_008_050),    # evidence/benchmarks/heavy-benchmark.json:3-7
    ((1_447_225, 1_447_225), 41_795_211_005),   # evidence/benchmarks/heavy-benchmark.json:39-44
]


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("interval,expect", AGGREGATES)
def test_reference_aggregates(ctx_big, mode, interval, expect):
    with tuned(ctx_big, **MODES[mode]) as ctx:
        tot, hits = ctx.search_interval(10_000_000, *interval)
        print(mode, interval, tot, ctx.stats(), flush=True)
        assert tot == expect and hits == []


def test_default_mode_equals_off_on_heavy_roots(ctx_big):
    """The uneven workload's heavy roots (at least four split primes, in
    [9e6, 1e7]): default tuning against window mode off, per position."""
    import workloads
    heavy, light = workloads.uneven_roots()
    roots = heavy[::16] + light[::64]
    r_def = ctx_big.search(10_000_000, roots)
    st = r_def.stats
    with tuned(ctx_big, **MODES["off"]) as ctx:
        r_off = ctx.search(10_000_000, roots)
    print("default", st["cells_ms"], st["window_tiles"], "off", r_off.stats["cells_ms"], flush=True)
    assert st["window_tiles"] > 0
    assert r_def.candidates == r_off.candidates
    assert r_def.leaves == r_off.leaves and r_def.pairs == r_off.pairs


def test_tuning_validation(hss_mod, ctx_small):
    for bad in (dict(tiles_per_wave=0), dict(win_max_frac=1.5), dict(win_item_cost=0.0)):
        with pytest.raises(hss_mod.HssError):
            ctx_small.set_tuning(**bad)
    t = ctx_small.get_tuning()
    assert t["tiles_per_wave"] >= 1 and all(0.0 <= f <= 1.0 for f in t["win_max_frac"])


def test_autotune_keeps_counts(hss_mod, ctx_big):
    """The autotuner times window thresholds per weight class and tile sizes on
    calibration roots; every setting it tries must give the same candidate
    total (it fails otherwise), and the tuning it applies still gives the
    reference's counts."""
    import workloads
    heavy, light = workloads.uneven_roots(seed=workloads.SEED + 1)
    calib = heavy[::32] + light[::64] + list(range(7_400_000, 7_400_016))
    old = ctx_big.get_tuning()
    try:
        tuned_t, ms = ctx_big.autotune(10_000_000, calib, repeats=1)
        print("autotuned", tuned_t, ms, flush=True)
        assert ms > 0 and tuned_t == ctx_big.get_tuning()
        tot, _ = ctx_big.search_interval(10_000_000, 7_500_000, 7_500_127)
        assert tot == 10_369_431_230  # evidence/benchmarks/v45-ab.json:11
    finally:
        ctx_big.set_tuning(**old)


def test_largest_tables_window_vs_plain(hss_mod):
    """Tables at the largest N_max (2^26): positions at LIMIT = 2^26 (partner
    indices k up to 2^26, the widest packed queue ids, the largest half-sets)
    give the same per-position counts with window mode forced everywhere and
    switched off, and the same hits (none expected this high)."""
    ctx = hss_mod.Context(1 << 26)
    try:
        limit = 1 << 26
        roots = [limit - 1, limit - 7, 66_000_000, 65_536_001]
        with tuned(ctx, **MODES["everywhere"]):
            r_win = ctx.search(limit, roots)
        with tuned(ctx, **MODES["off"]):
            r_off = ctx.search(limit, roots)
        print("2^26", r_win.candidates, r_win.stats["window_tiles"], flush=True)
        assert r_win.stats["window_tiles"] > 0
        assert r_win.candidates == r_off.candidates and r_win.leaves == r_off.leaves
        assert r_win.pairs == r_off.pairs and r_win.hits == r_off.hits
        assert sum(r_win.candidates) > 0
    finally:
        ctx.close()
