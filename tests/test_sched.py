"""Admission and batch formation of the served form (hss_sched.cpp), driven
on the CPU through tests/hostcheck: FIFO admission bounded by request slots
and by the device state pool, every position scheduled exactly once and in
order, an equal work share per request in every launch, completion and the
finished prefix with launches that finish out of order."""
from __future__ import annotations

import ctypes
import random
import subprocess

import pytest

from conftest import PKG, ROOT

LIB = ROOT / "tests" / "hostcheck" / "libhss_hostcheck.so"
u64, u32, dbl = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_double


@pytest.fixture(scope="module")
def hc():
    if not LIB.is_file():
        subprocess.run(["make", "-C", str(PKG), "hostcheck"], check=True, capture_output=True)
    lib = ctypes.CDLL(str(LIB))
    lib.hc_sched_new.restype = ctypes.c_void_p
    lib.hc_sched_new.argtypes = [u32, u64, dbl, u32]
    for name in ("hc_sched_free", "hc_sched_retire"):
        getattr(lib, name).restype = None
    vp = ctypes.c_void_p
    lib.hc_sched_free.argtypes = [vp]
    lib.hc_sched_enqueue.argtypes = [vp, u64, u32]
    lib.hc_sched_admit.argtypes = [vp, ctypes.POINTER(u64), u32]
    lib.hc_sched_admit.restype = u32
    lib.hc_sched_set_costs.argtypes = [vp, u64, ctypes.POINTER(dbl), u32]
    lib.hc_sched_set_costs.restype = None
    lib.hc_sched_next.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u32), ctypes.POINTER(u32), u32]
    lib.hc_sched_next.restype = u32
    lib.hc_sched_finished.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u32), ctypes.POINTER(u32), u32,
                                      ctypes.POINTER(u64), u32]
    lib.hc_sched_finished.restype = u32
    lib.hc_sched_retire.argtypes = [vp, u64]
    lib.hc_sched_cancel.argtypes = [vp, u64]
    lib.hc_sched_cancel.restype = ctypes.c_int
    for name in ("hc_sched_pool_offset", "hc_sched_prefix"):
        getattr(lib, name).argtypes = [vp, u64]
        getattr(lib, name).restype = u32
    lib.hc_sched_pool_free.argtypes = [vp]
    lib.hc_sched_pool_free.restype = u64
    for name in ("hc_sched_queued", "hc_sched_active"):
        getattr(lib, name).argtypes = [vp]
        getattr(lib, name).restype = u32
    return lib


class Sched:
    def __init__(self, lib, slots, pool, work, maxpos=1 << 16):
        self.lib, self.h = lib, lib.hc_sched_new(slots, pool, work, maxpos)

    def close(self):
        self.lib.hc_sched_free(self.h)

    def enqueue(self, rid, npos):
        return bool(self.lib.hc_sched_enqueue(self.h, rid, npos))

    def admit(self):
        out = (u64 * 4096)()
        n = self.lib.hc_sched_admit(self.h, out, 4096)
        return list(out[:n])

    def set_costs(self, rid, costs):
        arr = (dbl * len(costs))(*costs)
        self.lib.hc_sched_set_costs(self.h, rid, arr, len(costs))

    def next(self):
        ids, b, e = (u64 * 4096)(), (u32 * 4096)(), (u32 * 4096)()
        n = self.lib.hc_sched_next(self.h, ids, b, e, 4096)
        return [(ids[i], b[i], e[i]) for i in range(n)]

    def finished(self, segs):
        n = len(segs)
        ids = (u64 * max(n, 1))(*[s[0] for s in segs])
        b = (u32 * max(n, 1))(*[s[1] for s in segs])
        e = (u32 * max(n, 1))(*[s[2] for s in segs])
        out = (u64 * 4096)()
        k = self.lib.hc_sched_finished(self.h, ids, b, e, n, out, 4096)
        return list(out[:k])

    def __getattr__(self, name):
        fn = getattr(self.lib, "hc_sched_" + name)
        return lambda *a: fn(self.h, *a)


def test_rejects_requests_that_never_fit(hc):
    s = Sched(hc, 4, 100, 1e9)
    assert not s.enqueue(1, 0)
    assert not s.enqueue(2, 101)
    assert s.enqueue(3, 100)
    assert s.admit() == [3]
    s.close()


def test_admission_by_memory_and_slots(hc):
    s = Sched(hc, 3, 1000, 1e9)
    for rid, npos in [(1, 600), (2, 300), (3, 200), (4, 50), (5, 10)]:
        assert s.enqueue(rid, npos)
    # 600 + 300 fit, 200 does not (pool), and FIFO keeps 4 and 5 behind it
    assert s.admit() == [1, 2]
    assert s.queued() == 3 and s.pool_free() == 100
    offs = {r: s.pool_offset(r) for r in (1, 2)}
    assert sorted([(offs[1], 600), (offs[2], 300)]) == [(0, 600), (600, 300)]
    for r, n in [(1, 600), (2, 300)]:
        s.set_costs(r, [1.0] * n)
    segs = s.next()
    assert {x[0] for x in segs} == {1, 2}
    assert s.finished(segs) == [1, 2]
    s.retire(1)
    # 600 free again: 3 and 4 enter, the slot limit (3) holds 5 back
    assert s.admit() == [3, 4]
    assert s.active() == 3 and s.queued() == 1
    s.retire(2)
    assert s.admit() == [5]
    s.close()


def test_pool_ranges_merge_on_retire(hc):
    s = Sched(hc, 16, 100, 1e9)
    for rid in range(1, 6):
        s.enqueue(rid, 20)
    assert s.admit() == [1, 2, 3, 4, 5]
    assert s.pool_free() == 0
    for rid in (2, 4, 3):  # frees 20..80 in pieces that must merge
        s.retire(rid)
    s.enqueue(6, 60)
    assert s.admit() == [6]
    assert s.pool_offset(6) == 20
    s.close()


def simulate(hc, seed, n_req, slots, pool, work, reorder):
    """Arrivals interleaved with launches; up to two launches in flight, the
    older sometimes finishing last (a launch that is run again)."""
    rng = random.Random(seed)
    s = Sched(hc, slots, pool, work)
    reqs = {}
    for rid in range(1, n_req + 1):
        npos = rng.choice([1, 2, 5, 17, 64, 300, 1000])
        npos = min(npos, pool)
        reqs[rid] = [rng.choice([1.0, 3.0, 50.0, 1e3]) * rng.random() + 0.5 for _ in range(npos)]
    pending = list(reqs)
    scheduled = {rid: [] for rid in reqs}
    inflight, done, admitted_order, retired = [], set(), [], set()
    fin_ranges = {rid: [] for rid in reqs}
    active_ranges = {}
    steps = 0
    while len(retired) < n_req:
        steps += 1
        assert steps < 100000
        # arrivals
        for _ in range(rng.randint(0, 3)):
            if pending:
                rid = pending.pop(0)
                assert s.enqueue(rid, len(reqs[rid]))
        adm = s.admit()
        for rid in adm:
            admitted_order.append(rid)
            off = s.pool_offset(rid)
            rng_ = (off, off + len(reqs[rid]))
            assert 0 <= rng_[0] and rng_[1] <= pool
            for o in active_ranges.values():
                assert rng_[1] <= o[0] or o[1] <= rng_[0], "pool ranges overlap"
            active_ranges[rid] = rng_
            s.set_costs(rid, reqs[rid])
        assert s.active() <= slots
        if len(inflight) < 2:
            live = [rid for rid in active_ranges if rid not in done and
                    sum(e - b for b, e in scheduled[rid]) < len(reqs[rid])]
            segs = s.next()
            if segs:
                # every live request gets work in every launch
                assert sorted(x[0] for x in segs) == sorted(live)
                firsts = 0.0
                total = 0.0
                for rid, b, e in segs:
                    prev_end = scheduled[rid][-1][1] if scheduled[rid] else 0
                    assert b == prev_end and e > b
                    scheduled[rid].append((b, e))
                    total += sum(reqs[rid][b:e])
                    firsts += reqs[rid][b]
                assert total <= work + firsts + 1e-6
                inflight.append(segs)
                continue
        if inflight:
            idx = 1 if (reorder and len(inflight) == 2 and rng.random() < 0.3) else 0
            segs = inflight.pop(idx)
            for rid, b, e in segs:
                fin_ranges[rid].append((b, e))
            fin = s.finished(segs)
            for rid in fin:
                assert sum(e - b for b, e in scheduled[rid]) == len(reqs[rid])
                done.add(rid)
            for rid, _, _ in segs:
                # the finished prefix: contiguous finished ranges from position 0
                expect = 0
                for b, e in sorted(fin_ranges[rid]):
                    if b != expect:
                        break
                    expect = e
                assert s.prefix(rid) == expect
            for rid in fin:
                s.retire(rid)
                retired.add(rid)
                del active_ranges[rid]
    assert admitted_order == sorted(admitted_order)  # FIFO
    for rid in reqs:
        cover = sorted(scheduled[rid])
        assert cover[0][0] == 0 and cover[-1][1] == len(reqs[rid])
        assert all(a[1] == b[0] for a, b in zip(cover, cover[1:]))
    assert s.pool_free() == pool and s.active() == 0 and s.queued() == 0
    s.close()


@pytest.mark.parametrize("seed", range(6))
def test_simulated_stream(hc, seed):
    simulate(hc, seed, n_req=60, slots=[2, 8, 256][seed % 3], pool=[1000, 4096, 1 << 20][seed % 3],
             work=[10.0, 500.0, 5e4][seed % 3], reorder=seed % 2 == 1)


def test_prefix_with_out_of_order_launches(hc):
    s = Sched(hc, 4, 100, 10.0)
    s.enqueue(1, 30)
    assert s.admit() == [1]
    s.set_costs(1, [1.0] * 30)
    a = s.next()
    b = s.next()
    c = s.next()
    assert [x[1:] for x in a + b + c] == [(0, 10), (10, 20), (20, 30)]
    assert s.finished(b) == []
    assert s.prefix(1) == 0  # [10, 20) done, [0, 10) not yet
    assert s.finished(a) == []
    assert s.prefix(1) == 20
    assert s.finished(c) == [1]
    assert s.prefix(1) == 30
    s.close()


def test_equal_share_keeps_short_requests_moving(hc):
    """A long request does not hold a short one back: both get work in the
    first launch, the short one finishes in it, and because the short one
    could finish within a quarter of the budget the launch takes half of it."""
    s = Sched(hc, 8, 10000, 100.0)
    s.enqueue(1, 5000)
    s.enqueue(2, 5)
    assert s.admit() == [1, 2]
    s.set_costs(1, [1.0] * 5000)
    s.set_costs(2, [1.0] * 5)
    segs = s.next()
    assert (2, 0, 5) in segs
    first = [x for x in segs if x[0] == 1][0]
    assert first[2] - first[1] == 45  # its share of the half budget (25) plus the rest (20)
    assert s.finished(segs) == [2]
    s.retire(2)
    # only the long request left: full launches again
    segs = s.next()
    assert segs == [(1, 45, 145)]
    s.close()


def test_rest_of_budget_to_least_work_left(hc):
    """After the equal shares, the rest of a launch's budget goes to the
    requests with the least work left, not to the oldest."""
    s = Sched(hc, 8, 100000, 120.0)
    s.enqueue(1, 10000)   # oldest, most work
    s.enqueue(2, 200)
    s.enqueue(3, 10)
    assert s.admit() == [1, 2, 3]
    for rid, n in [(1, 10000), (2, 200), (3, 10)]:
        s.set_costs(rid, [1.0] * n)
    segs = {x[0]: x[2] - x[1] for x in s.next()}
    # 3 can finish within a quarter of 120: half budget 60, shares of 20; 3
    # takes its 10 and the 10 left over go to 2 (180 left) before 1
    assert segs == {1: 20, 2: 30, 3: 10}
    # no short request left: full budget 120 in shares of 60
    s.finished([(3, 0, 10)])
    s.retire(3)
    segs = {x[0]: x[2] - x[1] for x in s.next()}
    assert segs == {1: 60, 2: 60}
    s.close()


def test_cancel(hc):
    """Cancel: a queued request leaves the queue; an admitted one keeps only
    the positions already handed to launches and completes when they finish."""
    s = Sched(hc, 4, 1000, 10.0)
    for rid, n in [(1, 100), (2, 50), (3, 10)]:
        s.enqueue(rid, n)
    assert s.admit() == [1, 2, 3]
    for rid, n in [(1, 100), (2, 50), (3, 10)]:
        s.set_costs(rid, [1.0] * n)
    s.enqueue(4, 5)
    assert s.cancel(4) == 1 and s.queued() == 0   # never admitted
    assert s.cancel(3) == 2                       # nothing handed out yet: complete now
    s.retire(3)
    a = s.next()                                  # 1 and 2 get work
    assert {x[0] for x in a} == {1, 2}
    assert s.cancel(1) == 3                       # positions of `a` are in flight
    b = s.next()
    assert {x[0] for x in b} == {2}               # 1 gets no more work
    done = s.finished(a)
    assert 1 in done
    assert s.prefix(1) == [x for x in a if x[0] == 1][0][2]
    assert s.cancel(99) == 0
    s.close()
