"""The served form on the GPU: continuous batching of interval requests over
one context, through the C ABI (hss_server_*).

The served result of every request must equal, bit for bit, the same interval
searched alone (hss_search, itself pinned to the reference by
test_gpu_parity.py): per-position candidates, leaves and pairs, and the hit
lines; and the reference's own fixtures where the request is one of them
(scripts/smoke_test.py:8-25).  Admission by request slots and by the device
state pool is checked through the server's statistics.
"""
from __future__ import annotations

import random
import threading
import time

import pytest

from conftest import result_lines

pytestmark = pytest.mark.gpu

NMAX = 10_000_000

# (limit, first, last): several LIMITs, lengths 1 .. 5000, overlapping roots
MIX = [
    (10_000_000, 7_500_000, 7_500_007),
    (5000, 1, 5000),
    (70000, 11789, 11789),
    (100_000, 40_000, 40_255),
    (1_000_000, 700_000, 700_031),
    (70000, 1, 300),
    (5000, 4990, 5000),
    (10_000_000, 9_773_725, 9_773_725),  # Gaussian weight 5625
]


@pytest.fixture(scope="module")
def hss_mod():
    import hss
    return hss


@pytest.fixture(scope="module")
def ctx(hss_mod):
    c = hss_mod.Context(NMAX)
    yield c
    c.close()


@pytest.fixture(scope="module")
def alone(ctx):
    """Every request of MIX searched alone, before any server is attached."""
    out = {}
    for req in MIX:
        limit, first, last = req
        r = ctx.search(limit, range(first, last + 1))
        out[req] = r
        print("alone", req, r.total, flush=True)
    return out


def served_equals_alone(res, ref):
    assert res.candidates == ref.candidates
    assert res.leaves == ref.leaves
    assert res.pairs == ref.pairs
    assert res.hits == ref.hits
    assert res.info["candidates"] == ref.total


def test_alone_vs_mix(hss_mod, ctx, alone, golden):
    with hss_mod.Server(ctx, launch_work=2e9) as srv:
        # the heavy request first; the others arrive while it runs
        ids = {MIX[0]: srv.submit(*MIX[0])}
        time.sleep(0.02)
        for req in MIX[1:]:
            ids[req] = srv.submit(*req)
            time.sleep(0.002)
        for req, rid in ids.items():
            res = srv.result(rid)
            print("served", req, res.info["candidates"], res.info["launches"], flush=True)
            assert res.info["state"] == hss_mod.REQ_DONE and res.info["resume"] == req[2] + 1
            served_equals_alone(res, alone[req])
        st = srv.stats()
        print("stats", st, flush=True)
        assert st["completed"] == len(MIX) and st["failed"] == 0
        assert st["shared_launches"] >= 1  # requests really shared launches
        assert st["max_active"] >= 2
    full = alone[(5000, 1, 5000)]
    assert full.total == golden["full5000"]["candidates"]
    assert full.hits == result_lines(golden["full5000"]["results"])
    assert alone[(70000, 11789, 11789)].total == golden["known2"]["candidates"]


def test_admission_by_state_memory(hss_mod, ctx, alone):
    """A pool of 700 positions admits one 500-position request at a time; the
    others wait in FIFO order and still give the alone results."""
    reqs = [(70000, 1, 300)] * 3 + [(5000, 4990, 5000)]
    with hss_mod.Server(ctx, state_bytes=24 * 700, launch_work=1e8) as srv:
        ids = [srv.submit(70000, 1, 500), srv.submit(70000, 1, 500), srv.submit(*reqs[0]), srv.submit(*reqs[3])]
        infos = [srv.wait(i) for i in ids]
        st = srv.stats()
        print("stats", st, flush=True)
        assert st["max_active"] <= 2  # 500 + 300 > 700: never the two 500s together
        # FIFO: a request is never admitted before an earlier one
        admits = [inf["admit_ms"] for inf in infos]
        assert admits == sorted(admits)
        assert infos[1]["admit_ms"] >= infos[0]["done_ms"]
        assert srv.result(ids[2]).candidates == alone[(70000, 1, 300)].candidates
        assert srv.result(ids[3]).candidates == alone[(5000, 4990, 5000)].candidates
        a, b = srv.result(ids[0]), srv.result(ids[1])
        assert a.candidates == b.candidates and a.hits == b.hits
        assert a.candidates[:300] == alone[(70000, 1, 300)].candidates


def test_one_slot(hss_mod, ctx, alone):
    with hss_mod.Server(ctx, max_slots=1) as srv:
        ids = {req: srv.submit(*req) for req in MIX[1:4]}
        for req, rid in ids.items():
            served_equals_alone(srv.result(rid), alone[req])
        st = srv.stats()
        assert st["max_active"] == 1 and st["shared_launches"] == 0


def test_errors(hss_mod, ctx):
    with hss_mod.Server(ctx, state_bytes=24 * 100) as srv:
        for bad in [(NMAX + 1, 1, 1), (5000, 0, 3), (5000, 10, 9), (5000, 4000, 5001), (70000, 1, 101)]:
            with pytest.raises(hss_mod.HssError) as e:
                srv.submit(*bad)
            assert e.value.code == hss_mod.HSS_EINVAL
        with pytest.raises(hss_mod.HssError):
            ctx.search(5000, [1])  # the server drives the context
        with pytest.raises(hss_mod.HssError):
            srv.wait(12345)
        rid = srv.submit(5000, 1, 10)
        assert srv.wait(rid)["state"] == hss_mod.REQ_DONE
        srv.release(rid)
        with pytest.raises(hss_mod.HssError):
            srv.info(rid)
    ctx.search(5000, [1])  # usable again once the server is gone


def test_seeded_stream_from_threads(hss_mod, ctx):
    """40 requests (LIMIT in {5000, 70000, 100000}, length log-uniform in
    [1, 256]) from two client threads; each served result equals the alone
    search done afterwards."""
    rng = random.Random(20261203)
    reqs = []
    for _ in range(40):
        limit = rng.choice([5000, 70000, 100_000])
        length = int(round(2 ** rng.uniform(0, 8)))
        first = rng.randint(1, limit - length + 1)
        reqs.append((limit, first, first + length - 1))
    results = {}
    with hss_mod.Server(ctx, launch_work=5e8) as srv:
        def client(part):
            for i in part:
                rid = srv.submit(*reqs[i])
                time.sleep(0.001 * (i % 3))
                results[i] = rid
        threads = [threading.Thread(target=client, args=(range(k, 40, 2),)) for k in range(2)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        served = {i: srv.result(rid) for i, rid in results.items()}
        st = srv.stats()
        print("stats", st, flush=True)
    assert st["completed"] == 40
    for i, req in enumerate(reqs):
        ref = ctx.search(req[0], range(req[1], req[2] + 1))
        served_equals_alone(served[i], ref)


def test_cancel_and_partial(hss_mod, ctx):
    """A request stopped while it runs ends REQ_CANCELED; its finished prefix
    [first, resume) has the alone counts; partial results while running show
    finished positions only."""
    big = (1_000_000, 1, 20_000)
    with hss_mod.Server(ctx, launch_work=2e9) as srv:
        rid = srv.submit(*big)
        small = srv.submit(70000, 11789, 11789)
        seen_partial = None
        for _ in range(400):
            info = srv.info(rid)
            if info["resume"] > big[1] + 100:
                seen_partial = srv.partial(rid)
                break
            time.sleep(0.005)
        srv.cancel(rid)
        info = srv.wait(rid)
        assert info["state"] == hss_mod.REQ_CANCELED
        assert big[1] <= info["resume"] <= big[2] + 1
        res = srv.result(small)
        assert res.total == 1_194_756
        part = srv.partial(rid)
    n_done = info["resume"] - big[1]
    if n_done:
        alone = ctx.search(big[0], range(big[1], info["resume"]))
        assert part.candidates[:n_done] == alone.candidates
        assert info["candidates"] == alone.total
    if seen_partial is not None:
        k = seen_partial.info["resume"] - big[1]
        assert seen_partial.candidates[:k] == part.candidates[:k]
