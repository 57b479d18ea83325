"""Graph replay of the kernel chain on the GPU.

The launch sequence of a batch (left-set sort and build, counter reset,
window and plain cell kernels, exact and acceptance kernels) is captured once
per buffer layout and replayed for every batch, all sizes coming from a
device header.  Each replay must give exactly the direct launch's results,
replay after replay, with other direct work in between, across batches of
different sizes and a buffer growth that forces a new capture.
"""
from __future__ import annotations

import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hss_mod():
    import hss
    return hss


@pytest.fixture(scope="module")
def ctx(hss_mod):
    c = hss_mod.Context(10_000_000)
    yield c
    c.close()


def run(ctx, limit, roots, graph):
    ctx.use_graphs(graph)
    r = ctx.search(limit, roots)
    assert r.stats["graph_launch"] == (1 if graph else 0), r.stats
    return (r.candidates, r.leaves, r.pairs, r.hits)


def test_replays_equal_direct(ctx):
    import workloads
    heavy, light = workloads.uneven_roots()
    batches = [
        (5000, list(range(1, 701))),
        (70000, [11789]),
        (100_000, list(range(40_000, 40_300))),
        (10_000_000, heavy[:24] + light[:24] + list(range(7_500_000, 7_500_032))),  # window and plain tiles
        (5000, list(range(1, 5001))),                                              # grows the buffers
        (70000, [11789, 65535, 65536]),
    ]
    direct = [run(ctx, limit, roots, False) for limit, roots in batches]
    for rep in range(3):
        for i, (limit, roots) in enumerate(batches):
            got = run(ctx, limit, roots, True)
            assert got == direct[i], (rep, i)
            # other direct work between replays
            assert run(ctx, *batches[(i + 1) % len(batches)], False) == direct[(i + 1) % len(batches)]
    ctx.use_graphs(True)


def test_known_answers_through_graph(ctx, golden):
    ctx.use_graphs(True)
    tot, hits = ctx.search_interval(10_000_000, 7_500_000, 7_500_127)
    assert tot == 10_369_431_230  # evidence/benchmarks/v45-ab.json:11
    st = ctx.stats()
    assert st["graph_launch"] == 1
    # timing events sit between the replayed stages
    assert st["window_ms"] > 0 and st["plain_ms"] > 0 and st["setup_ms"] > 0
    g = golden["full5000"]
    r = ctx.search(5000, range(1, 5001))
    assert r.total == g["candidates"] and r.stats["graph_launch"] == 1
