"""The two measured workloads on the GPU, position by position against the
reference's own outputs (tests/golden/reference_cases.jsonl, written by the
reference CPU search, src/search_v18.cpp:532-573, one case per position).

uneven1536: the batch the throughput figure is quoted on -- the slice
7500000..7500511, 512 heavy and 512 light roots at LIMIT 10^7 in ONE search
call, so that plain tiles, window tiles and the exact stage all run together.
Every one of the 1536 positions has its own fixture: its candidate count and
its result lines must match.

served stream: the first 200 requests of workloads.served_stream (LIMITs
10^5, 10^6, 10^7, lengths 1 .. 4096) submitted at once to one server
(continuous batching, hss_server_*).  Every LIMIT-10^5 request has a fixture
for the whole interval (total and result lines); the others have fixtures for
two (10^6) or one (10^7) seeded positions inside them.
"""
from __future__ import annotations

import pytest

from conftest import result_lines

pytestmark = pytest.mark.gpu

LIMIT = 10_000_000
STREAM_REQUESTS = 200


@pytest.fixture(scope="module")
def ctx():
    import hss
    c = hss.Context(LIMIT)
    yield c
    c.close()


def hits_at(hits, l):
    return sorted(h for h in hits if h[0] == l)


def test_uneven1536_every_root(ctx, golden):
    import workloads

    limit, roots = workloads.workload("uneven1536")
    assert limit == LIMIT and len(roots) == 1536
    want = {}
    for rec in golden.values():
        if rec["limit"] == LIMIT and rec["first"] == rec["last"]:
            want[rec["first"]] = rec
    missing = [l for l in roots if l not in want]
    assert not missing, f"{len(missing)} roots without a fixture"
    print("uneven1536: one batch of", len(roots), "positions", flush=True)
    r = ctx.search(LIMIT, roots)
    print("uneven1536:", r.total, "candidates", r.stats["cells_ms"], "ms cells", flush=True)
    assert r.stats["window_tiles"] > 0 and r.stats["tiles"] > r.stats["window_tiles"]
    bad = [l for l, c in zip(roots, r.candidates)
           if c != want[l]["candidates"] or hits_at(r.hits, l) != sorted(result_lines(want[l]["results"]))]
    assert not bad, f"{len(bad)} positions differ from the reference, first {bad[:5]}"
    # the slice's aggregate, published by the reference
    # (evidence/benchmarks/gpu-opt-benchmark.json:15-20)
    assert sum(r.candidates[:512]) == 39_091_522_776


def test_served_stream_against_fixtures(ctx, golden):
    import hss
    import workloads

    stream = workloads.served_stream(STREAM_REQUESTS)
    print("served stream:", STREAM_REQUESTS, "requests,",
          sum(last - first + 1 for _, _, first, last in stream), "positions", flush=True)
    with hss.Server(ctx) as srv:
        ids = [srv.submit(limit, first, last) for _, limit, first, last in stream]
        served = [srv.result(i) for i in ids]
        st = srv.stats()
    print("served stream: done,", st["launches"], "launches,", st["shared_launches"], "shared", flush=True)
    assert st["shared_launches"] > 0  # requests really were batched together
    whole = positions = 0
    bad = []
    for i, ((_, limit, first, last), r) in enumerate(zip(stream, served)):
        assert len(r.candidates) == last - first + 1
        prefix = f"stream{i:03d}_{limit}_"
        rec = golden.get(f"{prefix}{first}_{last}")
        if rec is not None:
            whole += 1
            if sum(r.candidates) != rec["candidates"] or sorted(r.hits) != sorted(result_lines(rec["results"])):
                bad.append(f"request {i} ({limit} {first} {last})")
        for name, rec in golden.items():
            if name.startswith(prefix + "pos_"):
                l = rec["first"]
                positions += 1
                if (r.candidates[l - first] != rec["candidates"]
                        or hits_at(r.hits, l) != sorted(result_lines(rec["results"]))):
                    bad.append(f"request {i} position {l}")
    print("served stream:", whole, "whole requests and", positions, "positions checked", flush=True)
    assert not bad, bad[:10]
    # every request is covered: LIMIT 10^5 whole, the others by sampled positions
    n_1e5 = sum(1 for _, limit, _, _ in stream if limit == 100_000)
    assert whole == n_1e5
    assert positions >= STREAM_REQUESTS - n_1e5
