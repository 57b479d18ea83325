"""The no-allocation entry point hss_search_noalloc, through the C ABI.

Its results must equal hss_search's on the same roots (bit-exact per-root
counters and hits) on two calls in a row, the second with a smaller batch
(the scratch array keeps the first call's contents), and a batch larger than
the scratch array must be refused with HSS_ENOMEM.
"""
from __future__ import annotations

import ctypes

import pytest

pytestmark = pytest.mark.gpu


def _noalloc(ctx, limit, roots):
    import hss
    lib = hss._lib
    u64, u32 = ctypes.c_uint64, ctypes.c_uint32
    lib.hss_search_noalloc.argtypes = [ctypes.c_void_p, u64, ctypes.POINTER(u64), u32,
                                       ctypes.POINTER(hss.RootResult), ctypes.POINTER(u64),
                                       ctypes.POINTER(hss.Hit), u32, ctypes.POINTER(u32)]
    n = len(roots)
    arr = (u64 * n)(*roots)
    per = (hss.RootResult * n)()
    total = u64()
    hits = (hss.Hit * 64)()
    nh = u32()
    rc = lib.hss_search_noalloc(ctx._h, limit, arr, n, per, ctypes.byref(total), hits, 64, ctypes.byref(nh))
    msg = lib.hss_last_error().decode(errors="replace")
    if rc != 0:
        return rc, msg, None
    r = hss.SearchResult([p.candidates for p in per], [p.leaves for p in per], [p.pairs for p in per],
                         [(h.l, h.m, h.n, h.area) for h in hits[: nh.value]], {})
    assert total.value == r.total
    return rc, msg, r


def _same(a, b):
    assert a.candidates == b.candidates
    assert a.leaves == b.leaves
    assert a.pairs == b.pairs
    assert a.hits == b.hits


def test_noalloc_equals_search_twice_then_smaller():
    import hss
    with hss.Context(100000) as ctx:
        big = list(range(1, 3001)) + [11789, 4380, 1853] + list(range(99000, 99100))
        rc, msg, r1 = _noalloc(ctx, 100000, big)
        assert rc == 0, msg
        _same(r1, ctx.search(100000, big))
        assert (1853, 4380, 4427, 32918611718880) in r1.hits
        small = [11789, 7, 70000]
        rc, msg, r2 = _noalloc(ctx, 70000, small)
        assert rc == 0, msg
        _same(r2, ctx.search(70000, small))
        assert r2.candidates[0] == 1194756 and r2.pairs[0] == 58212 and r2.leaves[0] == 1580864
        assert r2.hits == [(11789, 68104, 68595, 284239560530875680)]


def test_noalloc_refuses_batch_beyond_scratch():
    import hss
    with hss.Context(10_000_000) as ctx:
        roots = list(range(1116384, 1116384 + 4096))  # about 460 MiB of per-call buffers
        rc, msg, _ = _noalloc(ctx, 10_000_000, roots)
        assert rc == hss.HSS_ENOMEM, (rc, msg)
        assert "split the batch" in msg
        # the context still works, with both entry points
        rc, msg, r = _noalloc(ctx, 10_000_000, [9_999_999, 10_000_000])
        assert rc == 0, msg
        _same(r, ctx.search(10_000_000, [9_999_999, 10_000_000]))
