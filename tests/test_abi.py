"""C ABI checks that need no GPU: libhss.so loads, exports every function that
include/hss.h declares, the binding's struct layouts agree with the header, and
the host-side exact triple check follows scripts/verify.py."""
from __future__ import annotations

import ctypes
import re
import subprocess

import pytest

from conftest import PKG, ROOT

HEADER = ROOT / "include" / "hss.h"
LIB = PKG / "libhss.so"


@pytest.fixture(scope="module")
def lib():
    if not LIB.is_file():
        subprocess.run(["make", "-C", str(PKG)], check=True, capture_output=True)
    return ctypes.CDLL(str(LIB))


def declared_functions():
    text = HEADER.read_text()
    body = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hss_[a-z_0-9]+)\s*\(", body)))


def test_exports_every_declared_function(lib):
    names = declared_functions()
    assert "hss_search" in names and "hss_ctx_create" in names
    for name in names:
        assert hasattr(lib, name), name


def test_struct_layouts():
    import hss
    text = HEADER.read_text()
    assert ctypes.sizeof(hss.Hit) == 40
    assert ctypes.sizeof(hss.RootResult) == 24
    # hss_stats: 5 doubles, 6 u64, 2 u32
    assert ctypes.sizeof(hss.Stats) == 5 * 8 + 6 * 8 + 2 * 4
    for field, _ in hss.Stats._fields_:
        assert re.search(rf"\b{field}\b", text), field


def test_verify_triple_known_answers():
    import hss
    assert hss.verify_triple(1853, 4380, 4427, 32918611718880)
    assert hss.verify_triple(11789, 68104, 68595, 284239560530875680)
    assert not hss.verify_triple(1853, 4380, 4427, 32918611718881)
    assert not hss.verify_triple(3706, 8760, 8854, 32918611718880 * 16)  # not primitive
    assert not hss.verify_triple(1, 1, 2, 0)


def test_verify_triple_matches_results_files():
    import hss
    evid = ROOT / "tests" / "golden" / "reference_cases.jsonl"
    import json
    lines = set()
    for ln in evid.read_text().splitlines():
        for r in json.loads(ln)["results"].splitlines():
            if r.strip():
                lines.add(tuple(int(x) for x in r.split()))
    assert lines
    for l, m, n, a in lines:
        assert hss.verify_triple(l, m, n, a)
