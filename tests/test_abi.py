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
    # hss_stats: 5 doubles, 6 u64, 2 u32, 4 u64, 2 doubles, 3 u64
    assert ctypes.sizeof(hss.Stats) == 5 * 8 + 6 * 8 + 2 * 4 + 4 * 8 + 2 * 8 + 3 * 8
    for field, _ in hss.Stats._fields_:
        assert re.search(rf"\b{field}\b", text), field


def test_struct_layouts_match_c(tmp_path):
    """sizeof and every field offset of the ABI structs, compiled from the header
    by the host C compiler, against the ctypes binding."""
    import hss
    structs = {"hss_hit": hss.Hit, "hss_root_result": hss.RootResult, "hss_stats": hss.Stats, "hss_tuning": hss.Tuning,
               "hss_server_config": hss.ServerConfig, "hss_request_info": hss.RequestInfo,
               "hss_server_stats": hss.ServerStats}
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(ln.split()[:2]): int(ln.split()[2]) for ln in out if ln}
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for field, _ in py._fields_:
            assert got[(cname, field)] == getattr(py, field).offset, (cname, field)


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


def _weight(x: int) -> int:
    """Gaussian weight by trial division (experiments/heavy/README.md:9)."""
    r, p = 1, 2
    while p * p <= x:
        e = 0
        while x % p == 0:
            x //= p
            e += 1
        if p % 4 == 1:
            r *= 4 * e + 1
        p += 1
    if x > 1 and x % 4 == 1:
        r *= 5
    return r


def test_gaussian_weight(lib):
    """hss_gaussian_weight against trial division, and the weights the
    reference quotes (evidence/benchmarks/heavy-benchmark.json: root 1116389
    has weight 125, root 1447225 weight 1125; max 5625 at 9773725)."""
    lib.hss_gaussian_weight.restype = ctypes.c_uint32
    lib.hss_gaussian_weight.argtypes = [ctypes.c_uint64]
    assert lib.hss_gaussian_weight(1116389) == 125
    assert lib.hss_gaussian_weight(1447225) == 1125
    assert lib.hss_gaussian_weight(9773725) == 5625
    assert lib.hss_gaussian_weight(0) == 0 and lib.hss_gaussian_weight((1 << 26) + 1) == 0
    import random
    rng = random.Random(7)
    xs = list(range(1, 3000)) + [rng.randrange(1, 1 << 26) for _ in range(300)] + [(1 << 26), 67108859]
    for x in xs:
        assert lib.hss_gaussian_weight(x) == _weight(x), x
