#!/usr/bin/env python3
"""Write golden fixtures by running the reference CPU search built from source.

The reference binary is ``oracle/_ref/search`` (``make -C oracle ref`` compiles
src/search_v18.cpp where it lies).  Each case runs it as a subprocess with the
reference's own argv shape (``search LIMIT FIRST LAST THREADS STATE RESULTS``,
src/search_v18.cpp:532-573) and an isolated coverage file, exactly as
scripts/smoke_test.py:33-61 does, and records the STATE fields and the results
text.  Output: ``reference_cases.jsonl`` beside this script, one JSON object per
case; cases already present are skipped, so the script can be re-run to extend.

Usage: python tests/golden/make_golden.py [--jobs 6] [--only PREFIX]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT / "heronian-square-search_amd"))
import numpy as np  # noqa: E402
import workloads  # noqa: E402

OUT = HERE / "reference_cases.jsonl"
BINARY = ROOT / "oracle" / "_ref" / "search"


def cases():
    rng = np.random.default_rng(workloads.SEED)
    out = []

    def add(name, limit, first, last, threads=1):
        out.append(dict(name=name, limit=limit, first=first, last=last, threads=threads))

    # regression cases of scripts/smoke_test.py:8-25
    add("full5000", 5000, 1, 5000, threads=8)
    add("known2", 70000, 11789, 11789)
    # per-position samples at small limits (edges included)
    l5k = sorted({1, 2, 3, 4, 5, 6, 25, 65, 1105, 1853, 4380, 4999, 5000}
                 | {int(x) for x in rng.integers(1, 5001, 40)})
    for l in l5k:
        add(f"l5000_{l}", 5000, l, l)
    l70k = sorted({11789, 1, 2, 65535, 65536, 69999, 70000}
                  | {int(x) for x in rng.integers(1, 70001, 16)})
    for l in l70k:
        add(f"l70000_{l}", 70000, l, l)
    # short intervals at LIMIT 10^5
    for a, b in [(1, 100), (1000, 1099), (4380, 4479), (50000, 50099), (99901, 100000)]:
        add(f"i100000_{a}_{b}", 100000, a, b, threads=4)
    # per-position samples at LIMIT 10^6
    for l in sorted({999999, 1000000} | {int(x) for x in rng.integers(500000, 1000001, 6)}):
        add(f"l1000000_{l}", 1000000, l, l)
    # LIMIT 10^7: light roots of the uneven workload, slice positions, two heavy roots
    heavy, light = workloads.uneven_roots()
    for l in light[::32]:
        add(f"l1e7_light_{l}", 10_000_000, l, l)
    for l in range(7_500_000, 7_500_004):
        add(f"l1e7_slice_{l}", 10_000_000, l, l)
    for l in (9_999_999, 10_000_000):
        add(f"l1e7_top_{l}", 10_000_000, l, l)
    for l in heavy[:2]:
        add(f"l1e7_heavy_{l}", 10_000_000, l, l)
    # second wave: more slice positions, heavy and light roots
    for l in range(7_500_004, 7_500_032):
        add(f"l1e7_slice_{l}", 10_000_000, l, l)
    for l in heavy[2:512:64]:
        add(f"l1e7_heavy_{l}", 10_000_000, l, l)
    for l in light[16::32]:
        add(f"l1e7_light_{l}", 10_000_000, l, l)
    # third wave: every root of the uneven1536 batch, one position per case
    for l in range(7_500_000, 7_500_512):
        add(f"l1e7_slice_{l}", 10_000_000, l, l)
    for l in heavy:
        add(f"l1e7_heavy_{l}", 10_000_000, l, l)
    for l in light:
        add(f"l1e7_light_{l}", 10_000_000, l, l)
    # the served stream: every LIMIT-10^5 request whole, seeded positions of the others
    for name, limit, first, last in stream_cases(STREAM_REQUESTS):
        add(name, limit, first, last)
    return out


STREAM_REQUESTS = 200


def stream_cases(n):
    """Fixture cases for the first n requests of workloads.served_stream.

    A LIMIT-10^5 request is recorded whole (one case over its interval).  A
    LIMIT-10^6 request contributes two positions and a LIMIT-10^7 request one,
    drawn with a generator seeded from workloads.SEED, so the sample does not
    depend on which requests a run serves.  Names carry the request index.
    """
    rng = np.random.default_rng(workloads.SEED + 2)
    out = []
    for i, (_, limit, first, last) in enumerate(workloads.served_stream(n)):
        picks = rng.integers(first, last + 1, 2)
        if limit == 100_000:
            out.append((f"stream{i:03d}_{limit}_{first}_{last}", limit, first, last))
        elif limit == 1_000_000:
            for l in sorted({int(x) for x in picks}):
                out.append((f"stream{i:03d}_{limit}_pos_{l}", limit, l, l))
        else:
            l = int(picks[0])
            out.append((f"stream{i:03d}_{limit}_pos_{l}", limit, l, l))
    return out


def run(case):
    with tempfile.TemporaryDirectory(prefix="golden-") as work:
        work = Path(work)
        state = work / "state.txt"
        results = work / "results.txt"
        env = os.environ.copy()
        env["HERON_COVERAGE_FILE"] = str(work / "no-coverage.json")
        cmd = [str(BINARY), str(case["limit"]), str(case["first"]), str(case["last"]),
               str(case["threads"]), str(state), str(results)]
        t0 = time.time()
        done = subprocess.run(cmd, cwd=work, env=env, text=True, capture_output=True)
        seconds = time.time() - t0
        if done.returncode:
            raise RuntimeError(f"{case['name']}: exit {done.returncode}: {done.stderr}")
        st = {k: int(v) for k, v in (ln.split() for ln in state.read_text().splitlines())}
        rec = dict(case)
        rec.update(pairs=st["pairs"], candidates=st["candidates"], resume=st["resume"],
                   begin=st["begin"], end=st["end"],
                   results=results.read_text() if results.exists() else "",
                   wall_seconds=round(seconds, 2))
        return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=6)
    ap.add_argument("--only", default="")
    args = ap.parse_args()
    if not BINARY.is_file():
        raise SystemExit(f"missing {BINARY}; run make -C oracle ref")
    have = set()
    if OUT.exists():
        have = {json.loads(ln)["name"] for ln in OUT.read_text().splitlines() if ln.strip()}
    todo = []
    for c in cases():
        if c["name"] not in have and c["name"].startswith(args.only):
            have.add(c["name"])
            todo.append(c)
    # cheap cases first so partial runs are still useful
    todo.sort(key=lambda c: (c["limit"], c["last"] - c["first"]))
    print(f"{len(todo)} cases to run", flush=True)
    with cf.ThreadPoolExecutor(args.jobs) as pool, OUT.open("a") as sink:
        futs = {pool.submit(run, c): c for c in todo}
        for fut in cf.as_completed(futs):
            rec = fut.result()
            sink.write(json.dumps(rec, sort_keys=True) + "\n")
            sink.flush()
            print(f"{rec['name']}: candidates {rec['candidates']} ({rec['wall_seconds']} s)", flush=True)


if __name__ == "__main__":
    main()
