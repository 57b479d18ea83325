#!/usr/bin/env python3
"""Benchmark: verified root positions per second at LIMIT = 10^7 on one MI355X.

Workload (config.workload) "uneven1536": one batch of 1536 root positions at
LIMIT 10^7 -- the reference's interval l in [7500000, 7500511]
(evidence/benchmarks/gpu-opt-benchmark.json:15-20), 512 heavy roots (at least
four distinct primes = 1 mod 4, Gaussian weight >= 625) and 512 light prime-power
roots, both drawn from [9e6, 1e7] with seed 20261203.  A step is one search of
the whole batch: one persistent launch of the cell kernel plus the left-set
build, exact and acceptance kernels, with the number-theory tables already
resident in HBM (built once, outside the timed region).  Every step's total
candidate count is checked against the first step's, and the slice's total
against the reference's 39,091,522,776 (bit-exact or the run fails).

Prints one JSON line.  roofline: the enumeration runs as two persistent
kernels, cell_win_kernel (tiles whose range window is narrow: only leaves of
the one area class and orientation that can count, inside the range window,
are evaluated) and cell_kernel (every leaf in the one orientation its area
class allows); the line describes the one with the longer HIP-event duration.
Both are bound by the vector ALU (FP64 and integer lane operations), not by
HBM or MFMA.  achieved = (8 lane-ops per evaluated leaf orientation + 16 per
candidate) / the kernel's duration, with its own evaluated and candidate
counts: 8 = the FP64 coordinate D = Re(u w) (mul + fma), |D| - mid and four
band compares, the count; 16 = zz = l^4 + m^4 -+ 2|D|, two reciprocal square
roots, the Newton step, rint and |n^4 - zz| <= 2^-47 zz.  peak = 256 CU x 64
lanes x 2.4 GHz.  The rest of a kernel's instructions (partner and tile
bookkeeping, queueing, masked lanes) is overhead this fraction exposes.
roofline.enumerated_ops: both kernels' time charged with every leaf of the
enumeration at 64 lane-ops per leaf plus 32 per candidate (the exact-product
convention, in which skipped leaves count as done; above 1 means the skipping
beats that roofline).
cpu_baseline: the reference search_v18 compiled from source (oracle/_ref/search),
timed on this host's cores (one single-position process per core, all at
once) on a bounded, seeded sample of the batch's roots from each of its three
groups; each process's time excludes its own table prepare (the "prepared in"
line it prints), and each group's leaves per core-second scale that group's
share of the batch.
plain_form: the same roots in the reference's scheduling shape, one launch per
contiguous run of at most 512 roots (batched_over_plain = value / its rate).
dense: config "dense LIMIT 1e5" in one launch (l in [1, 100000], 141,401,679,490
candidates and both known solutions, checked), beside the reference's 1083.5 s
on 8 threads measured when the fixtures were made.
served: the seeded request stream of workloads.served_stream (200 interval
requests, LIMIT in {1e5, 1e6, 1e7}, length log-uniform in [1, 4096], Poisson
arrivals at 5 req/s) submitted in real time to the request server; latency is
submit -> results ready per request (p50 / p99 by linear interpolation); every
served result is compared with the same request searched alone, and with the
reference's fixtures for the stream (every LIMIT-1e5 request whole, seeded
positions of the others).
per-root: after the warmup step, every root's count is compared with the
reference fixture of that root where one exists (outside the timed region).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
PKG = ROOT / "heronian-square-search_amd"
sys.path.insert(0, str(PKG))

METRIC = "verified root-side positions/sec at LIMIT=1e7 (bit-exact); p50/p99 req latency"
LIMIT = 10_000_000
SLICE_CANDIDATES = 39_091_522_776  # evidence/benchmarks/gpu-opt-benchmark.json:15-20
DENSE_CANDIDATES = 141_401_679_490  # search_v18 on l in [1, 100000] at LIMIT 100000
DENSE_HITS = [(1853, 4380, 4427, 32918611718880), (11789, 68104, 68595, 284239560530875680)]
DENSE_REF_SECONDS, DENSE_REF_THREADS = 1083.5, 8  # that run's wall time on the fixture host
GOLDEN = ROOT / "tests" / "golden" / "reference_cases.jsonl"
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # lane-ops per second, in T
OPS_PER_EVAL = 8   # lane-ops of one evaluated leaf orientation (FP64 coordinate and band test)
OPS_PER_CAND = 16  # lane-ops of one candidate's fourth-root proximity test
OPS_PER_LEAF = 64  # the exact-product convention: one enumerated leaf


def profiled_traffic(kernel: str, evaluated: int) -> int | None:
    """HBM bytes per launch of `kernel` measured by the committed rocprofv3
    counter passes (profiles/cell_kernels_uneven1536_counters.md), when they
    describe this exact launch (same kernel, same evaluated-leaf count)."""
    path = ROOT / "profiles" / "cell_kernels_uneven1536_counters.md"
    if not path.is_file():
        return None
    kv = {}
    for line in path.read_text().splitlines():
        if line.startswith(f"traffic.{kernel}.") and ":" in line:
            k, v = line.split(":", 1)
            kv[k.strip().split(".")[-1]] = int(v.strip())
    if kv.get("evaluated_per_launch") != evaluated:
        return None
    return kv.get("bytes_per_launch")


def plain_form(ctx, roots: list[int]) -> dict:
    """The reference's scheduling shape: one launch per contiguous run of at
    most 512 roots (src/gpu_search_v42.cu:456-460), synchronised after each."""
    runs, i = [], 0
    while i < len(roots):
        j = i
        while j + 1 < len(roots) and roots[j + 1] == roots[j] + 1 and j + 1 - i < 512:
            j += 1
        runs.append((roots[i], roots[j]))
        i = j + 1
    t0 = time.perf_counter()
    total = 0
    for a, b in runs:
        c, _ = ctx.search_interval(LIMIT, a, b)
        total += c
    wall = time.perf_counter() - t0
    return {"value": len(roots) / wall, "unit": "positions/s", "launches": len(runs), "seconds": wall,
            "candidates": total}


def served_leg(ctx, n_req: int, rate_hz: float) -> dict:
    """Run the seeded request stream through the server in real time, then
    search each request alone and compare (bit-exact)."""
    import numpy as np
    import hss
    import workloads

    stream = workloads.served_stream(n_req, rate_hz)
    ids = []
    with hss.Server(ctx) as srv:
        t0 = time.perf_counter()
        for at, limit, first, last in stream:
            delay = t0 + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(limit, first, last))
        served = [srv.result(i) for i in ids]
        st = srv.stats()
    lat = np.array([r.info["done_ms"] - r.info["submit_ms"] for r in served])
    span_ms = max(r.info["done_ms"] for r in served) - min(r.info["submit_ms"] for r in served)
    positions = sum(last - first + 1 for _, _, first, last in stream)
    log(f"served {n_req} requests ({positions} positions) in {span_ms / 1e3:.2f} s; "
        f"p50 {np.percentile(lat, 50):.1f} ms, p99 {np.percentile(lat, 99):.1f} ms; checking each alone")
    equal, alone_ms = True, []
    for (_, limit, first, last), r in zip(stream, served):
        t1 = time.perf_counter()
        a = ctx.search(limit, range(first, last + 1))
        alone_ms.append((time.perf_counter() - t1) * 1e3)
        if a.candidates != r.candidates or a.leaves != r.leaves or a.pairs != r.pairs or a.hits != r.hits:
            equal = False
            log(f"MISMATCH served vs alone for request {limit} {first} {last}")
    slowdown = lat / np.maximum(np.array(alone_ms), 1e-3)
    nreq_fix, npos_fix, bad = check_stream_fixtures(stream, served, golden())
    for b in bad:
        log(f"MISMATCH served vs reference fixture: {b}")
    log(f"reference fixtures: {nreq_fix} whole requests, {npos_fix} positions checked, {len(bad)} mismatches")
    return {"requests": n_req, "positions": positions, "rate_hz": rate_hz, "seed": workloads.SEED,
            "fixture_requests": nreq_fix, "fixture_positions": npos_fix, "fixture_equal": not bad,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
            "throughput_positions_per_s": positions / (span_ms / 1e3),
            "leaves_per_s": st["leaves"] / (span_ms / 1e3), "span_s": span_ms / 1e3,
            "gpu_busy_frac": st["gpu_ms"] / span_ms, "launches": st["launches"],
            "shared_launches": st["shared_launches"], "max_active": st["max_active"],
            "candidates": st["candidates"], "alone_equal": equal,
            "alone_sequential_s": sum(alone_ms) / 1e3, "median_slowdown_vs_alone": float(np.median(slowdown)),
            "p99_slowdown_vs_alone": float(np.percentile(slowdown, 99))}


def golden() -> dict:
    """The reference fixtures written by tests/golden/make_golden.py (data only)."""
    out = {}
    if GOLDEN.is_file():
        for line in GOLDEN.read_text().splitlines():
            if line.strip():
                rec = json.loads(line)
                out[rec["name"]] = rec
    return out


def result_lines(text: str) -> list[tuple[int, int, int, int]]:
    return [tuple(int(x) for x in ln.split()) for ln in text.splitlines() if ln.strip()]


def check_stream_fixtures(stream, served, gold) -> tuple[int, int, list[str]]:
    """Compare served requests with the reference's stream fixtures: a whole
    request (name streamNNN_LIMIT_FIRST_LAST) by its candidate total and result
    lines, a sampled position (streamNNN_LIMIT_pos_L) by that position's count
    and hits.  Returns (requests checked, positions checked, mismatches)."""
    nreq = npos = 0
    bad = []
    for i, ((_, limit, first, last), r) in enumerate(zip(stream, served)):
        prefix = f"stream{i:03d}_{limit}_"
        whole = gold.get(f"{prefix}{first}_{last}")
        if whole:
            nreq += 1
            npos += last - first + 1
            if sum(r.candidates) != whole["candidates"] or sorted(r.hits) != sorted(result_lines(whole["results"])):
                bad.append(f"request {i} ({limit} {first} {last})")
        for name, rec in gold.items():
            if not name.startswith(prefix + "pos_"):
                continue
            l = rec["first"]
            npos += 1
            hits_l = [h for h in r.hits if h[0] == l]
            if r.candidates[l - first] != rec["candidates"] or sorted(hits_l) != sorted(result_lines(rec["results"])):
                bad.append(f"request {i} position {l}")
    return nreq, npos, bad


def check_roots(roots, r, gold, limit) -> tuple[int, list[int]]:
    """Per-root counts of a search against the single-position fixtures."""
    want = {}
    for rec in gold.values():
        if rec["limit"] == limit and rec["first"] == rec["last"]:
            want[rec["first"]] = rec
    checked, bad = 0, []
    for l, c in zip(roots, r.candidates):
        rec = want.get(l)
        if rec is None:
            continue
        checked += 1
        if c != rec["candidates"] or sorted(h for h in r.hits if h[0] == l) != sorted(result_lines(rec["results"])):
            bad.append(l)
    return checked, bad


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_share() -> int:
    """Host cores this process may use (the box gives one GPU's share of 16)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(roots: list[int], per_root_leaves: dict[int, int], positions: int) -> dict | None:
    """Time the reference CPU search on a bounded sample of the workload.

    One single-position process per core, all started together; a process's
    time is its wall time minus the table prepare it reports on stderr.  The
    sample draws roots from each group of the batch (the slice, the heavy and
    the light roots) among those of at most 1.2e8 leaves (about 30 s on one
    core), and each group's measured leaves per core-second is applied to
    that group's share of the batch's leaves."""
    import random
    ref = ROOT / "oracle" / "_ref" / "search"
    port = ROOT / "oracle" / "hss_oracle_cli"
    if ref.is_file():
        kind, binary = "reference", ref
    elif port.is_file():
        kind, binary = "port", port
    else:
        return None
    cores = cpu_share()
    slice_set = set(range(7_500_000, 7_500_512))
    heavy_set = set(roots[512:1024])
    groups = {"slice": [l for l in roots if l in slice_set], "heavy": [l for l in roots if l in heavy_set],
              "light": [l for l in roots if l not in slice_set and l not in heavy_set]}
    rng = random.Random(20261203)
    per_group = max(1, cores // 3)
    sample = {}
    for g, ls in groups.items():
        pool = [l for l in ls if per_root_leaves[l] <= 1.2e8]
        sample[g] = sorted(rng.sample(pool, min(per_group, len(pool))))
    jobs = [(g, l) for g, ls in sample.items() for l in ls]
    with tempfile.TemporaryDirectory() as work:
        env = dict(os.environ, HERON_COVERAGE_FILE=str(Path(work) / "none.json"))
        procs = []
        for g, l in jobs:
            d = Path(work) / str(l)
            d.mkdir()
            if kind == "reference":
                cmd = [str(binary), str(LIMIT), str(l), str(l), "1", str(d / "s"), str(d / "r")]
            else:
                cmd = [str(binary), str(LIMIT), str(l), str(l), "1"]
            procs.append((g, l, time.perf_counter(), subprocess.Popen(cmd, cwd=d, env=env, stdout=subprocess.PIPE,
                                                                      stderr=subprocess.PIPE, text=True)))
        busy = {g: 0.0 for g in groups}
        leaves = {g: 0 for g in groups}
        prep = []
        for g, l, t0, pr in procs:
            _, err = pr.communicate()
            wall = time.perf_counter() - t0
            if pr.returncode:
                raise RuntimeError(f"cpu baseline root {l}: exit {pr.returncode}")
            p = 0.0
            for line in err.splitlines():
                if line.startswith("prepared in"):
                    p = float(line.split()[2])
            prep.append(p)
            busy[g] += max(wall - p, 1e-3)
            leaves[g] += per_root_leaves[l]
    batch_leaves = {g: sum(per_root_leaves[l] for l in ls) for g, ls in groups.items()}
    core_seconds = sum(batch_leaves[g] / (leaves[g] / busy[g]) for g in groups if batch_leaves[g] and leaves[g])
    value = positions * cores / core_seconds
    return {"value": value, "unit": "positions/s", "cores": cores, "kind": kind,
            "sample": (f"{len(jobs)} single-position processes at LIMIT 1e7, all at once on {cores} cores: "
                       + "; ".join(f"{g} {sample[g]} {leaves[g] / busy[g]:.3e} leaves/core-s" for g in groups)
                       + f"; table prepare ({min(prep):.1f}-{max(prep):.1f} s per process) excluded; reference "
                       "timer granularity 1 s; the batch's leaves per group scaled by each group's rate"),
            "batch_core_seconds": core_seconds}


def dense_leg(ctx) -> dict:
    """The dense LIMIT 1e5 configuration in one launch, checked."""
    roots = list(range(1, 100_001))
    t0 = time.perf_counter()
    r = ctx.search(100_000, roots)
    wall = time.perf_counter() - t0
    ok = r.total == DENSE_CANDIDATES and r.hits == DENSE_HITS
    return {"positions": len(roots), "seconds": wall, "positions_per_s": len(roots) / wall,
            "candidates": r.total, "hits": len(r.hits), "bit_exact": ok, "leaves": sum(r.leaves),
            "cells_ms": r.stats["cells_ms"],
            "reference_cpu": {"seconds": DENSE_REF_SECONDS, "threads": DENSE_REF_THREADS,
                              "speedup": DENSE_REF_SECONDS / wall}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="uneven1536", choices=["uneven1536", "slice512"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the one-interval-per-launch form of the same roots")
    ap.add_argument("--no-dense", action="store_true", help="skip the dense LIMIT 1e5 configuration")
    ap.add_argument("--no-served", action="store_true", help="skip the served request stream")
    ap.add_argument("--autotune", action="store_true",
                    help="run the autotuner on calibration roots (seed+1 draws, not the measured ones) first")
    ap.add_argument("--served-requests", type=int, default=200)
    ap.add_argument("--served-rate", type=float, default=5.0, help="Poisson arrival rate, requests/s")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(root positions shard across GPUs with no data exchange)", file=sys.stderr)
        return 2

    import hss
    import workloads

    limit, roots = workloads.workload(args.workload)
    assert limit == LIMIT
    t0 = time.perf_counter()
    ctx = hss.Context(LIMIT, hss.visible_device())
    log(f"tables for N_max={LIMIT} built in {time.perf_counter() - t0:.2f} s "
        f"({ctx.device_bytes / 2**30:.2f} GiB resident)")

    if args.autotune:
        heavy, light = workloads.uneven_roots(seed=workloads.SEED + 1)
        calib = heavy[::8] + light[::16] + list(range(7_400_000, 7_400_064))
        t0 = time.perf_counter()
        tuned, ms = ctx.autotune(LIMIT, calib, repeats=2)
        log(f"autotune on {len(calib)} calibration roots: {tuned} ({ms:.1f} ms, {time.perf_counter() - t0:.1f} s)")

    def step():
        r = ctx.search(LIMIT, roots)
        return r

    ref_total = None
    for i in range(args.warmup):
        r = step()
        ref_total = r.total
        log(f"warmup {i}: {r.stats['total_ms']:.1f} ms, candidates {r.total}, cells {r.stats['cells_ms']:.1f} ms")
    per_root_leaves = dict(zip(roots, r.leaves)) if args.warmup else None
    slice_total = sum(c for l, c in zip(roots, r.candidates) if 7_500_000 <= l <= 7_500_511) if args.warmup else None
    if slice_total is not None and slice_total != SLICE_CANDIDATES:
        log(f"PARITY FAILURE: slice candidates {slice_total} != {SLICE_CANDIDATES}")
        return 1
    per_root_checked = None
    if args.warmup:
        per_root_checked, bad = check_roots(roots, r, golden(), LIMIT)
        log(f"per-root reference fixtures: {per_root_checked} of {len(roots)} roots checked, {len(bad)} mismatches")
        if bad:
            log(f"PARITY FAILURE: per-root counts differ from the reference at {bad[:8]}")
            return 1
    cells_ms, exact_ms, win_ms, plain_ms, stats = [], [], [], [], None
    t_start = time.perf_counter()
    for i in range(args.steps):
        r = step()
        if ref_total is None:
            ref_total = r.total
        if r.total != ref_total:
            log(f"PARITY FAILURE: step {i} candidates {r.total} != {ref_total}")
            return 1
        cells_ms.append(r.stats["cells_ms"])
        exact_ms.append(r.stats["exact_ms"])
        win_ms.append(r.stats["window_ms"])
        plain_ms.append(r.stats["plain_ms"])
        stats = r.stats
        log(f"step {i}: {r.stats['total_ms']:.1f} ms (cells {r.stats['cells_ms']:.1f} ms)")
    elapsed = time.perf_counter() - t_start
    if per_root_leaves is None:
        per_root_leaves = dict(zip(roots, r.leaves))
    positions = len(roots) * args.steps
    value = positions / elapsed
    leaves, cands = stats["leaves"], stats["candidates"]
    cell_s = sum(cells_ms) / len(cells_ms) / 1e3
    w_ms, p_ms = sum(win_ms) / len(win_ms), sum(plain_ms) / len(plain_ms)
    w_eval = stats["window_evaluated"]
    p_eval = stats["evaluated"] - w_eval
    w_cand, p_cand = stats["window_candidates"], stats["plain_candidates"]
    kernel, k_ms, k_eval, k_cand = (("cell_win_kernel", w_ms, w_eval, w_cand) if w_ms >= p_ms
                                    else ("cell_kernel", p_ms, p_eval, p_cand))
    ops = OPS_PER_EVAL * k_eval + OPS_PER_CAND * k_cand
    achieved = ops / (k_ms / 1e3) / 1e12
    other_ops = OPS_PER_EVAL * (w_eval + p_eval - k_eval) + OPS_PER_CAND * (w_cand + p_cand - k_cand)
    other_ms = p_ms if kernel == "cell_win_kernel" else w_ms
    # the enumerated-leaf convention: every leaf of the whole enumeration, and
    # every candidate, charged over both cell kernels' time
    enum_ops = OPS_PER_LEAF * leaves + 32 * cands
    enum_achieved = enum_ops / cell_s / 1e12
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "positions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+i32 (exact i128 settle)",
        "data": "synthetic (seeded root positions; no dataset)",
        "config": {"workload": f"{args.workload}: {len(roots)} root positions at LIMIT=1e7 "
                               "(slice 7500000-7500511 + 512 heavy + 512 light roots, seed 20261203)",
                   "limit": LIMIT, "positions": len(roots), "parallelism": "1 GPU, persistent grid"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TOPS, "unit": "Tops/s",
                     "frac": achieved / VALU_PEAK_TOPS, "traffic": profiled_traffic(kernel, k_eval),
                     "kernel": kernel, "kernel_ms": k_ms, "evaluated": k_eval, "candidates": k_cand,
                     "ops_per_eval": OPS_PER_EVAL, "ops_per_candidate": OPS_PER_CAND, "ops_per_launch": ops,
                     "other_kernel_frac": other_ops / (other_ms / 1e3) / 1e12 / VALU_PEAK_TOPS,
                     "enumerated_ops": enum_ops, "enumerated_achieved": enum_achieved,
                     "enumerated_frac": enum_achieved / VALU_PEAK_TOPS, "enumerated_over_s": cell_s},
        "work": {"leaves": leaves, "evaluated": stats["evaluated"], "candidates": cands, "tiles": stats["tiles"],
                 "window_tiles": stats["window_tiles"], "window_ms": w_ms, "window_evaluated": w_eval,
                 "plain_ms": p_ms, "plain_evaluated": p_eval, "window_candidates": w_cand,
                 "plain_candidates": p_cand, "deferred": stats["deferred"],
                 "setup_ms": stats["setup_ms"], "plan_ms": stats["plan_ms"],
                 "exact_ms": sum(exact_ms) / len(exact_ms), "leaves_per_s": leaves / cell_s},
        "per_root_fixtures_checked": per_root_checked,
    }
    line["tuning"] = ctx.get_tuning()
    if not args.no_plain:
        log("timing the plain one-interval-per-launch form")
        pf = plain_form(ctx, sorted(roots))
        if pf["candidates"] != ref_total:
            log(f"PARITY FAILURE: plain form candidates {pf['candidates']} != {ref_total}")
            return 1
        pf["batched_over_plain"] = value / pf["value"]
        line["plain_form"] = pf
    if not args.no_served:
        log(f"served stream: {args.served_requests} requests at {args.served_rate} req/s")
        sv = served_leg(ctx, args.served_requests, args.served_rate)
        if not sv["alone_equal"]:
            log("PARITY FAILURE: served results differ from the alone searches")
            return 1
        if not sv["fixture_equal"]:
            log("PARITY FAILURE: served results differ from the reference fixtures")
            return 1
        line["served"] = sv
        line["latency_ms"] = {"p50": sv["p50_ms"], "p99": sv["p99_ms"]}
    if not args.no_dense:
        log("dense LIMIT 1e5, l in [1, 100000], one launch")
        dn = dense_leg(ctx)
        if not dn["bit_exact"]:
            log(f"PARITY FAILURE: dense LIMIT 1e5 candidates {dn['candidates']} != {DENSE_CANDIDATES}")
            return 1
        line["dense"] = dn
    if not args.no_cpu_baseline:
        log("timing the CPU baseline sample")
        line["cpu_baseline"] = cpu_baseline(roots, per_root_leaves, len(roots))
    ctx.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
