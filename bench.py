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
kernels, cell_win_kernel (tiles whose range window is narrow: only leaves that
can pass the range test are evaluated) and cell_kernel (every leaf); the line
describes the one with the longer HIP-event duration.  Both are bound by the
vector ALU (integer and FP64 lane operations), not by HBM or MFMA; achieved =
64 lane-ops per evaluated leaf (the fast path's essential arithmetic) times
the kernel's evaluated leaves per launch, divided by its duration; peak = 256
CU x 64 lanes x 2.4 GHz.
cpu_baseline: the reference search_v18 compiled from source (oracle/_ref/search),
timed on this host on a bounded sample of the same workload and scaled by the
exact leaf counts to the whole batch.
served: the seeded request stream of workloads.served_stream (40 interval
requests, LIMIT in {1e5, 1e6, 1e7}, length log-uniform in [1, 4096], Poisson
arrivals at 5 req/s) submitted in real time to the request server; latency is
submit -> results ready per request (p50 / p99 by linear interpolation), and
every served result is compared with the same request searched alone.
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
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # lane-ops per second, in T
OPS_PER_LEAF = 64  # lane-ops of one evaluated leaf's fast-path arithmetic


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
    return {"requests": n_req, "positions": positions, "rate_hz": rate_hz, "seed": workloads.SEED,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
            "throughput_positions_per_s": positions / (span_ms / 1e3),
            "leaves_per_s": st["leaves"] / (span_ms / 1e3), "span_s": span_ms / 1e3,
            "gpu_busy_frac": st["gpu_ms"] / span_ms, "launches": st["launches"],
            "shared_launches": st["shared_launches"], "max_active": st["max_active"],
            "candidates": st["candidates"], "alone_equal": equal,
            "alone_sequential_s": sum(alone_ms) / 1e3, "median_slowdown_vs_alone": float(np.median(slowdown))}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(per_root_leaves: dict[int, int], total_leaves: int, positions: int) -> dict | None:
    """Time the reference CPU search on a bounded sample of the workload."""
    ref = ROOT / "oracle" / "_ref" / "search"
    port = ROOT / "oracle" / "hss_oracle_cli"
    first, last = 7_500_000, 7_500_007
    threads = min(8, os.cpu_count() or 1)
    with tempfile.TemporaryDirectory() as work:
        env = dict(os.environ, HERON_COVERAGE_FILE=str(Path(work) / "none.json"))
        if ref.is_file():
            kind = "reference"
            cmd = [str(ref), str(LIMIT), str(first), str(last), str(threads), str(Path(work) / "s"),
                   str(Path(work) / "r")]
        elif port.is_file():
            kind = "port"
            cmd = [str(port), str(LIMIT), str(first), str(last), str(threads)]
        else:
            return None
        t0 = time.perf_counter()
        subprocess.run(cmd, cwd=work, env=env, check=True, capture_output=True)
        wall = time.perf_counter() - t0
    sample_leaves = sum(per_root_leaves[l] for l in range(first, last + 1))
    leaves_per_s = sample_leaves / wall
    value = positions * leaves_per_s / total_leaves
    return {"value": value, "unit": "positions/s", "cores": threads, "kind": kind,
            "sample": (f"l in [{first},{last}] (8 positions of the batch) at LIMIT 1e7, {threads} threads, "
                       f"{wall:.1f} s wall incl. table prepare; {leaves_per_s:.3e} leaves/s scaled to the "
                       f"batch's {total_leaves} leaves")}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="uneven1536", choices=["uneven1536", "slice512"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain", action="store_true",
                    help="also time the one-interval-per-launch form of the same roots")
    ap.add_argument("--no-served", action="store_true", help="skip the served request stream")
    ap.add_argument("--autotune", action="store_true",
                    help="run the autotuner on calibration roots (seed+1 draws, not the measured ones) first")
    ap.add_argument("--served-requests", type=int, default=40)
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
    kernel, k_ms, k_eval = ("cell_win_kernel", w_ms, w_eval) if w_ms >= p_ms else ("cell_kernel", p_ms, p_eval)
    ops = OPS_PER_LEAF * k_eval
    achieved = ops / (k_ms / 1e3) / 1e12
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
                     "kernel": kernel, "kernel_ms": k_ms, "evaluated_leaves": k_eval, "ops_per_leaf": OPS_PER_LEAF,
                     "ops_per_launch": ops},
        "work": {"leaves": leaves, "evaluated": stats["evaluated"], "candidates": cands, "tiles": stats["tiles"],
                 "window_tiles": stats["window_tiles"], "window_ms": w_ms, "window_evaluated": w_eval,
                 "plain_ms": p_ms, "plain_evaluated": p_eval, "deferred": stats["deferred"],
                 "setup_ms": stats["setup_ms"], "plan_ms": stats["plan_ms"],
                 "exact_ms": sum(exact_ms) / len(exact_ms), "leaves_per_s": leaves / cell_s},
    }
    line["tuning"] = ctx.get_tuning()
    if args.plain:
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
        line["served"] = sv
        line["latency_ms"] = {"p50": sv["p50_ms"], "p99": sv["p99_ms"]}
    if not args.no_cpu_baseline:
        log("timing the CPU baseline sample")
        line["cpu_baseline"] = cpu_baseline(per_root_leaves, sum(per_root_leaves[l] for l in roots), len(roots))
    ctx.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
