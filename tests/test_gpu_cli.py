"""Drop-in contract of the executables on the GPU.

Follows scripts/smoke_test.py:8-61 (cases, argv shapes, STATE fields, exact
results text), the checkpoint chain rule of scripts/coverage_audit.py:48-68 and
scripts/gpu_queue.py:40-63 (PROGRESS resume steps of exactly 512, or the final
short batch, from FIRST to LAST+1), resume from STATE, SIGTERM between
batches, de-duplication against an existing results file, and the `search`
front end's pairs counter and coverage skip (src/search_v18.cpp:254-283).
"""
from __future__ import annotations

import json
import os
import signal
import subprocess
import time

import pytest

from conftest import PKG, result_lines

pytestmark = pytest.mark.gpu

GPU = PKG / "gpu_search"
SERVE = PKG / "gpu_serve"
CPU = PKG / "search"
CASES = {
    "full5000": dict(limit=5000, first=1, last=5000, pairs=12502500, candidates=151842556,
                     results="1853 4380 4427 32918611718880\n"),
    "known2": dict(limit=70000, first=11789, last=11789, pairs=58212, candidates=1194756,
                   results="11789 68104 68595 284239560530875680\n"),
}


def read_state(path):
    return {k: int(v) for k, v in (ln.split() for ln in path.read_text().splitlines())}


def chain_end(log_text: str, first: int, last: int):
    """resume reached by an unbroken +512 PROGRESS chain, or None if broken."""
    expected = first
    for line in log_text.splitlines():
        f = line.split()
        if len(f) < 3 or f[:2] != ["PROGRESS", "resume"]:
            continue
        cur = int(f[2])
        if cur <= expected:
            continue
        if cur != min(expected + 512, last + 1):
            return None
        expected = cur
    return expected


def run(cmd, cwd, **kw):
    env = dict(os.environ, HERON_COVERAGE_FILE=str(cwd / "no-coverage.json"))
    return subprocess.run([str(c) for c in cmd], cwd=cwd, env=env, capture_output=True, text=True, timeout=600,
                          **kw)


@pytest.mark.parametrize("name", list(CASES))
def test_gpu_search_smoke_cases(tmp_path, name):
    c = CASES[name]
    state, results = tmp_path / "gpu-0x-state.txt", tmp_path / "gpu-0x-results.txt"
    r = run([GPU, c["limit"], c["first"], c["last"], results, 0, state], tmp_path)
    assert r.returncode == 0, r.stderr
    st = read_state(state)
    assert st == {"version": 1, "limit": c["limit"], "begin": c["first"], "end": c["last"],
                  "resume": c["last"] + 1, "candidates": c["candidates"]}
    assert results.read_text() == c["results"]
    assert chain_end(r.stdout, c["first"], c["last"]) == c["last"] + 1


@pytest.mark.parametrize("name", list(CASES))
def test_search_cli_smoke_cases(tmp_path, name):
    c = CASES[name]
    state, results = tmp_path / "state.txt", tmp_path / "results.txt"
    r = run([CPU, c["limit"], c["first"], c["last"], 2, state, results], tmp_path)
    assert r.returncode == 0, r.stderr
    st = read_state(state)
    assert st["limit"] == c["limit"] and st["begin"] == c["first"] and st["end"] == c["last"]
    assert st["resume"] == c["last"] + 1
    assert st["pairs"] == c["pairs"] and st["candidates"] == c["candidates"]
    assert results.read_text() == c["results"]
    assert "HIT " + c["results"].split()[0] in r.stdout


def test_resume_from_checkpoint(tmp_path):
    limit, first, last = 100000, 1, 2000
    full = tmp_path / "full"
    full.mkdir()
    r = run([GPU, limit, first, last, full / "r.txt", 0, full / "s.txt"], full)
    assert r.returncode == 0, r.stderr
    total = read_state(full / "s.txt")["candidates"]
    head = tmp_path / "head"
    head.mkdir()
    r = run([GPU, limit, 1, 1024, head / "r.txt", 0, head / "s.txt"], head)
    part1 = read_state(head / "s.txt")["candidates"]
    # a checkpoint as a killed worker leaves it after two batches
    resumed = tmp_path / "resumed"
    resumed.mkdir()
    (resumed / "s.txt").write_text(f"version 1\nlimit {limit}\nbegin 1\nend {last}\nresume 1025\ncandidates 7\n")
    r = run([GPU, limit, first, last, resumed / "r.txt", 0, resumed / "s.txt"], resumed)
    assert r.returncode == 0, r.stderr
    st = read_state(resumed / "s.txt")
    assert st["begin"] == 1025 and st["resume"] == last + 1
    assert part1 + st["candidates"] == total
    assert chain_end(r.stdout, 1025, last) == last + 1


def test_checkpoint_mismatch_exits_zero(tmp_path):
    (tmp_path / "s.txt").write_text("version 1\nlimit 5000\nbegin 1\nend 10\nresume 3\ncandidates 0\n")
    r = run([GPU, 5000, 1, 20, tmp_path / "r.txt", 0, tmp_path / "s.txt"], tmp_path)
    assert r.returncode == 0 and "checkpoint does not match requested range" in r.stderr
    assert read_state(tmp_path / "s.txt")["end"] == 10


def test_sigterm_between_batches_then_resume(tmp_path):
    limit, first, last = 1_000_000, 1, 12_288  # 24 batches
    state, results, log = tmp_path / "gpu-0x-state.txt", tmp_path / "gpu-0x-results.txt", tmp_path / "gpu-0x.log"
    env = dict(os.environ, HERON_COVERAGE_FILE=str(tmp_path / "none.json"))
    with log.open("w") as out:
        p = subprocess.Popen([str(GPU), str(limit), str(first), str(last), str(results), "0", str(state)],
                             cwd=tmp_path, env=env, stdout=out, stderr=subprocess.STDOUT)
        deadline = time.time() + 120
        while time.time() < deadline and "PROGRESS" not in log.read_text():
            time.sleep(0.05)
        p.send_signal(signal.SIGTERM)
        assert p.wait(timeout=120) == 0
    stopped = read_state(state)["resume"]
    assert (stopped - first) % 512 == 0 or stopped == last + 1
    with log.open("a") as out:  # the supervisor appends the relaunch to the same log
        r = subprocess.run([str(GPU), str(limit), str(first), str(last), str(results), "0", str(state)],
                           cwd=tmp_path, env=env, stdout=out, stderr=subprocess.STDOUT, timeout=600)
    assert r.returncode == 0
    assert read_state(state)["resume"] == last + 1
    assert chain_end(log.read_text(), first, last) == last + 1


def test_results_dedup(tmp_path):
    c = CASES["known2"]
    results = tmp_path / "r.txt"
    results.write_text(c["results"])
    r = run([GPU, c["limit"], c["first"], c["last"], results, 0, tmp_path / "s.txt"], tmp_path)
    assert r.returncode == 0
    assert results.read_text() == c["results"]


def test_search_coverage_skip(tmp_path):
    cov = tmp_path / "cov.json"
    cov.write_text(json.dumps({"limit": 5000, "complete_intervals": [[1, 4000]], "root_coverage": {}}))
    env = dict(os.environ, HERON_COVERAGE_FILE=str(cov))
    r = subprocess.run([str(CPU), "5000", "3990", "4010", "1", str(tmp_path / "s"), str(tmp_path / "r")],
                       cwd=tmp_path, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    st = read_state(tmp_path / "s")
    ref = subprocess.run([str(CPU), "5000", "4001", "4010", "1", str(tmp_path / "s2"), str(tmp_path / "r2")],
                         cwd=tmp_path, env=dict(os.environ, HERON_COVERAGE_FILE=str(tmp_path / "none")),
                         capture_output=True, text=True, timeout=600)
    assert ref.returncode == 0
    # covered roots count as done (pairs) but contribute no candidates
    assert st["candidates"] == read_state(tmp_path / "s2")["candidates"]
    assert st["pairs"] == sum(5000 - l + 1 for l in range(3990, 4011))
    assert result_lines((tmp_path / "r").read_text()) == []


def test_search_coverage_refresh_mid_run(tmp_path):
    """The coverage report is read again while the run goes on
    (src/search_v18.cpp:213-222, 254-283): a report that appears mid-run stops
    the search of the intervals it covers.  The run is paused (SIGSTOP) after
    its first batch, the report written, and the run continued; with the
    refresh period 0 the next batch sees it."""
    import hss
    limit, first, last = 1_000_000, 1, 12_000
    cov, state, results = tmp_path / "cov.json", tmp_path / "s.txt", tmp_path / "r.txt"
    env = dict(os.environ, HERON_COVERAGE_FILE=str(cov), HSS_SEARCH_REFRESH_S="0", HSS_SEARCH_PROGRESS_S="0")
    p = subprocess.Popen([str(CPU), str(limit), str(first), str(last), "1", str(state), str(results)], cwd=tmp_path,
                         env=env, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
    try:
        line = p.stdout.readline()
        assert line.startswith("PROGRESS"), line
        p.send_signal(signal.SIGSTOP)
        resume = read_state(state)["resume"]  # the batch in flight starts here
        cut = resume + 2048
        assert cut < last
        cov.write_text(json.dumps({"limit": limit, "complete_intervals": [[cut, last]], "root_coverage": {}}))
        p.send_signal(signal.SIGCONT)
        out, _ = p.communicate(timeout=600)
    finally:
        if p.poll() is None:
            p.kill()
    assert p.returncode == 0
    st = read_state(state)
    assert st["resume"] == last + 1 and st["pairs"] == sum(limit - l + 1 for l in range(first, last + 1))
    with hss.Context(limit) as ctx:
        assert st["candidates"] == ctx.search(limit, range(first, cut)).total
    # PROGRESS after every batch with period 0, ending at last + 1
    resumes = [int(ln.split()[2]) for ln in (line + out).splitlines() if ln.startswith("PROGRESS")]
    assert resumes[-1] == last + 1 and all(a <= b for a, b in zip(resumes, resumes[1:]))


def test_search_progress_cadence(tmp_path):
    """With the default 10-second period a short run prints one PROGRESS line,
    at its end (src/search_v18.cpp:213-226)."""
    r = run([CPU, 100000, 1, 3000, 1, tmp_path / "s.txt", tmp_path / "r.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("PROGRESS")]
    assert len(lines) == 1 and lines[0].split()[:5] == ["PROGRESS", "resume", "3001", "completed_l", "3000"]


def weight(x: int) -> int:
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


def check_done(state_path, first, resume, limit, min_weight):
    """The checks experiments/heavy/verify_heavy.py:44-96 and
    scripts/coverage_audit.py:71-99 make on a heavy checkpoint's .done file."""
    lines = (state_path.parent / (state_path.name + ".done")).read_text().splitlines()
    header, roots = {}, []
    for ln in lines:
        f = ln.split()
        if len(f) == 2 and f[0] in {"limit", "min_weight"}:
            header[f[0]] = int(f[1])
        else:
            assert len(f) == 1, ln
            roots.append(int(f[0]))
    assert header == {"limit": limit, "min_weight": min_weight}
    assert all(a < b for a, b in zip(roots, roots[1:]))
    assert roots == [l for l in range(first, resume) if weight(l) >= min_weight]
    return roots


def test_heavy_weight_mode(tmp_path):
    """MIN_WEIGHT (experiments/heavy/gpu_search_v36.cu): batches of 4096,
    min_weight in STATE with begin = FIRST_L, the .done file, and candidates
    equal to searching exactly the heavy roots."""
    import hss
    limit, first, last, w = 70000, 1, 20000, 125
    state, results = tmp_path / "heavy-v36-w125-1-20000-state.txt", tmp_path / "heavy-v36-w125-1-20000-results.txt"
    r = run([GPU, limit, first, last, results, 0, state, w], tmp_path)
    assert r.returncode == 0, r.stderr
    st = read_state(state)
    assert st["min_weight"] == w and st["begin"] == first and st["resume"] == last + 1
    resumes = [int(ln.split()[2]) for ln in r.stdout.splitlines() if ln.startswith("PROGRESS")]
    assert resumes == [min(first + 4096 * (i + 1), last + 1) for i in range(len(resumes))]
    roots = check_done(state, first, last + 1, limit, w)
    with hss.Context(limit) as ctx:
        want = ctx.search(limit, roots)
    assert st["candidates"] == want.total
    assert result_lines(results.read_text()) == want.hits


def test_heavy_weight_reference_batch(tmp_path):
    """The reference's own heavy-weight run: gpu_search_v36 on [1116384, 1120479]
    at LIMIT 1e7 with min_weight 100 (experiments/heavy/heavy_queue.py:12-14)
    searched 166 rows and counted 1,105,278,882,993 candidates
    (evidence/benchmarks/heavy-benchmark.json:10-23)."""
    limit, first, last, w = 10_000_000, 1116384, 1120479, 100
    stem = f"heavy-v36-w{w}-{first}-{last}"
    state, results = tmp_path / f"{stem}-state.txt", tmp_path / f"{stem}-results.txt"
    r = run([GPU, limit, first, last, results, 0, state, w], tmp_path)
    assert r.returncode == 0, r.stderr
    st = read_state(state)
    assert st == {"version": 1, "limit": limit, "begin": first, "end": last, "resume": last + 1,
                  "candidates": 1_105_278_882_993, "min_weight": w}
    roots = check_done(state, first, last + 1, limit, w)
    assert len(roots) == 166
    assert results.read_text() == ""  # new_solutions 0


def test_heavy_weight_resume_and_mismatch(tmp_path):
    import hss
    limit, first, last, w = 70000, 1, 20000, 125
    state = tmp_path / "h-state.txt"
    # a checkpoint as a stopped worker leaves it after two batches
    state.write_text(f"version 1\nlimit {limit}\nbegin {first}\nend {last}\nresume 8193\ncandidates 5\n"
                     f"min_weight {w}\n")
    r = run([GPU, limit, first, last, tmp_path / "r.txt", 0, state, w], tmp_path)
    assert r.returncode == 0, r.stderr
    st = read_state(state)
    assert st["begin"] == first and st["resume"] == last + 1
    roots = check_done(state, first, last + 1, limit, w)
    with hss.Context(limit) as ctx:
        assert st["candidates"] == ctx.search(limit, [l for l in roots if l >= 8193]).total
    # another weight does not match the checkpoint
    r = run([GPU, limit, first, last, tmp_path / "r.txt", 0, state, 45], tmp_path)
    assert r.returncode == 0 and "checkpoint does not match requested range or weight" in r.stderr
    assert read_state(state)["min_weight"] == w


def test_min_weight_zero_keeps_512_batches(tmp_path):
    state = tmp_path / "s.txt"
    r = run([GPU, 5000, 1, 1200, tmp_path / "r.txt", 0, state, 0], tmp_path)
    assert r.returncode == 0, r.stderr
    st = read_state(state)
    assert st["min_weight"] == 0 and st["resume"] == 1201
    assert [int(ln.split()[2]) for ln in r.stdout.splitlines() if ln.startswith("PROGRESS")] == [513, 1025, 1201]
    assert not (tmp_path / "s.txt.done").exists()


def serve(tmp_path, lines, **kw):
    env = dict(os.environ, HERON_COVERAGE_FILE=str(tmp_path / "none.json"))
    return subprocess.run([str(SERVE), "100000", "0"], input="".join(l + "\n" for l in lines), cwd=tmp_path,
                          env=env, capture_output=True, text=True, timeout=600, **kw)


def test_gpu_serve_files(tmp_path, golden):
    """The served loop writes per request what one gpu_search_v42 process
    writes for its interval: STATE, results lines and a log whose PROGRESS
    chain steps by 512 from FIRST to LAST + 1 (scripts/coverage_audit.py:48-68)."""
    reqs = {"gpu-0a": (5000, 1, 5000), "gpu-1a": (70000, 11789, 11789), "gpu-2a": (100000, 99901, 100000)}
    r = serve(tmp_path, [f"{a} {b} {c} {p}" for p, (a, b, c) in reqs.items()])
    assert r.returncode == 0, r.stderr
    done = {ln.split()[1]: ln for ln in r.stdout.splitlines() if ln.startswith("DONE")}
    assert set(done) == set(reqs)
    want = {"gpu-0a": golden["full5000"], "gpu-1a": golden["known2"], "gpu-2a": golden["i100000_99901_100000"]}
    for p, (limit, first, last) in reqs.items():
        st = read_state(tmp_path / f"{p}-state.txt")
        assert st == {"version": 1, "limit": limit, "begin": first, "end": last, "resume": last + 1,
                      "candidates": want[p]["candidates"]}
        assert result_lines((tmp_path / f"{p}-results.txt").read_text()) == result_lines(want[p]["results"])
        assert chain_end((tmp_path / f"{p}.log").read_text(), first, last) == last + 1
        assert f"resume {last + 1} candidates {want[p]['candidates']}" in done[p]


def test_gpu_serve_checkpoints(tmp_path):
    """Resume from a STATE on the +512 chain, mismatching STATE skipped,
    SIGTERM stops with a STATE on the chain, and the rerun completes it."""
    # a resumed request: the chain continues from the STATE's resume
    (tmp_path / "r-state.txt").write_text("version 1\nlimit 5000\nbegin 1\nend 3000\nresume 1025\ncandidates 9\n")
    (tmp_path / "m-state.txt").write_text("version 1\nlimit 5000\nbegin 1\nend 10\nresume 3\ncandidates 0\n")
    r = serve(tmp_path, ["5000 1 3000 r", "5000 1 20 m"])
    assert r.returncode == 0, r.stderr
    assert "m: checkpoint does not match requested range" in r.stderr
    st = read_state(tmp_path / "r-state.txt")
    assert st["begin"] == 1025 and st["resume"] == 3001
    assert chain_end((tmp_path / "r.log").read_text(), 1025, 3000) == 3001
    import hss
    with hss.Context(5000) as ctx:
        assert st["candidates"] == ctx.search(5000, range(1025, 3001)).total
    # SIGTERM while a long request runs
    limit, first, last = 1_000_000, 1, 30_000
    env = dict(os.environ, HERON_COVERAGE_FILE=str(tmp_path / "none.json"))
    p = subprocess.Popen([str(SERVE), str(limit), "0"], stdin=subprocess.PIPE, stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, cwd=tmp_path, env=env, text=True)
    p.stdin.write(f"{limit} {first} {last} big\n")
    p.stdin.flush()
    log = tmp_path / "big.log"
    deadline = time.time() + 120
    while time.time() < deadline and not (log.exists() and "PROGRESS" in log.read_text()):
        time.sleep(0.02)
    p.send_signal(signal.SIGTERM)
    out, err = p.communicate(timeout=300)
    assert p.returncode == 0, err
    stopped = read_state(tmp_path / "big-state.txt")
    assert (stopped["resume"] - first) % 512 == 0 or stopped["resume"] == last + 1
    r = subprocess.run([str(SERVE), str(limit), "0"], input=f"{limit} {first} {last} big\n", cwd=tmp_path,
                       env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    final = read_state(tmp_path / "big-state.txt")
    assert final["resume"] == last + 1
    assert chain_end(log.read_text(), first, last) == last + 1
    with hss.Context(limit) as ctx:
        total = ctx.search_interval(limit, first, last)[0]
    assert stopped["candidates"] + final["candidates"] == total


def test_queue_on_gpu(tmp_path):
    """The interval queue (hss_queue.py, counterpart of scripts/gpu_queue.py)
    on this box's GPU: the split by work is computed on GPU 0, one
    gpu_search_v42 process runs the part, and its files are the reference's
    (STATE complete, +512 PROGRESS chain in the log, the full5000 result)."""
    import sys
    c = CASES["full5000"]
    p = subprocess.run([sys.executable, str(PKG / "hss_queue.py"), str(c["limit"]), str(c["first"]), str(c["last"]),
                        "--gpus", "0", "--suffix", "q", "--dir", str(tmp_path)],
                       capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, (p.stdout, p.stderr)
    assert "PART gpu=0 range=1-5000" in p.stdout and "DONE gpu=0 range=1-5000" in p.stdout
    st = read_state(tmp_path / "gpu-0q-state.txt")
    assert (st["limit"], st["end"], st["resume"], st["candidates"]) == (5000, 5000, 5001, c["candidates"])
    assert chain_end((tmp_path / "gpu-0q.log").read_text(), 1, 5000) == 5001
    assert (tmp_path / "gpu-0q-results.txt").read_text() == c["results"]
