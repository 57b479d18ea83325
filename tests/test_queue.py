"""The interval queue (heronian-square-search_amd/hss_queue.py, counterpart of
the reference's scripts/gpu_queue.py) on the CPU, with a test double of the
search binary's file contract (tests/queue/fake_gpu_search.py): parts, file
names, the PROGRESS-chain check of checkpoints, relaunch after an incomplete
run, unverified checkpoints moved aside, and the failure limit."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
QUEUE = ROOT / "heronian-square-search_amd" / "hss_queue.py"
FAKE = ROOT / "tests" / "queue" / "fake_gpu_search.py"


def run(tmp_path, *args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    p = subprocess.run([sys.executable, str(QUEUE), *map(str, args), "--dir", str(tmp_path), "--binary", str(FAKE)],
                       capture_output=True, text=True, env=e, timeout=120)
    return p.returncode, p.stdout, p.stderr


def state(tmp_path, name):
    return {k: int(v) for k, v in (l.split() for l in (tmp_path / name).read_text().splitlines())}


def chain_ok(log: Path, first: int, last: int) -> bool:
    expected = first
    for line in log.read_text().splitlines():
        f = line.split()
        if f[:2] != ["PROGRESS", "resume"]:
            continue
        cur = int(f[2])
        if cur <= expected:
            continue
        if cur != min(expected + 512, last + 1):
            return False
        expected = cur
    return expected == last + 1


def test_two_parts_complete(tmp_path):
    rc, out, err = run(tmp_path, 100000, 1, 3000, "--gpus", "3,5", "--bounds", "1,1700,3001", "--suffix", "ab")
    assert rc == 0, (out, err)
    assert "PART gpu=3 range=1-1699" in out and "PART gpu=5 range=1700-3000" in out
    assert "DONE gpu=3 range=1-1699" in out and "DONE gpu=5 range=1700-3000" in out and "QUEUE FINISHED" in out
    for g, (a, b) in ((3, (1, 1699)), (5, (1700, 3000))):
        s = state(tmp_path, f"gpu-{g}ab-state.txt")
        assert (s["limit"], s["end"], s["resume"]) == (100000, b, b + 1)
        assert chain_ok(tmp_path / f"gpu-{g}ab.log", a, b)
        assert (tmp_path / f"gpu-{g}ab-results.txt").exists()


def test_incomplete_run_is_resumed(tmp_path):
    rc, out, err = run(tmp_path, 100000, 1, 2000, "--gpus", "0", "--bounds", "1,2001", env={"FAKE_STOP_AFTER": "2"})
    assert rc == 0, (out, err)
    assert out.count("START gpu=0") == 2  # 2 x 2 batches of 512 cover 2000 positions
    assert chain_ok(tmp_path / "gpu-0.log", 1, 2000)
    assert state(tmp_path, "gpu-0-state.txt")["resume"] == 2001


def test_unverified_checkpoint_moved_aside(tmp_path):
    (tmp_path / "gpu-0-state.txt").write_text("version 1\nlimit 100000\nbegin 1\nend 700\nresume 701\ncandidates 5\n")
    rc, out, err = run(tmp_path, 100000, 1, 700, "--bounds", "1,701")
    assert rc == 0, (out, err)
    assert "PRESERVED UNVERIFIED CHECKPOINT gpu-0-state.txt" in out
    assert list(tmp_path.glob("gpu-0-state.txt.unverified-*"))
    assert state(tmp_path, "gpu-0-state.txt")["candidates"] == 700  # recomputed, not the stale 5


def test_complete_part_is_not_relaunched(tmp_path):
    assert run(tmp_path, 100000, 1, 600, "--bounds", "1,601")[0] == 0
    rc, out, _ = run(tmp_path, 100000, 1, 600, "--bounds", "1,601")
    assert rc == 0 and "START" not in out and "DONE gpu=0 range=1-600" in out


def test_failure_limit(tmp_path):
    rc, out, _ = run(tmp_path, 100000, 1, 600, "--bounds", "1,601", env={"FAKE_NO_PROGRESS": "1"})
    assert rc == 1
    assert out.count("START gpu=0") == 3 and "ERROR gpu=0" in out


def test_bad_arguments(tmp_path):
    assert run(tmp_path, 100000, 10, 9)[0] == 2
    assert run(tmp_path, 100000, 1, 10, "--gpus", "1,1", "--bounds", "1,5,11")[0] == 2
    assert run(tmp_path, 100000, 1, 10, "--gpus", "0,1", "--bounds", "1,11")[0] == 2
