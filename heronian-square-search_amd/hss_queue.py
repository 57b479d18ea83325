"""Interval queue over GPUs: the counterpart of the reference supervisor
scripts/gpu_queue.py with its intervals cut by estimated work.

The reference gives GPU g of a stage the fixed block [first + 12500 g,
first + 12500 g + 12499] (scripts/gpu_queue.py:224-226) and rebalances by
killing the queue when some GPUs fall idle (scripts/gpu_rebalance.py:65-108).
Here the interval [FIRST, LAST] is split into one contiguous part per GPU of
about equal planner work (hss_split_by_work), so the parts finish together.
Everything else keeps the reference's contract, so its audit and recovery
tools read the files unchanged:

* one `gpu_search_v42 LIMIT first last RESULTS GPU STATE` process per GPU and
  part (8 argv fields, the shape `ps -C gpu_search_v42` matching expects,
  scripts/gpu_queue.py:85-98), stdout and stderr appended to the part's log;
* files gpu-<g><suffix>-state.txt, gpu-<g><suffix>-results.txt,
  gpu-<g><suffix>.log in the working directory (scripts/gpu_queue.py:185-189);
* a checkpoint counts only when its resume is backed by the log's PROGRESS
  chain of +512 steps (scripts/gpu_queue.py:40-61); an unbacked one is moved
  aside as `.unverified-<ns>` before relaunching (:64-82);
* an incomplete part is relaunched (it resumes from its STATE) until complete,
  at most three failed attempts in a row (:272-283);
* SIGTERM / SIGINT are forwarded to the running processes (:286-292);
* lines START gpu=.. pid=.. range=.., DONE gpu=.. range=.., ERROR gpu=..: ..,
  QUEUE FINISHED on stdout.

    python hss_queue.py LIMIT FIRST LAST [--gpus 0,1,...] [--suffix S]
                        [--dir D] [--binary PATH] [--bounds b0,b1,...]

--bounds gives the part boundaries directly (len(gpus) + 1 values) instead of
computing them on the first GPU.  Exit status 0 when every part completed.
"""
from __future__ import annotations

import argparse
import datetime
import signal
import subprocess
import sys
import threading
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
STEP = 512  # positions per gpu_search batch: the PROGRESS chain step


def read_state(path: Path) -> dict:
    if not path.exists():
        return {}
    values = {}
    for line in path.read_text().splitlines():
        fields = line.split()
        if len(fields) == 2:
            try:
                values[fields[0]] = int(fields[1])
            except ValueError:
                pass
    return values


def checkpoint_resume(limit: int, first: int, last: int, state: Path, log: Path) -> int | None:
    """The STATE's resume when the log's PROGRESS chain backs it, else None."""
    values = read_state(state)
    resume = values.get("resume", 0)
    if (values.get("limit") != limit or values.get("end") != last
            or not first <= values.get("begin", 0) <= resume <= last + 1):
        return None
    if not log.exists():
        return None
    expected = first
    for line in log.read_text().splitlines():
        fields = line.split()
        if len(fields) < 3 or fields[:2] != ["PROGRESS", "resume"]:
            continue
        current = int(fields[2])
        if current <= expected:
            continue
        if current != min(expected + STEP, last + 1):
            return None
        expected = current
    return expected if expected == resume else None


def complete(limit: int, first: int, last: int, state: Path, log: Path) -> bool:
    return checkpoint_resume(limit, first, last, state, log) == last + 1


def prepare_checkpoint(limit: int, first: int, last: int, state: Path, log: Path) -> None:
    """Move aside a STATE (and its log) whose resume the log does not back."""
    if checkpoint_resume(limit, first, last, state, log) is not None:
        return
    suffix = f".unverified-{time.time_ns()}"
    for source in (state, log):
        if source.exists():
            destination = source.with_name(source.name + suffix)
            source.replace(destination)
            print(f"PRESERVED UNVERIFIED CHECKPOINT {source.name} -> {destination.name}", flush=True)


class Queue:
    def __init__(self, limit: int, parts: list[tuple[int, int, int]], suffix: str, workdir: Path, binary: Path):
        self.limit, self.parts, self.suffix, self.dir, self.binary = limit, parts, suffix, workdir, binary
        self.stop = threading.Event()
        self.children: dict[int, subprocess.Popen] = {}
        self.lock = threading.Lock()
        self.ok: dict[int, bool] = {}

    def files(self, gpu: int) -> tuple[Path, Path, Path]:
        prefix = self.dir / f"gpu-{gpu}{self.suffix}"
        return (prefix.with_name(prefix.name + "-state.txt"), prefix.with_name(prefix.name + "-results.txt"),
                prefix.with_name(prefix.name + ".log"))

    def launch(self, gpu: int, first: int, last: int) -> subprocess.Popen:
        state, results, log = self.files(gpu)
        with log.open("a") as output:
            process = subprocess.Popen([str(self.binary), str(self.limit), str(first), str(last), str(results),
                                        str(gpu), str(state)], cwd=self.dir, stdout=output,
                                       stderr=subprocess.STDOUT)
        with self.lock:
            self.children[gpu] = process
        print(f"START gpu={gpu} pid={process.pid} range={first}-{last} state={state.name} "
              f"at={datetime.datetime.now(datetime.timezone.utc).isoformat()}", flush=True)
        return process

    def worker(self, gpu: int, first: int, last: int) -> None:
        state, results, log = self.files(gpu)
        try:
            failures = 0
            while not self.stop.is_set() and not complete(self.limit, first, last, state, log):
                prepare_checkpoint(self.limit, first, last, state, log)
                process = self.launch(gpu, first, last)
                process.wait()
                with self.lock:
                    self.children.pop(gpu, None)
                if complete(self.limit, first, last, state, log):
                    break
                if self.stop.is_set():
                    break
                failures += 1
                if failures >= 3:
                    raise RuntimeError(f"part {first}-{last} made no complete checkpoint after three launches")
                self.stop.wait(1)
            if complete(self.limit, first, last, state, log):
                self.ok[gpu] = True
                print(f"DONE gpu={gpu} range={first}-{last}", flush=True)
        except Exception as error:  # reported per GPU, like the reference
            print(f"ERROR gpu={gpu}: {error}", flush=True)

    def request_stop(self, signum, frame) -> None:
        self.stop.set()
        with self.lock:
            processes = list(self.children.values())
        for process in processes:
            if process.poll() is None:
                process.send_signal(signal.SIGTERM)

    def run(self) -> bool:
        threads = [threading.Thread(target=self.worker, args=part, daemon=True) for part in self.parts]
        for thread in threads:
            thread.start()
        for thread in threads:
            while thread.is_alive():
                thread.join(0.2)
        print("QUEUE FINISHED", flush=True)
        return all(self.ok.get(gpu) for gpu, _, _ in self.parts)


def work_bounds(limit: int, first: int, last: int, parts: int, device: int) -> list[int]:
    """Split by estimated work, computed on `device`; the context (and the
    device) is released before any search process starts."""
    sys.path.insert(0, str(HERE))
    import hss
    ctx = hss.Context(limit, device)
    try:
        return ctx.split_by_work(limit, first, last, parts)
    finally:
        ctx.close()


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("limit", type=int)
    ap.add_argument("first", type=int)
    ap.add_argument("last", type=int)
    ap.add_argument("--gpus", default="0", help="comma-separated GPU ids (distinct)")
    ap.add_argument("--suffix", default="", help="stage suffix of the file names")
    ap.add_argument("--dir", default=".", help="directory of the state, results and log files")
    ap.add_argument("--binary", default=str(HERE / "gpu_search_v42"))
    ap.add_argument("--bounds", default=None, help="part boundaries, len(gpus) + 1 comma-separated values")
    a = ap.parse_args(argv)
    gpus = [int(g) for g in a.gpus.split(",")]
    if len(set(gpus)) != len(gpus) or any(g < 0 for g in gpus):
        print("invalid --gpus", file=sys.stderr)
        return 2
    if not 1 <= a.first <= a.last <= a.limit:
        print("need 1 <= FIRST <= LAST <= LIMIT", file=sys.stderr)
        return 2
    if a.bounds:
        bounds = [int(b) for b in a.bounds.split(",")]
        if (len(bounds) != len(gpus) + 1 or bounds[0] != a.first or bounds[-1] != a.last + 1
                or bounds != sorted(bounds)):
            print("--bounds must run from FIRST to LAST + 1, one part per GPU", file=sys.stderr)
            return 2
    else:
        bounds = work_bounds(a.limit, a.first, a.last, len(gpus), gpus[0])
    parts = [(g, bounds[i], bounds[i + 1] - 1) for i, g in enumerate(gpus) if bounds[i + 1] > bounds[i]]
    for g, b, e in parts:
        print(f"PART gpu={g} range={b}-{e}", flush=True)
    queue = Queue(a.limit, parts, a.suffix, Path(a.dir).resolve(), Path(a.binary))
    signal.signal(signal.SIGTERM, queue.request_stop)
    signal.signal(signal.SIGINT, queue.request_stop)
    return 0 if queue.run() else 1


if __name__ == "__main__":
    sys.exit(main())
