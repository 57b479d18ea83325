#!/usr/bin/env python3
"""Test double of gpu_search's file contract (no GPU): argv LIMIT FIRST LAST
RESULTS GPU STATE; resumes from STATE; per batch of 512 positions rewrites STATE
(version/limit/begin/end/resume/candidates) and prints the PROGRESS line.
Environment: FAKE_STOP_AFTER=<n> exits after n batches of this process (an
incomplete run); FAKE_NO_PROGRESS=1 exits at once without any batch."""
import os
import sys
from pathlib import Path

limit, first, last = (int(x) for x in sys.argv[1:4])
results, state = Path(sys.argv[4]), Path(sys.argv[6])
if os.environ.get("FAKE_NO_PROGRESS"):
    sys.exit(0)
values = {}
if state.exists():
    values = {k: int(v) for k, v in (line.split() for line in state.read_text().splitlines())}
    if values.get("limit") != limit or values.get("end") != last:
        print("checkpoint does not match requested range", file=sys.stderr)
        sys.exit(0)
begin = values.get("resume", first)
stop_after = int(os.environ.get("FAKE_STOP_AFTER", "0"))
results.touch()
nxt, candidates, batches = begin, 0, 0
while nxt <= last:
    batch = min(512, last - nxt + 1)
    candidates += batch  # one candidate per position
    nxt += batch
    tmp = state.with_name(state.name + ".tmp")
    tmp.write_text(f"version 1\nlimit {limit}\nbegin {begin}\nend {last}\nresume {nxt}\ncandidates {candidates}\n")
    tmp.replace(state)
    print(f"PROGRESS resume {nxt} candidates {candidates} seconds 0", flush=True)
    batches += 1
    if stop_after and batches >= stop_after and nxt <= last:
        sys.exit(0)
