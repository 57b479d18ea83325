"""Command-line contract checks that never reach the device: argument
validation and checkpoint rules of the reference executables
(src/gpu_search_v42.cu:610-614, src/search_v18.cpp:532-561)."""
from __future__ import annotations

import subprocess

import pytest

from conftest import PKG

GPU = PKG / "gpu_search"
CPU = PKG / "search"


@pytest.fixture(scope="module", autouse=True)
def built():
    if not (GPU.is_file() and CPU.is_file()):
        subprocess.run(["make", "-C", str(PKG)], check=True, capture_output=True)


def run(args, cwd):
    return subprocess.run([str(a) for a in args], cwd=cwd, capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("argv", [[], ["100"], ["100", "1", "2"], ["1", "2", "3", "r", "0", "s", "x"]])
def test_gpu_search_usage(tmp_path, argv):
    r = run([GPU] + argv, tmp_path)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.parametrize("argv", [[], ["100", "1", "2", "1"], ["0", "1", "1", "1", "s"], ["100", "0", "1", "1", "s"],
                                  ["100", "5", "4", "1", "s"], ["100", "1", "101", "1", "s"],
                                  ["100", "1", "2", "0", "s"]])
def test_search_validation(tmp_path, argv):
    r = run([CPU] + argv, tmp_path)
    assert r.returncode == 2


def test_search_checkpoint_mismatch(tmp_path):
    (tmp_path / "s").write_text("version 1\nlimit 100\nbegin 1\nend 50\nresume 10\npairs 0\ncandidates 0\n")
    r = run([CPU, 100, 1, 60, 1, "s"], tmp_path)
    assert r.returncode == 2 and "does not match" in r.stderr
    r = run([CPU, 200, 1, 50, 1, "s"], tmp_path)
    assert r.returncode == 2


def test_search_checkpoint_complete(tmp_path):
    (tmp_path / "s").write_text("version 1\nlimit 100\nbegin 1\nend 50\nresume 51\npairs 0\ncandidates 0\n")
    r = run([CPU, 100, 1, 50, 1, "s"], tmp_path)
    assert r.returncode == 0 and "checkpoint already complete" in r.stdout


def test_supervisor_names_link_to_executables():
    assert (PKG / "gpu_search_v42").resolve() == GPU.resolve()
    assert (PKG / "search_v13").resolve() == CPU.resolve()
