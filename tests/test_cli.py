"""ccs_search command line: host-only behaviour (self-test, argument validation).  The
GPU campaign runs are in tests/test_gpu_cli.py."""
import os
import subprocess

import pytest

from conftest import ROOT

CLI = os.path.join(ROOT, "cunningham-chain-search_amd", "ccs_search")

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="ccs_search is not built (build.py)")


def run(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)


def test_self_test_passes_without_gpu():
    r = run("--test")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "self-test: all passed" in r.stdout and "CC13  PASS" in r.stdout


def test_help():
    r = run("--help")
    assert r.returncode == 0 and "--prefix-lanes" in r.stdout and "--checkpoint" in r.stdout


@pytest.mark.parametrize("args, msg", [
    (["--bits", "20"], "out of supported range"),
    (["--target", "70"], "out of range"),
    (["--batch", "0"], "out of range"),
    (["--prefix", "0b0101"], "must start with 1"),
    (["--prefix", "101"], "binary format"),
    (["--prefix", "0b1" + "0" * 89], "prefix_bits must be < bits"),
    (["--prefix-mode", "random"], "requires --prefix"),
    (["--prefix", "0b1", "--prefix-mode", "spiral"], "must be 'sequential' or 'random'"),
    (["--prefix-lanes", "4"], "requires --prefix"),
    (["--prefix", "0b1", "--prefix-lanes", "4", "--prefix-lane-id", "4"], "out of range"),
    (["--coverage", "legacy"], "not supported"),
    (["--frobnicate"], "Unknown argument"),
])
def test_argument_errors_before_any_gpu_work(args, msg):
    r = run(*args)
    assert r.returncode == 1 and msg in r.stderr, r.stderr
