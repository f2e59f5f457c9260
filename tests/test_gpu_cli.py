"""ccs_search on the GPU: the output file of a campaign equals the oracle's filter + prover
over the same tiles; checkpoint / resume continue exactly; a sequential prefix sweep visits
every tile once and stops; host proving and host confirmation agree with the GPU prove."""
import os
import subprocess
import sys

import numpy as np
import pytest

import ccs
from ccs import workloads as WL
from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "cunningham-chain-search_amd", "ccs_search")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(*args, timeout=240):
    say("[cli]", " ".join(args))
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return r.stdout


def result_lines(path):
    return sorted(l for l in open(path).read().splitlines() if l and not l.startswith("#"))


def oracle_lines(oracle, t0, tiles, log, target=18, bits=89):
    p = WL.shape_params("reference")
    o = oracle.Shape(p["crt"], p["wheel"], p["target"], p["depth"], p["filter_primes"], p["pos0_max_q"])
    surv, _ = o.filter(t0, tiles, cap=1 << 20)
    pr = o.prove(t0, surv, bits=bits)
    res = np.zeros(len(pr), ccs.CHAIN_RESULT_DTYPE)
    res["survivor_idx"] = np.arange(len(pr))
    res["flags"], res["chain_len"], res["back_steps"] = pr[:, 0], pr[:, 1], pr[:, 2]
    res["root_lo"], res["root_hi"] = pr[:, 4], pr[:, 5]
    return sorted(ccs.format_results(res, log_threshold=log, target=target).splitlines())


def test_sequential_run_checkpoint_and_resume(tmp_path, oracle):
    t0, _ = WL.reference_range(seed=0x5EED, bits=89)
    out1, out2, ck = tmp_path / "a.txt", tmp_path / "b.txt", tmp_path / "run.ckpt"
    common = ["--bits", "89", "--log", "5", "--batch", "2000", "--report", "1000"]
    run(*common, "--start", str(t0), "--tiles", "6000", "--output", str(out1), "--checkpoint", str(ck))
    got1 = result_lines(out1)
    assert got1 == oracle_lines(oracle, t0, 6000, 5) and len(got1) > 0
    kv = dict(l.split("=", 1) for l in open(ck).read().splitlines()[1:])
    assert (int(kv["tile_hi"], 16) << 64) | int(kv["tile_lo"], 16) == t0 + 6000
    assert kv["total_tiles"] == "6000" and kv["prefix"] == "none"
    run(*common, "--resume", str(ck), "--tiles", "4000", "--output", str(out2))
    assert result_lines(out2) == oracle_lines(oracle, t0 + 6000, 4000, 5)
    text = open(out1).read()
    assert text.startswith("# CCS search") and "# Finished:" in text


def test_prefix_sweep_visits_every_tile_once(tmp_path, ctx):
    s = WL.make_shape(ctx, "reference")
    prefix = "0b1" + "0101" * 9 + "11"            # 39 bits: a small slice of the 89-bit range
    lo, hi = s.tile_range(89, prefix, lanes=3, lane_id=1)
    out = run("--bits", "89", "--prefix", prefix, "--prefix-mode", "sequential", "--prefix-lanes", "3",
              "--prefix-lane-id", "1", "--batch", "4096", "--log", "30", "--report", "1000")
    assert "[PREFIX] Sequential sweep complete" in out
    tiles = int(out.split("Tiles:")[1].split()[0])
    assert tiles == hi - lo


def test_host_prove_and_confirm_agree_with_gpu(tmp_path):
    t0, _ = WL.reference_range(seed=0x77, bits=89)
    a, b = tmp_path / "gpu.txt", tmp_path / "cpu.txt"
    common = ["--bits", "89", "--log", "3", "--batch", "3000", "--start", str(t0), "--tiles", "9000",
              "--report", "1000"]
    out = run(*common, "--output", str(a), "--confirm")
    assert "all logged chains agree" in out
    run(*common, "--output", str(b), "--cpu-prove", "--prove-threads", "8")
    assert result_lines(a) == result_lines(b) and len(result_lines(a)) > 0


def test_verify_root_and_time_limit():
    import json
    vec = json.load(open(os.path.join(GOLDEN, "reference_vectors.json")))
    root = vec["q7_immunized_roots"][0]
    out = run("--verify-root", root)
    assert "REACHABLE" in out and "NOT REACHABLE" not in out and "chain: prime" in out
    out = run("--verify-root", "0x" + format((1 << 88) + 1, "x"))
    assert "NOT REACHABLE" in out
    out = run("--bits", "89", "--time", "2", "--seed", "11", "--batch", "16384", "--report", "1")
    assert "[TIME]" in out and "RANDOM tile selection" in out
