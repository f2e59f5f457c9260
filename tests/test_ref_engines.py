"""The oracle against the reference's own CPU engines (v33 `cc_gmp_v33_03.c`, v34
`cc_gmp_v34_bit-vector_10.c`), built by `make -C oracle ref` from their sources into
oracle/_ref/ (GMP 6.2.1 from the image).  Each engine sweeps a complete sequential prefix
range (`--prefix ... --sequential`, cc_gmp_v33_03.c:1690-2293) and prints its stage
counters and chain results; the oracle must reproduce them over the same tile range.

Skipped where the engines cannot be built (no reference sources or no GMP): the GPU host
runs only the -m gpu tests.
"""
import collections
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

REF = os.path.join(ROOT, "oracle", "_ref")
V33 = os.path.join(REF, "cc_gmp_v33")
V34 = os.path.join(REF, "cc_gmp_v34")
REF_SRC = "/root/reference/src/cpu"


@pytest.fixture(scope="module")
def engines():
    if not (os.access(V33, os.X_OK) and os.access(V34, os.X_OK)):
        if os.path.isdir(REF_SRC) and shutil.which("make"):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=False,
                           capture_output=True)
    if not (os.access(V33, os.X_OK) and os.access(V34, os.X_OK)):
        pytest.skip("reference CPU engines not built (no sources or GMP here)")
    return V33, V34


def run_engine(binary, prefix, bits, target, sieve_only, tmp_path=None, threads=8):
    """One complete sequential sweep; returns the FINAL RESULTS counters and, in full mode,
    the (length, root) pairs of the results file (--log 1: every prime's chain)."""
    cmd = [binary, "--prefix", prefix, "--sequential", "--bits", str(bits), "--target", str(target),
           "--threads", str(threads), "--report", "1", "--log", "1"]
    out_file = None
    if sieve_only:
        cmd.append("--sieve-only")
    else:
        out_file = os.path.join(str(tmp_path), "results.txt")
        cmd += ["--output", out_file]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    text = r.stdout
    assert "SEARCH COMPLETE" in text, text[-2000:]

    def num(pat):
        m = re.search(pat, text)
        assert m, pat
        return int(m.group(1))

    c = {"wheel": num(r"Wheel candidates \(L1\):\s+(\d+)"), "l2": num(r"Passed L2 bitmask:\s+(\d+)"),
         "ext": num(r"Passed ext L2 \(OPT-7\):\s+(\d+)"), "line": num(r"Line-sieve \(OPT-17\):\s+(\d+) passed")}
    if not sieve_only:
        c["primes"] = num(r"Primes found:\s+(\d+)")
        c["non_roots"] = num(r"Non-roots skipped:\s+(\d+)")
        chains = collections.Counter()
        for line in open(out_file):
            m = re.match(r"CC(\d+)a (0x[0-9a-f]+)\s*$", line)
            if m:
                chains[(int(m.group(1)), int(m.group(2), 16))] += 1
        c["chains"] = chains
    return c


def oracle_shape(oracle, target):
    from ccs import workloads as WL
    p = WL.shape_params("reference")
    return oracle.Shape(p["crt"], p["wheel"], target, target, p["filter_primes"], p["pos0_max_q"])


def prefix_window(oracle, o, bits, prefix):
    """tile range of the sweep, and the prefix-clamped n range v33's prover gates on
    (cc_gmp_v33_03.c:3036-3039, :2112-2113)"""
    b = [int(x) for x in o.bases()]
    lo, hi = oracle.tile_range(o.M, o.WP, min(b), max(b), bits, prefix)
    pv, pb = int(prefix[2:], 2), len(prefix) - 2
    nmin = max(pv << (bits - pb), 1 << (bits - 1))
    nmax = min(((pv + 1) << (bits - pb)) - 1, (1 << bits) - 1)
    return lo, hi, nmin, nmax


SWEEPS = [(89, 18, "0b1011001110001111000011110000111100001111"),
          (80, 14, "0b110100101100011101011110110101010")]


@pytest.mark.parametrize("bits,target,prefix", SWEEPS)
def test_v33_stage_counters_equal_oracle(engines, oracle, bits, target, prefix):
    """v33's wheel / L2 / ext-L2 / line-sieve counts equal orc_filter's stage counts
    (L2 = primes 37..61, ext-L2 = 67..97 at positions 0..target-1, line = 101..863 at
    positions 1..target-1) over the same tiles."""
    ref = run_engine(engines[0], prefix, bits, target, sieve_only=True)
    o = oracle_shape(oracle, target)
    lo, hi, _, _ = prefix_window(oracle, o, bits, prefix)
    surv, counts = o.filter(lo, hi - lo, cap=1 << 22, cuts=[7, 14])
    assert ref["wheel"] == counts[0] == o.count_candidates(lo, hi - lo)
    assert (ref["l2"], ref["ext"], ref["line"]) == (counts[1], counts[2], len(surv))


@pytest.mark.parametrize("bits,target,prefix", SWEEPS[:1])
def test_v34_stage_counters_equal_oracle_bv(engines, oracle, bits, target, prefix):
    """v34's bit-vector filter (L2 and ext-L2 in one 64-candidate mask) reports the same
    wheel and line-sieve counts, and orc_filter_bv the same survivors."""
    ref = run_engine(engines[1], prefix, bits, target, sieve_only=True)
    o = oracle_shape(oracle, target)
    lo, hi, _, _ = prefix_window(oracle, o, bits, prefix)
    surv, cand = o.filter_bv(lo, hi - lo, cap=1 << 22)
    _, counts = o.filter(lo, hi - lo, cap=1 << 22, cuts=[14])
    assert ref["wheel"] == cand == counts[0]
    assert ref["ext"] == counts[1]
    assert ref["line"] == len(surv)


@pytest.mark.parametrize("bits,target,prefix", SWEEPS)
def test_v33_chains_equal_oracle_prove(engines, oracle, tmp_path, bits, target, prefix):
    """Full mode (--log 1): every prime survivor's chain (length from its true root, the
    root itself) from v33's GMP prover equals the oracle's prove over the survivors whose
    n lies in the prefix window; primes found and non-roots agree too."""
    ref = run_engine(engines[0], prefix, bits, target, sieve_only=False, tmp_path=tmp_path)
    o = oracle_shape(oracle, target)
    lo, hi, nmin, nmax = prefix_window(oracle, o, bits, prefix)
    surv, _ = o.filter(lo, hi - lo, cap=1 << 22)
    ns = np.array([nmin <= o.n_of(lo, row) <= nmax for row in surv], bool)
    res = o.prove(lo, surv[ns], bits=bits)
    prime = (res[:, 0] & 4) != 0
    mine = collections.Counter((int(r[1]), int(r[4]) | (int(r[5]) << 64)) for r in res[prime])
    assert sum(mine.values()) == ref["primes"]
    assert int((prime & (res[:, 2] > 0)).sum()) == ref["non_roots"]
    assert mine == ref["chains"]
    assert ref["primes"] > 100   # the window holds enough primes to mean something
