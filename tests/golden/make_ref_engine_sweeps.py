"""Writes ref_engine_sweeps.json: what the reference's own v33 engine (cc_gmp_v33_03.c,
built by `make -C oracle ref` into oracle/_ref/ from its sources) prints for complete
sequential sweeps of two prefix windows, in full mode with --log 1: its stage counters and
every chain it reports (length, root), the latter as a count, a histogram and a SHA-256 of
the sorted "length 0xroot" lines.  tests/test_gpu_ref_sweeps.py checks the GPU path against
it; tests/test_ref_engines.py runs the same sweeps against the oracle where the engine is
at hand.

Run from the repository root: python tests/golden/make_ref_engine_sweeps.py
"""
import collections
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
V33 = os.path.join(ROOT, "oracle", "_ref", "cc_gmp_v33")
SWEEPS = [(89, 18, "0b1011001110001111000011110000111100001111"),
          (80, 14, "0b110100101100011101011110110101010")]


def sweep(bits, target, prefix):
    with tempfile.TemporaryDirectory() as tmp:
        out_file = os.path.join(tmp, "results.txt")
        cmd = [V33, "--prefix", prefix, "--sequential", "--bits", str(bits), "--target", str(target),
               "--threads", "8", "--report", "1", "--log", "1", "--output", out_file]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600, check=True)
        text = r.stdout
        assert "SEARCH COMPLETE" in text

        def num(pat):
            return int(re.search(pat, text).group(1))

        chains = []
        for line in open(out_file):
            m = re.match(r"CC(\d+)a (0x[0-9a-f]+)\s*$", line)
            if m:
                chains.append((int(m.group(1)), int(m.group(2), 16)))
    lines = sorted(f"{ln} {hex(root)}" for ln, root in chains)
    hist = collections.Counter(ln for ln, _ in chains)
    return {"bits": bits, "target": target, "prefix": prefix,
            "wheel": num(r"Wheel candidates \(L1\):\s+(\d+)"), "l2": num(r"Passed L2 bitmask:\s+(\d+)"),
            "ext": num(r"Passed ext L2 \(OPT-7\):\s+(\d+)"), "line": num(r"Line-sieve \(OPT-17\):\s+(\d+) passed"),
            "primes": num(r"Primes found:\s+(\d+)"), "non_roots": num(r"Non-roots skipped:\s+(\d+)"),
            "chains": len(chains), "chains_histogram": {str(k): v for k, v in sorted(hist.items())},
            "chains_sha256": hashlib.sha256("\n".join(lines).encode()).hexdigest()}


def main():
    if not os.access(V33, os.X_OK):
        sys.exit("oracle/_ref/cc_gmp_v33 is not built (make -C oracle ref)")
    out = {"source": "reference engine cc_gmp_v33_03.c (v33-03-CC18a), full mode, --log 1, "
                     "complete sequential prefix sweeps; chains sorted as '<length> 0x<root>' lines",
           "sweeps": [sweep(*s) for s in SWEEPS]}
    path = os.path.join(ROOT, "tests", "golden", "ref_engine_sweeps.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path)


if __name__ == "__main__":
    main()
