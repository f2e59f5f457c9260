"""The GPU path against the reference's own v33 engine (cc_gmp_v33_03.c): for two complete
sequential prefix sweeps (89-bit at target 18, 80-bit at target 14), the GPU's stage
counters equal the engine's wheel / L2 / ext-L2 / line-sieve counts, and the GPU prove of
the survivors reports the same chains (length from the true root, and the root) for every
prime, as the fixture tests/golden/ref_engine_sweeps.json records them
(tests/golden/make_ref_engine_sweeps.py ran the engine to write it)."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import ccs
from ccs import workloads as WL
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

SWEEPS = json.load(open(os.path.join(GOLDEN, "ref_engine_sweeps.json")))["sweeps"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("sw", SWEEPS, ids=[f"bits{s['bits']}-target{s['target']}" for s in SWEEPS])
def test_gpu_equals_reference_engine(ctx, sw):
    say(f"[ref sweep] bits={sw['bits']} target={sw['target']} prefix={sw['prefix']}")
    s = WL.make_shape(ctx, "reference", target=sw["target"], depth=sw["target"])
    # the GPU's stage ends at the engine's: L2 = 37..61, ext-L2 = 67..97 (then the line sieve)
    s.set_tuning(n_dense=7, n_tierb=7)
    lo, hi = s.tile_range(sw["bits"], sw["prefix"])
    surv, st = s.filter(lo, hi - lo, cap=1 << 20)
    assert (st["candidates"], st["alive_dense"], st["alive_tierb"], st["survivors"]) == \
        (sw["wheel"], sw["l2"], sw["ext"], sw["line"])
    # the engine's prover gates on the prefix-clamped n range (cc_gmp_v33_03.c:3036-3039,
    # :2112-2113): the survivors of the edge tiles past it are left out
    pv, pb = int(sw["prefix"][2:], 2), len(sw["prefix"]) - 2
    nmin = max(pv << (sw["bits"] - pb), 1 << (sw["bits"] - 1))
    nmax = min(((pv + 1) << (sw["bits"] - pb)) - 1, (1 << sw["bits"]) - 1)
    B = [int(b) for b in s.bases()]
    inside = np.array([nmin <= ((lo + int(t)) * s.WP + int(k)) * s.M + B[int(b)] <= nmax
                       for b, k, t in zip(surv["base_idx"], surv["k_offset"], surv["tile_idx"])], bool)
    res = s.prove(lo, surv[inside], bits=sw["bits"])
    prime = (res["flags"] & ccs.F_PRIME) != 0
    assert int(prime.sum()) == sw["primes"]
    assert int((prime & (res["back_steps"] > 0)).sum()) == sw["non_roots"]
    roots = [int(r["root_lo"]) | (int(r["root_hi"]) << 64) for r in res[prime]]
    lines = sorted(f"{int(ln)} {hex(rt)}" for ln, rt in zip(res["chain_len"][prime], roots))
    hist = {}
    for ln in res["chain_len"][prime]:
        hist[str(int(ln))] = hist.get(str(int(ln)), 0) + 1
    assert dict(sorted(hist.items(), key=lambda kv: int(kv[0]))) == sw["chains_histogram"]
    assert hashlib.sha256("\n".join(lines).encode()).hexdigest() == sw["chains_sha256"]
