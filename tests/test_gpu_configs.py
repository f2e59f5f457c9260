"""Parity of the configurations at their own sizes and depths:

* config 4 (classes mod 37# at depth 18): tens of jobs of 2^24 candidates against the
  oracle's class-form filter (survivors and stage counters), and the whole 4096-job set
  through size-independent properties (every survivor passes the direct predicate, the
  count is within 6 sigma of the analytic expectation);
* config 3 (M = 37#, primes to 2^18) at depth 10, where candidates reach the deep primes
  65537..262139, against the oracle's tile filter;
* config 1 over its full range (t in [t0, t0 + 2^24), 2.1e11 candidates) against the
  oracle's counts and survivor digest committed in tests/golden/c1_full_range.json.
"""
import hashlib
import json
import math
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import ccs
from ccs import workloads as WL
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

CRT = [5, 7, 11, 13, 17, 19]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def class_oracle(oracle, name):
    """the class-form predicate needs only the filter primes and the depth: a small lattice"""
    p = WL.shape_params(name)
    return oracle.Shape(CRT, [23], 18, p["depth"], p["filter_primes"], p["pos0_max_q"])


def threads():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return 4


@pytest.fixture(scope="module")
def c4(ctx):
    return WL.make_shape(ctx, "c4", classes=False)


def test_config4_depth18_jobs_equal_oracle(ctx, oracle, c4):
    jobs_all, W = WL.batched_jobs(n_jobs=32)
    jobs = [(a, u0, 1 << 24) for a, u0, _ in jobs_all]
    say(f"[c4] {len(jobs)} jobs of 2^24 at depth 18, tuning {c4.tuning}")
    js = ccs.JobSet(c4, jobs)
    cap = 1 << 16
    d_out, d_cnt = ctx.dalloc(cap * 8), ctx.dalloc(64)
    try:
        js.launch(d_out, cap, d_cnt)
        st = js.stats(d_cnt)
        n = st["survivors"]
        out = np.zeros(n, ccs._lib.JOB_SURVIVOR_DTYPE)
        if n:
            ctx.d2h(out, d_out)
        assert not js.status()["overflow"]
    finally:
        js.close()
        ctx.dfree(d_out)
        ctx.dfree(d_cnt)
    o = class_oracle(oracle, "c4")
    cuts = [st["dense_end"], st["tierb_end"], len(c4.filter_primes)]
    with ThreadPoolExecutor(threads()) as ex:
        res = list(ex.map(lambda j: o.filter_class(j[0], W, j[1], j[2], cap=1 << 16, cuts=cuts), jobs))
    want = sorted((i, int(u)) for i, (us, _) in enumerate(res) for u in us)
    got = sorted(zip(out["job_idx"].tolist(), out["u_off"].tolist()))
    counts = np.sum([c for _, c in res], axis=0)
    assert st["candidates"] == counts[0] == 32 << 24
    assert (st["alive_dense"], st["alive_tierb"], st["survivors"]) == tuple(int(x) for x in counts[1:])
    assert got == want and len(got) > 0


def test_config4_full_set_properties(ctx, oracle, c4):
    """the bench's 4096-job set (2.05e12 candidates) in one launch"""
    jobs, W = WL.batched_jobs(n_jobs=4096)
    say(f"[c4] full set: {len(jobs)} jobs, {sum(c for _, _, c in jobs)} candidates")
    js = ccs.JobSet(c4, jobs)
    cap = 1 << 20
    d_out, d_cnt = ctx.dalloc(cap * 8), ctx.dalloc(64)
    try:
        js.launch(d_out, cap, d_cnt)
        st = js.stats(d_cnt)
        n = st["survivors"]
        assert n < cap
        out = np.zeros(n, ccs._lib.JOB_SURVIVOR_DTYPE)
        ctx.d2h(out, d_out)
        assert not js.status()["overflow"]
    finally:
        js.close()
        ctx.dfree(d_out)
        ctx.dfree(d_cnt)
    total = sum(c for _, _, c in jobs)
    assert st["candidates"] == total
    fs = [(q, len(WL.forbidden(q, 0 if q <= 97 else 1, 18))) for q in c4.filter_primes]
    _, surv = WL.expected_tests(fs)
    expect = total * surv
    say(f"[c4] {n} survivors, {expect:.1f} expected")
    assert abs(n - expect) < 6 * math.sqrt(expect)
    o = class_oracle(oracle, "c4")
    nfp = len(c4.filter_primes)
    keys = set()
    for j, u in zip(out["job_idx"].tolist(), out["u_off"].tolist()):
        a, u0, c = jobs[j]
        assert u < c
        keys.add((j, u))
        assert o.first_kill_brute(a + W * (u0 + u)) == nfp
    assert len(keys) == n          # no survivor reported twice


def test_config3_deep_primes_equal_oracle(ctx, oracle):
    """depth 10 keeps ~2e-5 of the candidates alive past 2^16, so the deep stage's primes
    65537..262139 (hash probes of the largest tables) decide real candidates"""
    s = WL.make_shape(ctx, "c3", depth=10)
    o_p = WL.shape_params("c3")
    o = oracle.Shape(o_p["crt"], o_p["wheel"], o_p["target"], 10, o_p["filter_primes"], o_p["pos0_max_q"])
    t0, _, _ = WL.c3_range()
    tc = 300
    say(f"[c3 depth 10] tiles={tc} tuning={s.tuning}")
    g_surv, st = s.filter(t0, tc, cap=1 << 18)
    fp = s.filter_primes
    big = sum(1 for q in fp if q < 65536)
    cuts = [st["dense_end"], st["tierb_end"], big, len(fp)]
    assert st["tierb_end"] < big < len(fp)
    o_surv, counts = o.filter(t0, tc, cap=1 << 18, cuts=cuts, nthreads=threads())
    assert st["candidates"] == counts[0]
    assert (st["alive_dense"], st["alive_tierb"]) == (counts[1], counts[2])
    say(f"[c3 depth 10] alive past 2^16: {counts[3]}, survivors {counts[4]}")
    assert counts[3] > counts[4] > 0       # the primes above 2^16 kill some candidates
    assert st["survivors"] == counts[4]
    got = sorted(zip(g_surv["tile_idx"].tolist(), g_surv["base_idx"].tolist(), g_surv["k_offset"].tolist()))
    want = sorted((int(x[2]), int(x[0]), int(x[1])) for x in o_surv)
    assert got == want


def test_config1_full_range_equals_golden(ctx):
    g = json.load(open(os.path.join(GOLDEN, "c1_full_range.json")))
    s = WL.make_shape(ctx, "c1")
    # stage cuts at two of the golden file's cut points: the 37 primes below 200 dense,
    # the other 463 in stage B (no deep stage)
    s.set_tuning(n_dense=37, n_tierb=463)
    t0, tc, _ = WL.c1_range()
    assert (str(t0), tc) == (g["tile_start"], g["tile_count"])
    say(f"[c1 full] tiles={tc} tuning={s.tuning}")
    surv, st = s.filter(t0, tc, cap=1 << 26)
    assert st["candidates"] == g["candidates"]
    alive = dict(zip(g["cuts"], g["alive_after"]))
    assert st["dense_end"] in alive and st["tierb_end"] in alive
    assert st["alive_dense"] == alive[st["dense_end"]]
    assert st["alive_tierb"] == alive[st["tierb_end"]]
    assert st["survivors"] == g["survivors"] == len(surv)
    keys = np.stack([surv["tile_idx"], surv["base_idx"], surv["k_offset"]], 1).astype("<u4")
    keys = np.ascontiguousarray(keys[np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))])
    assert keys[:8].tolist() == g["first_survivors"] and keys[-8:].tolist() == g["last_survivors"]
    assert hashlib.sha256(keys.tobytes()).hexdigest() == g["survivor_sha256"]
