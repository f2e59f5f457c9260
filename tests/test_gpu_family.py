"""Fingerprint families on the GPU (the v16 Q-iteration front end): family shapes built from
pool entries sieve their Q windows exactly like the oracle's restatement of the v16
predicate, the reference's q-iter chains come out with their lengths, the K1.5 top gate
keeps exactly the survivors whose top position passes strong MR, and the prove of family
survivors equals the host confirmation engine."""
import json
import os
import random
import sys

import numpy as np
import pytest

import ccs
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

PRIMES = [q for q in range(3, 1010) if all(q % d for d in range(2, int(q ** 0.5) + 1))]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def spec_value(spec):
    D = 1
    for tok in spec.split("*"):
        p, _, k = tok.partition("^")
        D *= int(p) ** int(k or 1)
    return D


def run_set(ctx, shape, jobs, cap=1 << 20):
    js = ccs.JobSet(shape, jobs)
    d_out, d_cnt = ctx.dalloc(cap * 8), ctx.dalloc(64)
    js.launch(d_out, cap, d_cnt)
    ctx.synchronize()
    n = np.zeros(1, np.uint32)
    ctx.d2h(n, d_cnt)
    assert n[0] <= cap
    out = np.zeros(int(n[0]), ccs.JOB_SURVIVOR_DTYPE)
    if len(out):
        ctx.d2h(out, d_out)
    return js, d_out, d_cnt, out


def strong_mr(n, a):
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    x = pow(a, d, n)
    if x in (1, n - 1):
        return True
    for _ in range(s - 1):
        x = x * x % n
        if x == n - 1:
            return True
    return False


def test_pool_families_match_oracle(ctx, oracle):
    pool = ccs.pool_load(os.path.join(GOLDEN, "pools", "known_fingerprints_top10.txt"))
    rng = random.Random(0xFA)
    for e in pool:
        D = e["D"]
        shape = ccs.FamilyShape(ctx, D, 1, 18, 8, PRIMES)
        qmin, qmax = shape.window(89, 92)
        jobs = []
        for _ in range(3):
            count = rng.randrange(1 << 14, 1 << 17)
            q0 = rng.randrange(qmin, qmax - count)
            jobs.append((0, q0, count))
        say(f"[family] {e['name']} D={D} jobs={len(jobs)}")
        js, d_out, d_cnt, out = run_set(ctx, shape, jobs)
        got = sorted(zip(out["job_idx"].tolist(), out["u_off"].tolist()))
        want = []
        for ji, (_, q0, count) in enumerate(jobs):
            want += [(ji, int(k)) for k in oracle.filter_family(D, 1, 8, PRIMES, q0, count)]
        assert got == sorted(want) and len(want) > 10, e["name"]
        st = js.stats(d_cnt)
        assert st["candidates"] == sum(j[2] for j in jobs) and st["survivors"] == len(want)
        js.close()
        ctx.dfree(d_out)
        ctx.dfree(d_cnt)
        shape.close()


def test_reference_qiter_chains(ctx):
    vecs = json.load(open(os.path.join(GOLDEN, "reference_vectors.json")))["v16_qiter_family"]
    for v in vecs:
        D, Q, L = spec_value(v["D_spec"]), int(v["Q"], 16), v["len"]
        say(f"[family] q-iter vector CC{L} D={v['D_spec']}")
        shape = ccs.FamilyShape(ctx, D, 0, L, L, PRIMES)
        js, d_out, d_cnt, out = run_set(ctx, shape, [(0, Q - 5000, 10000)])
        qs = [Q - 5000 + int(u) for u in out["u_off"]]
        assert Q in qs
        res = shape.prove_jobs([(0, Q - 5000, 10000)], out)
        r = res[qs.index(Q)]
        assert (int(r["root_lo"]) | (int(r["root_hi"]) << 64)) == int(v["p0"], 16)
        assert int(r["chain_len"]) == L and int(r["back_steps"]) == 0
        js.close()
        ctx.dfree(d_out)
        ctx.dfree(d_cnt)
        shape.close()


def test_top_gate_equals_host_mr(ctx):
    D = 2 * 3 * 5 * 11 * 13 * 19
    shape = ccs.FamilyShape(ctx, D, 1, 18, 10, PRIMES)
    qmin, qmax = shape.window(89, 90)
    jobs = [(0, qmin + 12345, 1 << 20), (0, qmax - (1 << 19), 1 << 19)]
    js, d_out, d_cnt, out = run_set(ctx, shape, jobs)
    assert len(out) > 100
    cap = 1 << 20
    g_out, g_cnt = ctx.dalloc(cap * 8), ctx.dalloc(64)
    W = D << 1
    for top, reps in [(17, 2), (17, 3), (9, 2), (0, 2)]:
        js.top_gate(d_out, d_cnt, cap, top, reps, g_out, cap, g_cnt)
        ctx.synchronize()
        n = np.zeros(1, np.uint32)
        ctx.d2h(n, g_cnt)
        g = np.zeros(int(n[0]), ccs.JOB_SURVIVOR_DTYPE)
        if len(g):
            ctx.d2h(g, g_out)
        got = sorted(zip(g["job_idx"].tolist(), g["u_off"].tolist()))
        want = []
        for s in out:
            q = jobs[int(s["job_idx"])][1] + int(s["u_off"])
            p_top = (W * q << top) - 1
            if all(strong_mr(p_top, a) for a in (2, 3, 5)[:reps]):
                want.append((int(s["job_idx"]), int(s["u_off"])))
        say(f"[family] top gate top={top} reps={reps}: {len(want)} of {len(out)}")
        assert got == sorted(want)
        assert 0 < len(want) < len(out)
    with pytest.raises(ccs.CcsError):
        js.top_gate(d_out, d_cnt, cap, 17, 4, g_out, cap, g_cnt)
    js.close()
    for p in (d_out, d_cnt, g_out, g_cnt):
        ctx.dfree(p)
    shape.close()


def test_family_prove_equals_host_engine(ctx):
    D = 2 * 3 ** 2 * 5 * 11 * 13 * 19
    shape = ccs.FamilyShape(ctx, D, 0, 18, 7, PRIMES)
    qmin, _ = shape.window(62, 64)
    jobs = [(0, qmin + 999, 1 << 19)]
    js, d_out, d_cnt, out = run_set(ctx, shape, jobs)
    assert len(out) > 300
    res = shape.prove_jobs(jobs, out)
    values = [(jobs[0][1] + int(s["u_off"])) * D - 1 for s in out]
    host, _ = ccs.confirm_values(values)
    for f in ("flags", "chain_len", "back_steps", "root_lo", "root_hi"):
        assert (res[f] == host[f]).all(), f
    assert int((res["flags"] & ccs.F_PRIME != 0).sum()) > 10
    js.close()
    ctx.dfree(d_out)
    ctx.dfree(d_cnt)
    shape.close()


def test_family_validation(ctx):
    D = 2 * 3 * 5 * 11 * 13 * 19
    shape = ccs.FamilyShape(ctx, D, 1, 18, 12, PRIMES)
    assert shape.info.family == 1 and shape.info.W == D << 1
    assert shape.info.n_filter == len([q for q in PRIMES if D % q])
    for jobs in ([(5, 1000, 10)], [(0, 0, 10)], [(0, (1 << 127) // (D << 1), 2)]):
        with pytest.raises(ccs.CcsError):
            ccs.JobSet(shape, jobs)
    with pytest.raises(ccs.CcsError):
        shape.filter(0, 10)               # no tiles on a family shape
    shape.close()
    for bad in [dict(D=0), dict(D=1 << 127), dict(filter_primes=[2, 3, 5])]:
        kw = dict(D=D, exp_start=0, target=18, depth=12, filter_primes=PRIMES)
        kw.update(bad)
        with pytest.raises(ccs.CcsError):
            ccs.FamilyShape(ctx, **kw)


def test_pool_group_launch_matches_oracle(ctx, oracle):
    """All ten families of a pool in one persistent launch (a group set), jobs of uneven
    length, some families with several windows and one with an empty job: the same
    survivors as the oracle per family, also through the per-job loop and a HIP graph."""
    pool = ccs.pool_load(os.path.join(GOLDEN, "pools", "known_fingerprints_top10.txt"))
    rng = random.Random(0x6A)
    shapes = [ccs.FamilyShape(ctx, e["D"], 1, 18, 8, PRIMES) for e in pool]
    jobs = []
    for k, sh in enumerate(shapes):
        qmin, qmax = sh.window(89, 92)
        for _ in range(1 + k % 3):
            count = 0 if (k, _) == (4, 1) else rng.randrange(1 << 12, 1 << 17)
            jobs.append((k, (0, rng.randrange(qmin, qmax - (1 << 17)), count)))
    want = []
    for i, (k, (_, q0, count)) in enumerate(jobs):
        want += [(i, int(x)) for x in oracle.filter_family(pool[k]["D"], 1, 8, PRIMES, q0, count)]
    want.sort()
    say(f"[family] group set: {len(shapes)} shapes, {len(jobs)} jobs, {len(want)} survivors")
    gs = ccs.GroupSet(shapes, jobs)
    cap = 1 << 20
    d_out, d_cnt = ctx.dalloc(cap * 8), ctx.dalloc(64)

    def read():
        n = np.zeros(1, np.uint32)
        ctx.d2h(n, d_cnt)
        out = np.zeros(int(n[0]), ccs.JOB_SURVIVOR_DTYPE)
        if len(out):
            ctx.d2h(out, d_out)
        return sorted(zip(out["job_idx"].tolist(), out["u_off"].tolist()))

    gs.launch(d_out, cap, d_cnt)
    ctx.synchronize()
    assert read() == want and len(want) > 100
    gs.launch_each(d_out, cap, d_cnt)
    ctx.synchronize()
    assert read() == want
    g = gs.graph(d_out, cap, d_cnt)
    for _ in range(3):
        gs.graph_launch(g)
        ctx.synchronize()
        assert read() == want
    # an update to the jobs of three families
    sub = [j for j in jobs if j[0] in (2, 5, 9)]
    gs.update(sub)
    gs.launch(d_out, cap, d_cnt)
    ctx.synchronize()
    want2 = []
    for i, (k, (_, q0, count)) in enumerate(sub):
        want2 += [(i, int(x)) for x in oracle.filter_family(pool[k]["D"], 1, 8, PRIMES, q0, count)]
    assert read() == sorted(want2)
    with pytest.raises(ccs.CcsError):
        gs.update([(5, (0, 10 ** 20, 10)), (2, (0, 10 ** 20, 10))])   # not grouped by shape
    gs.close()
    ctx.dfree(d_out)
    ctx.dfree(d_cnt)
    for sh in shapes:
        sh.close()


def test_large_batch_prove_on_base2_pseudoprimes(ctx):
    """More than 4096 survivors take the large-batch prove (strong MR to base 2 per survivor,
    then the Lucas tests of the base-2 probable primes gathered one per lane).  Family shape
    D = 2, E = 0 makes any odd n a survivor (n = 2Q - 1 with Q = (n + 1) / 2), so base-2
    strong pseudoprimes without small factors (rejected only by the Lucas test), primes and
    random odd values go through it: every record equals the host confirmation engine."""
    import sympy
    from test_confirm import strong_pseudoprimes_base2_without_small_factors
    rng = random.Random(0xB2)
    spsp = strong_pseudoprimes_base2_without_small_factors(24, rng)
    primes = [int(sympy.randprime(1 << 80, 1 << 100)) for _ in range(200)]
    values = spsp + primes + [rng.randrange(1 << 60, 1 << 100) | 1 for _ in range(4400)]
    rng.shuffle(values)
    shape = ccs.FamilyShape(ctx, 2, 0, 2, 1, [3])
    jobs = [(0, (v + 1) // 2, 1) for v in values]
    surv = np.zeros(len(values), ccs.JOB_SURVIVOR_DTYPE)
    surv["job_idx"] = np.arange(len(values), dtype=np.uint32)
    say(f"[family] large-batch prove: {len(values)} values, {len(spsp)} base-2 strong pseudoprimes")
    res = shape.prove_jobs(jobs, surv)
    host, _ = ccs.confirm_values(values)
    for f in ("flags", "chain_len", "back_steps", "root_lo", "root_hi"):
        assert (res[f] == host[f]).all(), f
    prime = {v for v, r in zip(values, res) if int(r["flags"]) & ccs.F_PRIME}
    assert not prime & set(spsp) and set(primes) <= prime
    shape.close()
