"""Batched form: one persistent work-stealing launch over many class-form jobs, the
per-job launch loop and the captured HIP graph give the same survivors as the oracle."""
import sys

import numpy as np
import pytest

import ccs
from ccs import workloads as WL

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_jobs(oracle, jobs, W, depth, cuts):
    p = WL.shape_params("c4")
    o = oracle.Shape(p["crt"], p["wheel"], 18, depth, p["filter_primes"], p["pos0_max_q"])
    want, alive = [], [0] * (len(cuts) + 1)
    for ji, (a, u0, count) in enumerate(jobs):
        js, al = o.filter_class(a, W, u0, count, cuts=cuts)
        want += [(ji, int(x)) for x in js]
        alive = [x + y for x, y in zip(alive, al)]
    return sorted(want), alive


def read_out(ctx, d_out, d_count):
    n = np.zeros(1, np.uint32)
    ctx.d2h(n, d_count)
    out = np.zeros(int(n[0]), ccs.JOB_SURVIVOR_DTYPE)
    if len(out):
        ctx.d2h(out, d_out)
    return sorted(zip(out["job_idx"].tolist(), out["u_off"].tolist()))


def test_jobset_launch_each_and_graph(ctx, oracle):
    jobs, W = WL.batched_jobs(n_jobs=40, seed=0xB5, lo_exp=8, hi_exp=19)
    s = WL.make_shape(ctx, "c4", classes=False, depth=9)
    js = ccs.JobSet(s, jobs)
    cap = 1 << 18
    d_out = ctx.dalloc(cap * 8)
    d_count = ctx.dalloc(64)
    say(f"[batch] {len(jobs)} jobs, {js.candidates} candidates")
    js.launch(d_out, cap, d_count)
    ctx.synchronize()
    got = read_out(ctx, d_out, d_count)
    st = js.stats(d_count)
    want, alive = oracle_jobs(oracle, jobs, W, 9, [st["dense_end"], st["tierb_end"], len(s.filter_primes)])
    assert got == want and len(got) > 20
    assert (st["candidates"], st["alive_dense"], st["alive_tierb"], st["survivors"]) == tuple(alive)
    say("[batch] per-job loop")
    js.launch_each(d_out, cap, d_count)
    ctx.synchronize()
    assert read_out(ctx, d_out, d_count) == want
    st2 = js.stats(d_count)
    assert (st2["alive_dense"], st2["alive_tierb"]) == (st["alive_dense"], st["alive_tierb"])
    say("[batch] graph replays")
    g = js.graph(d_out, cap, d_count)
    for rep in range(4):
        js.graph_launch(g)
        ctx.synchronize()
        assert read_out(ctx, d_out, d_count) == want, rep
        # eager work between replays
        js.launch_each(d_out, cap, d_count)
        ctx.synchronize()
    js.close()
    ctx.dfree(d_out)
    ctx.dfree(d_count)


def test_jobs_with_empty_and_tiny_counts(ctx, oracle):
    jobs, W = WL.batched_jobs(n_jobs=12, seed=0xB6, lo_exp=0, hi_exp=12)
    jobs[3] = (jobs[3][0], jobs[3][1], 0)
    s = WL.make_shape(ctx, "c4", classes=False, depth=6)
    say("[batch] empty and tiny jobs")
    g, st = s.batch(jobs)
    want, alive = oracle_jobs(oracle, jobs, W, 6, [st["dense_end"], st["tierb_end"], len(s.filter_primes)])
    assert sorted(zip(g["job_idx"].tolist(), g["u_off"].tolist())) == want
    assert st["candidates"] == alive[0]


def test_autotune_keeps_results(ctx, oracle):
    s = WL.make_shape(ctx, "reference")
    t0, _ = WL.reference_range(seed=0x31, bits=89)
    say("[autotune] reference shape, 6 trials")
    tuning, ms = s.autotune(t0, 6000, max_trials=6)
    assert ms > 0 and tuning == s.tuning
    p = WL.shape_params("reference")
    o = oracle.Shape(p["crt"], p["wheel"], p["target"], p["depth"], p["filter_primes"], p["pos0_max_q"])
    g, st = s.filter(t0, 1200)
    o_surv, counts = o.filter(t0, 1200, cap=1 << 18, cuts=[st["dense_end"], st["tierb_end"]])
    assert sorted(zip(g["tile_idx"].tolist(), g["base_idx"].tolist(), g["k_offset"].tolist())) == \
        sorted((int(x[2]), int(x[0]), int(x[1])) for x in o_surv)
    assert (st["candidates"], st["alive_dense"], st["alive_tierb"]) == tuple(counts)


def test_autotune_jobs_keeps_results(ctx, oracle):
    jobs, W = WL.batched_jobs(n_jobs=16, seed=0xB7, lo_exp=12, hi_exp=18)
    s = WL.make_shape(ctx, "c4", classes=False, depth=9)
    say("[autotune] job-set form, 6 trials")
    tuning, ms = s.autotune_jobs(jobs, max_trials=6)
    assert ms > 0 and tuning == s.tuning
    g, st = s.batch(jobs)
    want, alive = oracle_jobs(oracle, jobs, W, 9, [st["dense_end"], st["tierb_end"], len(s.filter_primes)])
    assert sorted(zip(g["job_idx"].tolist(), g["u_off"].tolist())) == want
    assert (st["candidates"], st["alive_dense"], st["alive_tierb"], st["survivors"]) == tuple(alive)


def test_jobset_unit_size_and_tuning_change(ctx, oracle):
    """A set sizes its own work units (smaller for small sets, per job in the per-job loop):
    results do not depend on the unit size, on a tuning change after the set was built, or
    on an update to fewer jobs."""
    jobs, W = WL.batched_jobs(n_jobs=24, seed=0xB8, lo_exp=14, hi_exp=21)
    s = WL.make_shape(ctx, "c4", classes=False, depth=8)
    js = ccs.JobSet(s, jobs)
    cap = 1 << 18
    d_out, d_cnt = ctx.dalloc(cap * 8), ctx.dalloc(64)
    say("[batch] unit sizes")
    js.launch(d_out, cap, d_cnt)
    ctx.synchronize()
    st0 = js.stats(d_cnt)
    want, _ = oracle_jobs(oracle, jobs, W, 8, [st0["dense_end"], st0["tierb_end"], len(s.filter_primes)])
    assert read_out(ctx, d_out, d_cnt) == want and len(want) > 20
    uj = s.tuning["unit_j"]
    s.set_tuning(unit_j=uj // 2 if uj > 65536 else uj * 2)
    js.launch(d_out, cap, d_cnt)
    ctx.synchronize()
    assert read_out(ctx, d_out, d_cnt) == want
    js.launch_each(d_out, cap, d_cnt)
    ctx.synchronize()
    assert read_out(ctx, d_out, d_cnt) == want
    js.update(jobs[:5])
    js.launch(d_out, cap, d_cnt)
    ctx.synchronize()
    assert read_out(ctx, d_out, d_cnt) == [x for x in want if x[0] < 5]
    js.close()
    ctx.dfree(d_out)
    ctx.dfree(d_cnt)


def test_mixed_depth_group_equals_separate_sets(ctx, oracle):
    """Jobs of three depth profiles of one CRT shape family (the served form's groups) in
    one group launch give each job exactly the survivors of its own one-shape set; prove
    results equal the one-shape sets' too."""
    shapes, jobs = [], []
    rng = np.random.default_rng(0xD3)
    for k, depth in enumerate((6, 9, 12)):
        s = WL.make_shape(ctx, "c4", classes=False, depth=depth)
        shapes.append(s)
        jl, _ = WL.batched_jobs(n_jobs=5 + k, seed=0xE0 + k, lo_exp=12, hi_exp=19)
        jobs += [(k, j) for j in jl]
    say(f"[batch] group of {len(shapes)} depths, {len(jobs)} jobs")
    gs = ccs.GroupSet(shapes, jobs)
    cap = 1 << 18
    d_out, d_cnt = ctx.dalloc(cap * 8), ctx.dalloc(64)
    gs.launch(d_out, cap, d_cnt)
    ctx.synchronize()
    got = read_out(ctx, d_out, d_cnt)
    want = []
    base = 0
    for k, s in enumerate(shapes):
        mine = [j for kk, j in jobs if kk == k]
        js = ccs.JobSet(s, mine)
        js.launch(d_out, cap, d_cnt)
        ctx.synchronize()
        want += [(base + i, u) for i, u in read_out(ctx, d_out, d_cnt)]
        js.close()
        base += len(mine)
    assert got == sorted(want) and len(want) > 50
    # prove through the group set equals the host engine on the same n
    gs.launch(d_out, cap, d_cnt)
    ctx.synchronize()
    n = np.zeros(1, np.uint32)
    ctx.d2h(n, d_cnt)
    d_res = ctx.dalloc(cap * 32)
    gs.prove_launch(d_out, d_cnt, cap, d_res)
    ctx.synchronize()
    surv = np.zeros(int(n[0]), ccs.JOB_SURVIVOR_DTYPE)
    ctx.d2h(surv, d_out)
    res = np.zeros(int(n[0]), ccs.CHAIN_RESULT_DTYPE)
    ctx.d2h(res, d_res)
    values = []
    for sv in surv:
        k, (a, u0, _) = jobs[int(sv["job_idx"])]
        values.append(a + shapes[k].W * (u0 + int(sv["u_off"])))
    host, _ = ccs.confirm_values(values)
    for f in ("flags", "chain_len", "back_steps", "root_lo", "root_hi"):
        assert (res[f] == host[f]).all(), f
    gs.close()
    for p in (d_out, d_cnt, d_res):
        ctx.dfree(p)
