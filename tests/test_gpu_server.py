"""Served form: requests of several depths / widths / lengths submitted while others run;
every job's survivors and prove results are the same alone and in the mix, and equal the
oracle's class-form filter."""
import sys
import threading
import time

import numpy as np
import pytest

import ccs
from ccs import workloads as WL

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def shapes(ctx):
    return {d: WL.make_shape(ctx, f"c5d{d}", classes=False) for d in (12, 16, 20)}


def run_server(ctx, shapes, reqs, stagger=0.0, **kw):
    srv = ccs.Server(ctx, shapes, **kw)
    ids = []
    for r in reqs:
        ids.append(srv.submit(r["depth"], r["a"], r["u0"], r["count"], tag=r["id"]))
        if stagger:
            time.sleep(stagger)
    srv.drain()
    done = {}
    while True:
        got = srv.poll()
        if not got:
            break
        for d in got:
            done[d["job_id"]] = d
    out = {}
    for jid, r in zip(ids, reqs):
        u_off, res = srv.results(jid)
        order = np.argsort(u_off, kind="stable")
        out[r["id"]] = ([int(x) for x in u_off[order]], res[order])
        assert done[jid]["t_done"] >= done[jid]["t_admit"] >= done[jid]["t_submit"]
        assert done[jid]["n_survivors"] == len(u_off)
        assert done[jid]["flags"] == 0
    st = srv.stats()
    srv.shutdown()
    return out, st


def test_alone_equals_mixed_and_oracle(ctx, oracle, shapes):
    reqs = [r for r in WL.served_requests(n_jobs=60, seed=0x5E, lo_exp=14, hi_exp=22) if r["depth"] in shapes]
    for r in reqs:          # lower depths keep survivors frequent enough to compare
        r["depth"] = {12: 12, 16: 16, 20: 20}.get(r["depth"], r["depth"])
    reqs = reqs[:24]
    say(f"[server] {len(reqs)} jobs mixed")
    mixed, st = run_server(ctx, shapes, reqs, stagger=0.002, slice_candidates=1 << 22, min_chunk=1 << 16)
    assert st["candidates"] == sum(r["count"] for r in reqs)
    say("[server] each job alone")
    for r in reqs[:8]:
        alone, _ = run_server(ctx, shapes, [r], slice_candidates=1 << 22, min_chunk=1 << 16)
        a_u, a_res = alone[r["id"]]
        m_u, m_res = mixed[r["id"]]
        assert a_u == m_u
        for f in ("chain_len", "back_steps", "flags", "root_lo", "root_hi"):
            assert np.array_equal(a_res[f], m_res[f]), f
    say("[server] oracle")
    for r in reqs:
        p = WL.shape_params(f"c5d{r['depth']}")
        o = oracle.Shape(p["crt"], p["wheel"], p["target"], p["depth"], p["filter_primes"], p["pos0_max_q"])
        js, _ = o.filter_class(r["a"], r["W"], r["u0"], r["count"])
        assert [int(x) for x in js] == mixed[r["id"]][0]
        m_res = mixed[r["id"]][1]
        for u, res in zip(js[:20], m_res[:20]):
            n = r["a"] + r["W"] * (r["u0"] + int(u))
            f, root, ln, back = oracle.prove_one(n)
            assert (int(res["flags"]), int(res["chain_len"]), int(res["back_steps"])) == (f, ln, back)


def test_admission_by_memory_budget(ctx, shapes):
    reqs = [r for r in WL.served_requests(n_jobs=40, seed=0x5F, lo_exp=14, hi_exp=18) if r["depth"] in shapes][:12]
    probe = ccs.Server(ctx, shapes, slice_candidates=1 << 20, min_chunk=1 << 14)
    st = probe.stats()
    probe.shutdown()
    # one job's resident state: the phases of its widest profile's sieve primes (at least the
    # stage-B primes of depth 12) plus its rows in the three lanes' slice tables
    assert st["job_bytes"] >= 4 * int(shapes[12].info.n_tierb) + 3 * 80
    # a budget for two resident jobs: the rest wait for admission, all complete
    srv = ccs.Server(ctx, shapes, mem_budget_bytes=2 * st["job_bytes"] + 100, slice_candidates=1 << 20,
                     min_chunk=1 << 14)
    assert srv.stats()["slots"] == 2
    for r in reqs:
        srv.submit(r["depth"], r["a"], r["u0"], r["count"], tag=r["id"])
    peak = peak_bytes = 0
    t_end = time.time() + 120
    seen = 0
    while seen < len(reqs) and time.time() < t_end:
        st = srv.stats()
        peak = max(peak, st["active"])
        peak_bytes = max(peak_bytes, st["resident_bytes"])
        seen += len(srv.poll())
        time.sleep(0.001)
    srv.shutdown()
    assert seen == len(reqs)
    assert peak <= 2 and peak_bytes <= 2 * st["job_bytes"]


def test_resident_state_equals_recomputed(ctx, shapes):
    """the same jobs with resident phases and with phases computed per unit give the same
    survivors and prove results (several slices per job, so the resident phases are
    advanced across slices)"""
    reqs = [r for r in WL.served_requests(n_jobs=60, seed=0x6A, lo_exp=20, hi_exp=24) if r["depth"] in shapes][:12]
    say(f"[server] resident vs recomputed, {len(reqs)} jobs")
    res_r, st_r = run_server(ctx, shapes, reqs, slice_candidates=1 << 24, min_chunk=1 << 16)
    res_c, _ = run_server(ctx, shapes, reqs, slice_candidates=1 << 24, min_chunk=1 << 16, recompute=True)
    assert st_r["slices"] > 4
    for r in reqs:
        assert res_r[r["id"]][0] == res_c[r["id"]][0]
        for f in ("chain_len", "back_steps", "flags", "root_lo", "root_hi"):
            assert np.array_equal(res_r[r["id"]][1][f], res_c[r["id"]][1][f]), f


def test_launch_groups_do_not_change_results(ctx, shapes):
    """the depth profiles in one launch per kernel instance (ratio 1), grouped by widths
    within a factor of two, and in the default grouping (one launch per row layout) give the
    same per-job results"""
    reqs = [r for r in WL.served_requests(n_jobs=60, seed=0x6D, lo_exp=18, hi_exp=22) if r["depth"] in shapes][:12]
    say(f"[server] launch groups, {len(reqs)} jobs")
    base, st = run_server(ctx, shapes, reqs, slice_candidates=1 << 23, min_chunk=1 << 16)
    launches = {}
    for ratio in (1, 2):
        other, st2 = run_server(ctx, shapes, reqs, slice_candidates=1 << 23, min_chunk=1 << 16,
                                group_width_ratio=ratio)
        launches[ratio] = st2["launches"] / st2["slices"]
        for r in reqs:
            assert base[r["id"]][0] == other[r["id"]][0]
            for f in ("chain_len", "back_steps", "flags", "root_lo", "root_hi"):
                assert np.array_equal(base[r["id"]][1][f], other[r["id"]][1][f]), f
    # a slice makes at most one launch per group: with ratio 1 a group per kernel instance
    # (row layout and stage-B register width), by default one per row layout
    def width(sh):
        g = max(1, -(-int(sh.info.n_tierb) // 64))
        return g if g <= 8 else (12 if g <= 12 else 16)
    rows = {sh.tuning["rows_per_group"] for sh in shapes.values()}
    instances = {(sh.tuning["rows_per_group"], width(sh)) for sh in shapes.values()}
    assert 1 <= launches[1] <= len(instances)
    assert 1 <= launches[2] <= len(instances)
    assert 1 <= st["launches"] / st["slices"] <= len(rows)


def test_shares_in_several_pieces_do_not_change_results(ctx, shapes):
    """a job's share of a slice above max_chunk goes into the launch as several pieces (rows
    of the job table with their own offsets): the per-job results equal those with one
    piece per job and slice"""
    reqs = [r for r in WL.served_requests(n_jobs=60, seed=0x6E, lo_exp=20, hi_exp=23) if r["depth"] in shapes][:10]
    say(f"[server] shares in pieces, {len(reqs)} jobs")
    base, st = run_server(ctx, shapes, reqs, slice_candidates=1 << 24, min_chunk=1 << 16)
    pieces, st_p = run_server(ctx, shapes, reqs, slice_candidates=1 << 24, min_chunk=1 << 16, max_chunk=1 << 17)
    assert st_p["candidates"] == st["candidates"] == sum(r["count"] for r in reqs)
    # with 2^24-candidate slices over at most 10 jobs a share spans several 2^17 pieces: the
    # largest job is done in fewer slices than it has pieces
    assert st_p["slices"] * (1 << 17) < max(r["count"] for r in reqs)
    for r in reqs:
        assert base[r["id"]][0] == pieces[r["id"]][0]
        for f in ("chain_len", "back_steps", "flags", "root_lo", "root_hi"):
            assert np.array_equal(base[r["id"]][1][f], pieces[r["id"]][1][f]), f


def test_survivor_offsets_beyond_2_32(ctx, oracle, shapes):
    """a request longer than 2^32 candidates: survivors past offset 2^32 are reported at
    their full 64-bit offset, equal to the oracle's class-form filter there"""
    r = [x for x in WL.served_requests(n_jobs=40, seed=0x6B) if x["depth"] == 12][0]
    count = (1 << 32) + (1 << 22)
    say(f"[server] one depth-12 request of {count} candidates")
    out, st = run_server(ctx, shapes, [dict(r, count=count)], slice_candidates=1 << 31)
    u_off = out[r["id"]][0]
    p = WL.shape_params("c5d12")
    o = oracle.Shape(p["crt"], p["wheel"], p["target"], p["depth"], p["filter_primes"], p["pos0_max_q"])
    lo = [u for u in u_off if u < (1 << 22)]
    hi = [u - (1 << 32) for u in u_off if u >= (1 << 32)]
    want_lo, _ = o.filter_class(r["a"], r["W"], r["u0"], 1 << 22)
    want_hi, _ = o.filter_class(r["a"], r["W"], r["u0"] + (1 << 32), 1 << 22)
    assert len(want_hi) > 0
    assert lo == [int(x) for x in want_lo]
    assert hi == [int(x) for x in want_hi]
    assert st["candidates"] == count


def test_overflow_marks_jobs_incomplete(ctx, shapes):
    """a stage-B buffer too small for a slice: the slice's jobs finish flagged incomplete and
    drain reports it"""
    reqs = [r for r in WL.served_requests(n_jobs=60, seed=0x6C) if r["depth"] == 12][:3]
    ctx.limit_stage_b(16)
    try:
        srv = ccs.Server(ctx, shapes, slice_candidates=1 << 24, min_chunk=1 << 16)
        for r in reqs:
            srv.submit(12, r["a"], r["u0"], 1 << 23, tag=r["id"])
        complete = srv.drain()
        done = srv.poll()
        st = srv.stats()
        srv.shutdown()
    finally:
        ctx.limit_stage_b(0)
    assert not complete
    assert len(done) == 3 and all(d["flags"] & ccs._lib.JOB_INCOMPLETE for d in done)
    assert st["incomplete"] == 3
