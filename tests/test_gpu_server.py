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
        surv, res = srv.results(jid)
        order = np.argsort(surv["u_off"], kind="stable")
        out[r["id"]] = (surv["u_off"][order].tolist(), res[order])
        assert done[jid]["t_done"] >= done[jid]["t_admit"] >= done[jid]["t_submit"]
        assert done[jid]["n_survivors"] == len(surv)
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
    # a budget for two resident jobs: the rest wait for admission, all complete
    srv = ccs.Server(ctx, shapes, mem_budget_bytes=2 * 2 * (32 + 16 * 40), slice_candidates=1 << 20,
                     min_chunk=1 << 14)
    for r in reqs:
        srv.submit(r["depth"], r["a"], r["u0"], r["count"], tag=r["id"])
    peak = 0
    t_end = time.time() + 120
    seen = 0
    while seen < len(reqs) and time.time() < t_end:
        st = srv.stats()
        peak = max(peak, st["active"])
        seen += len(srv.poll())
        time.sleep(0.001)
    srv.shutdown()
    assert seen == len(reqs)
    assert peak <= 2
