"""Stage-B overflow of asynchronous launches is reported (ccs_stream_status /
ccs_jobset_status), as the reference warns on survivor overflow (v13.cu:3846-3852).

The stage-B buffer is shrunk from the host with the test knob ccs_debug_limit_stage_b, so
the overflow path runs on ordinary inputs; every kernel still stays inside its buffers."""
import sys

import numpy as np
import pytest

import ccs
from ccs import workloads as WL

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name,tiles", [("reference", 6000), ("c2", 60000)])
def test_filter_launch_overflow_reported(ctx, name, tiles):
    """plain launches: a shape without deep primes (the emit kernel) and one with them
    (the deep kernel)"""
    say(f"[overflow] {name} tiles={tiles}")
    s = WL.make_shape(ctx, name)
    t0 = WL.reference_range(seed=0x5A, bits=89)[0] if name == "reference" else WL.c2_range()[0]
    surv, st = s.filter(t0, tiles)            # the true counts, unconstrained
    true_b = st["alive_tierb"]
    assert true_b > 200, true_b
    cap = 1 << 16
    d_out, d_cnt = ctx.dalloc(cap * 16), ctx.dalloc(64)
    try:
        assert not ctx.status()["overflow"]    # the synchronous run did not overflow
        ctx.limit_stage_b(100)
        try:
            for _ in range(2):
                s.filter_launch(t0, tiles, d_out, cap, d_cnt)
            got = ctx.status()
            with pytest.raises(ccs.CcsError):   # the synchronous form fails outright
                s.filter(t0, tiles)
        finally:
            ctx.limit_stage_b(0)
        assert got["overflow"] and got["overflowed_launches"] == 2
        assert got["worst_stage_b"] == true_b and got["stage_b_cap"] == 100
        assert got["dropped"] == 2 * (true_b - 100)
        n = np.zeros(1, np.uint32)
        ctx.d2h(n, d_cnt)
        assert n[0] <= len(surv)                # what the deep stage saw: a subset
        ctx.status()                            # the failed synchronous run's record
        assert not ctx.status()["overflow"]     # read once, cleared
        s.filter_launch(t0, tiles, d_out, cap, d_cnt)
        assert not ctx.status()["overflow"]     # the full buffer again
        ctx.d2h(n, d_cnt)
        assert n[0] == len(surv)
    finally:
        ctx.dfree(d_out)
        ctx.dfree(d_cnt)


def test_jobset_overflow_reported(ctx):
    say("[overflow] job set, depth 12")
    s = WL.make_shape(ctx, "c5d12", classes=False)
    reqs = [r for r in WL.served_requests(n_jobs=200, seed=0x0F) if r["depth"] == 12][:6]
    jobs = [(r["a"], r["u0"], 1 << 22) for r in reqs]
    js = ccs.JobSet(s, jobs)
    cap = 1 << 18
    d_out, d_cnt = ctx.dalloc(cap * 8), ctx.dalloc(64)
    try:
        js.launch(d_out, cap, d_cnt)
        true_b = js.stats(d_cnt)["alive_tierb"]
        assert true_b > 100
        assert not js.status()["overflow"]
        ctx.limit_stage_b(50)
        try:
            js.launch(d_out, cap, d_cnt)
            got = js.status()
        finally:
            ctx.limit_stage_b(0)
        assert got["overflow"] and got["overflowed_launches"] == 1
        assert got["worst_stage_b"] == true_b and got["dropped"] == true_b - 50
        js.launch(d_out, cap, d_cnt)
        assert not js.status()["overflow"]
    finally:
        js.close()
        ctx.dfree(d_out)
        ctx.dfree(d_cnt)
