"""The deep stage's two forms give the same survivors: the batched form (a wave tests a batch
of stage-B records against each group of 64 deep primes, the primes' constants built once
per batch) and the one-record form.  Small launches are forced through either form with the
context's deep_batch test knob and compared with the oracle's class-form filter (one-shape
and group launches); a full survivor-rich launch, where the batched form is the one chosen
by count, equals the same launch in the one-record form."""
import sys

import numpy as np
import pytest

import ccs
from ccs import workloads as WL

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def read_out(ctx, d_out, d_count):
    n = np.zeros(1, np.uint32)
    ctx.d2h(n, d_count)
    out = np.zeros(int(n[0]), ccs.JOB_SURVIVOR_DTYPE)
    if len(out):
        ctx.d2h(out, d_out)
    return sorted(zip(out["job_idx"].tolist(), out["u_off"].tolist()))


def launch(ctx, js, per, cap=1 << 20):
    d_out, d_count = ctx.dalloc(cap * 8), ctx.dalloc(64)
    ctx.deep_batch(per)
    try:
        js.launch(d_out, cap, d_count)
        ctx.synchronize()
        return read_out(ctx, d_out, d_count), js.stats(d_count)
    finally:
        ctx.deep_batch(0)
        ctx.dfree(d_out)
        ctx.dfree(d_count)


def oracle_shape(oracle, depth):
    p = WL.shape_params(f"c5d{depth}")
    return oracle.Shape(p["crt"], p["wheel"], p["target"], p["depth"], p["filter_primes"], p["pos0_max_q"])


def test_deep_batch_knob_bounds(ctx):
    with pytest.raises(ccs.CcsError):
        ctx.deep_batch(65)
    ctx.deep_batch(0)


def test_each_form_equals_oracle(ctx, oracle):
    reqs = [r for r in WL.served_requests(n_jobs=300, seed=0x7A, lo_exp=19, hi_exp=22) if r["depth"] == 12][:16]
    s = WL.make_shape(ctx, "c5d12", classes=False)
    js = ccs.JobSet(s, [(r["a"], r["u0"], r["count"]) for r in reqs])
    o = oracle_shape(oracle, 12)
    want = sorted((ji, int(u)) for ji, r in enumerate(reqs)
                  for u in o.filter_class(r["a"], r["W"], r["u0"], r["count"])[0])
    say(f"[deep batch] depth 12, {len(reqs)} jobs, {js.candidates} candidates, {len(want)} survivors")
    assert len(want) > 10
    base = None
    for per in (1, 2, 3, 17, 32, 64):
        got, st = launch(ctx, js, per)
        say(f"[deep batch] per_wave {per}: {len(got)} survivors, {st['alive_tierb']} stage-B records")
        assert got == want
        assert st["alive_tierb"] > 2 * per
        if base is None:
            base = st
        assert (st["alive_dense"], st["alive_tierb"], st["survivors"]) == \
            (base["alive_dense"], base["alive_tierb"], base["survivors"])
    js.close()
    s.close()


def test_group_launch_each_form_equals_oracle(ctx, oracle):
    depths = (12, 14, 20)
    shapes = [WL.make_shape(ctx, f"c5d{d}", classes=False) for d in depths]
    for s in shapes:   # a group launch runs one row layout
        s.set_tuning(rows_per_group=2)
    reqs = WL.served_requests(n_jobs=400, seed=0x7B, lo_exp=18, hi_exp=21)
    jobs, want_reqs = [], []
    for k, d in enumerate(depths):
        for r in [r for r in reqs if r["depth"] == d][:6]:
            jobs.append((k, (r["a"], r["u0"], r["count"])))
            want_reqs.append(r)
    gs = ccs.GroupSet(shapes, jobs)
    want = sorted((ji, int(u)) for ji, r in enumerate(want_reqs)
                  for u in oracle_shape(oracle, r["depth"]).filter_class(r["a"], r["W"], r["u0"], r["count"])[0])
    say(f"[deep batch] group of depths {depths}, {len(jobs)} jobs, {len(want)} survivors")
    for per in (1, 5, 32):
        got, _ = launch(ctx, gs, per)
        assert got == want, per
    gs.close()
    for s in shapes:
        s.close()


def test_plain_launch_forms_agree(ctx):
    """plain (class-table) launches: the batched form forced on a config-3-shaped range equals
    the one-record form, whose survivors the plain-form parity tests compare with the oracle"""
    s = WL.make_shape(ctx, "c3", depth=10)
    t0 = (1 << 90) // (s.WP * s.M)
    tiles = 6 * 400
    ctx.deep_batch(1)
    try:
        one, st1 = s.filter(t0, tiles, cap=1 << 20)
    finally:
        ctx.deep_batch(0)
    ctx.deep_batch(7)
    try:
        bat, st7 = s.filter(t0, tiles, cap=1 << 20)
    finally:
        ctx.deep_batch(0)
    say(f"[deep batch] plain: {st1['alive_tierb']} stage-B records, {st1['survivors']} survivors")
    assert st1["alive_tierb"] > 100 and st1["survivors"] > 0
    key = lambda v: sorted(zip(v["base_idx"].tolist(), v["k_offset"].tolist(), v["tile_idx"].tolist()))
    assert key(one) == key(bat)
    assert st1["survivors"] == st7["survivors"]
    s.close()


def test_full_launch_chosen_form_equals_one_record_form(ctx):
    """2^34 candidates at depth 12 (~4.4e5 stage-B records, several per wave: the batched
    form by count) against the same launch with one record per wave"""
    reqs = [r for r in WL.served_requests(n_jobs=2000, seed=0xC5) if r["depth"] == 12][:64]
    s = WL.make_shape(ctx, "c5d12", classes=False)
    js = ccs.JobSet(s, [(r["a"], r["u0"], 1 << 28) for r in reqs])
    auto, st_a = launch(ctx, js, 0)
    one, st_1 = launch(ctx, js, 1)
    say(f"[deep batch] full: {st_a['alive_tierb']} stage-B records, {len(auto)} survivors")
    assert st_a["alive_tierb"] > 2 * 65536
    assert auto == one and len(auto) == st_a["survivors"] == st_1["survivors"]
    js.close()
    s.close()
