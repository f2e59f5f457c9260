"""ccs_filter_launch_noalloc (the launch that allocates no device memory: its scratch is a
fixed per-device array of the library) gives what ccs_filter_launch gives, on two calls in
a row (the second smaller, so stale contents of the shared scratch would show), and rejects
a tile range whose stage-B survivors the fixed scratch cannot hold."""
import sys

import numpy as np
import pytest

import ccs
from ccs import workloads as WL

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(ctx, launch, t0, tiles, cap=1 << 16):
    d_out, d_cnt = ctx.dalloc(cap * 16), ctx.dalloc(64)
    try:
        launch(t0, tiles, d_out, cap, d_cnt)
        ctx.synchronize()
        n = np.zeros(1, np.uint32)
        ctx.d2h(n, d_cnt)
        assert n[0] <= cap
        surv = np.zeros(n[0], ccs.SURVIVOR_DTYPE)
        if n[0]:
            ctx.d2h(surv, d_out)
    finally:
        ctx.dfree(d_out)
        ctx.dfree(d_cnt)
    return int(n[0]), sorted(zip(surv["tile_idx"].tolist(), surv["base_idx"].tolist(), surv["k_offset"].tolist()))


@pytest.mark.parametrize("name", ["c2", "reference"])
def test_noalloc_equals_filter_launch(ctx, name):
    s = WL.make_shape(ctx, name)
    t0 = WL.c2_range()[0] if name == "c2" else WL.reference_range(seed=0x77, bits=89)[0]
    for tiles in ((600000, 200000) if name == "c2" else (60000, 6000)):   # the second call smaller
        say(f"[noalloc] {name} tiles={tiles}")
        want = run(ctx, s.filter_launch, t0, tiles)
        got = run(ctx, s.filter_launch_noalloc, t0, tiles)
        assert got == want and want[0] > 0


def test_noalloc_rejects_oversized_range(ctx):
    s = WL.make_shape(ctx, "c2")
    t0, tc, _ = WL.c2_range()
    d_out, d_cnt = ctx.dalloc(1 << 20), ctx.dalloc(64)
    try:
        with pytest.raises(ccs.CcsError, match="no-allocation scratch"):
            s.filter_launch_noalloc(t0, 8 * tc, d_out, 1 << 16, d_cnt)   # 2^39 candidates
        # the full 2^36-candidate step of config 2 fits
        s.filter_launch_noalloc(t0, tc, d_out, 1 << 16, d_cnt)
        ctx.synchronize()
        n = np.zeros(1, np.uint32)
        ctx.d2h(n, d_cnt)
        assert 0 < n[0] < 1000
    finally:
        ctx.dfree(d_out)
        ctx.dfree(d_cnt)
