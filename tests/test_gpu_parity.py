"""Parity of the HIP pipeline (through the C ABI) with the CPU oracle.

Bit-exact: survivor sets (sorted, since the GPU emits in arbitrary order), the stage
counters at the GPU's stage boundaries, and every field of the prove results.
"""
import math
import sys

import numpy as np
import pytest

import ccs
from ccs import workloads as WL

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_shape(oracle, name, **over):
    p = WL.shape_params(name)
    p.update(over)
    return oracle.Shape(p["crt"], p["wheel"], p["target"], p["depth"], p["filter_primes"], p["pos0_max_q"])


def sorted_keys(surv):
    """(tile, base, k) keys of plain survivor records, sorted"""
    if isinstance(surv, np.ndarray) and surv.dtype.names:
        k = list(zip(surv["tile_idx"].tolist(), surv["base_idx"].tolist(), surv["k_offset"].tolist()))
    else:
        k = [(int(s[2]), int(s[0]), int(s[1])) for s in surv]
    return sorted(k)


def check_plain(ctx, oracle, name, tile_start, tile_count, shape=None, oshape=None, **tuning):
    say(f"[plain] {name} tile_start={tile_start} tile_count={tile_count} tuning={tuning}")
    s = shape or WL.make_shape(ctx, name)
    if tuning:
        s.set_tuning(**tuning)
    o = oshape or oracle_shape(oracle, name)
    g_surv, st = s.filter(tile_start, tile_count, cap=1 << 20)
    cuts = [st["dense_end"], st["tierb_end"], len(s.filter_primes)]
    o_surv, counts = o.filter(tile_start, tile_count, cap=1 << 20, cuts=cuts)
    assert st["candidates"] == counts[0] == s.count_candidates(tile_start, tile_count)
    assert st["alive_dense"] == counts[1]
    assert st["alive_tierb"] == counts[2]
    assert st["survivors"] == counts[3] == len(o_surv)
    assert sorted_keys(g_surv) == sorted_keys(o_surv)
    return g_surv, o_surv, st


@pytest.fixture(scope="module")
def ref(ctx):
    return WL.make_shape(ctx, "reference")


@pytest.fixture(scope="module")
def oref(oracle):
    return oracle_shape(oracle, "reference")


def test_reference_shape_tables(ref, oref):
    assert (ref.M, ref.WP, ref.n_bases, ref.n_classes) == (1616615, 20677, 36, 123552)
    assert [int(b) for b in ref.bases()] == [int(b) for b in oref.bases()]
    a, b, k, r = ref.classes()
    bases = ref.bases()
    assert np.all(a == (r.astype(object) * 20677 * 1616615 + k.astype(object) * 1616615 +
                        bases[b].astype(object)).astype(np.uint64))
    assert np.all(a % 6 == 5)


def test_reference_campaign_batch(ctx, oracle, ref, oref):
    t0, _ = WL.reference_range(seed=0x13, bits=89, n_tiles=16384)
    g, o, st = check_plain(ctx, oracle, "reference", t0, 2400, shape=ref, oshape=oref)
    assert len(g) > 300


@pytest.mark.parametrize("off", range(6))
@pytest.mark.parametrize("tile_count", [1, 2, 5, 7, 13, 6 * 2048 + 3])
def test_ragged_tile_ranges(ctx, oracle, ref, oref, off, tile_count):
    t0 = (1 << 86) // (20677 * 1616615) // 6 * 6 + off
    check_plain(ctx, oracle, "reference", t0, tile_count, shape=ref, oshape=oref)


def test_tile_range_equals_oracle(oracle, ref):
    """ccs_tile_range (campaign --bits / --prefix / --prefix-lanes) equals the oracle's
    restatement of compute_prefix_tile_range, errors included."""
    say("[plain] tile ranges")
    b = ref.bases()
    args = (ref.M, ref.WP, int(b.min()), int(b.max()))
    rnd = np.random.default_rng(0x7A)
    cases = [(89, None, 0, 0), (89, "0b101", 0, 0), (89, "0b101", 4, 2), (127, "0b11", 3, 0),
             (60, "0b1" + "0" * 58, 0, 0), (60, "0b1" + "0" * 59, 0, 0), (21, None, 0, 0)]
    for _ in range(300):
        bits = int(rnd.integers(40, 128))
        pb = int(rnd.integers(1, min(bits, 20) + 1))
        pv = int(rnd.integers(1 << (pb - 1), 1 << pb))
        lanes = int(rnd.integers(0, 9))
        cases.append((bits, "0b" + format(pv, "b"), lanes, int(rnd.integers(0, max(lanes, 1)))))
    for bits, prefix, lanes, lane in cases:
        try:
            want = oracle.tile_range(*args, bits, prefix, lanes, lane)
        except ValueError:
            with pytest.raises(ccs.CcsError):
                ref.tile_range(bits, prefix, lanes, lane)
            continue
        assert ref.tile_range(bits, prefix, lanes, lane) == want, (bits, prefix, lanes, lane)
    for bad in [(89, None, 2, 0), (89, "0b1", 2, 2), (10, None, 0, 0)]:
        with pytest.raises(ccs.CcsError):
            ref.tile_range(*bad)


def test_empty_range(ref):
    say("[plain] empty range")
    surv, st = ref.filter(12345, 0)
    assert len(surv) == 0 and st["candidates"] == 0 and st["survivors"] == 0


def test_u128_carry_and_high_bits(ctx, oracle, ref, oref):
    # u = U0 + j crosses 2^64, and n reaches ~2^125
    t0 = 6 * ((1 << 64) - 3000)
    check_plain(ctx, oracle, "reference", t0, 36000, shape=ref, oshape=oref)
    t1 = ((1 << 125) // (20677 * 1616615))
    check_plain(ctx, oracle, "reference", t1, 600, shape=ref, oshape=oref)


@pytest.mark.parametrize("tuning", [
    dict(rows_per_group=1), dict(rows_per_group=4), dict(n_dense=0, n_tierb=100),
    dict(n_dense=7, n_tierb=7), dict(n_dense=43, n_tierb=96), dict(n_dense=14, n_tierb=0),
    dict(n_dense=21, n_tierb=64), dict(n_dense=21, n_tierb=65), dict(unit_j=8192), dict(unit_j=24576),
    dict(blocks_per_cu=1), dict(pair_max=4096), dict(pair_max=16384, n_dense=43, n_tierb=96),
    dict(pair_max=16384, n_dense=20, rows_per_group=4), dict(pair_max=16384, n_dense=64, n_tierb=75, rows_per_group=1),
])
def test_tuning_does_not_change_results(ctx, oracle, oref, tuning):
    s = WL.make_shape(ctx, "reference")
    t0 = (1 << 88) // (20677 * 1616615) + 777
    check_plain(ctx, oracle, "reference", t0, 900, shape=s, oshape=oref, **tuning)


def test_config1_shape(ctx, oracle):
    t0, tc, _ = WL.c1_range(seed=0xC1, t_count=1 << 11)
    check_plain(ctx, oracle, "c1", t0, tc)


def test_config2_shape_counters(ctx, oracle):
    s = WL.make_shape(ctx, "c2")
    o = oracle_shape(oracle, "c2")
    t0, _, _ = WL.c2_range(seed=0xC2)
    check_plain(ctx, oracle, "c2", t0, 600, shape=s, oshape=o)


def test_deep_stage_with_survivors(ctx, oracle):
    # shallow depth and few dense/stage-B primes: thousands of survivors pass through the
    # deep stage's hash tests
    s = WL.make_shape(ctx, "c2", depth=8, tuning=dict(n_dense=5, n_tierb=10, rows_per_group=2,
                                                        unit_j=65536, blocks_per_cu=4))
    o = oracle_shape(oracle, "c2", depth=8)
    t0, _, _ = WL.c2_range(seed=0x77)
    g, og, st = check_plain(ctx, oracle, "c2", t0, 300, shape=s, oshape=o)
    assert len(g) > 100


@pytest.mark.parametrize("tuning", [dict(n_tierb=1024), dict(n_tierb=700, rows_per_group=1),
                                    dict(n_dense=20, n_tierb=1000)])
def test_wide_stage_b_vs_oracle(ctx, oracle, tuning):
    """Stage B over up to 1024 primes (16 groups of 64, primes up to ~8700, most bitsets in
    device memory; 700 primes run on the 12-group kernel with padding lanes): at depth 12
    the stage counters and the survivors equal the oracle's."""
    s = WL.make_shape(ctx, "c2", depth=12)
    o = oracle_shape(oracle, "c2", depth=12)
    t0, _, _ = WL.c2_range(seed=0x12B)
    g, og, st = check_plain(ctx, oracle, "c2", t0, 1000, shape=s, oshape=o, **tuning)
    assert st["tierb_end"] - st["dense_end"] == tuning["n_tierb"]
    assert len(g) > 20 and st["alive_tierb"] > st["survivors"]


def test_wide_stage_b_jobs_and_group_vs_oracle(ctx, oracle):
    """Job sets with a 1024-prime stage B, alone and in a group launch beside a shape with
    the default stage B (the group runs the 16-group kernel for both), equal the oracle."""
    jobs, W = WL.batched_jobs(n_jobs=12, seed=0xB16, lo_exp=18, hi_exp=22)
    wide = WL.make_shape(ctx, "c4", classes=False, depth=11, target=18)
    wide.set_tuning(n_tierb=1024)
    narrow = WL.make_shape(ctx, "c4", classes=False, depth=13, target=18)
    want = {}
    for k, (sh, depth) in enumerate(((wide, 11), (narrow, 13))):
        o = oracle_shape(oracle, "c4", depth=depth, target=18)
        want[k] = []
        for ji, (a, u0, count) in enumerate(jobs):
            js, _ = o.filter_class(a, W, u0, count)
            want[k] += [(ji, int(x)) for x in js]
    say(f"[batch] wide stage B: {len(want[0])} and {len(want[1])} survivors")
    g, st = wide.batch(jobs, cap=1 << 20)
    assert sorted(zip(g["job_idx"].tolist(), g["u_off"].tolist())) == sorted(want[0])
    assert st["tierb_end"] - st["dense_end"] == 1024
    gs = ccs.GroupSet([wide, narrow], [(0, j) for j in jobs] + [(1, j) for j in jobs])
    cap = 1 << 20
    d_out, d_cnt = ctx.dalloc(cap * 8), ctx.dalloc(64)
    gs.launch(d_out, cap, d_cnt)
    ctx.synchronize()
    n = np.zeros(1, np.uint32)
    ctx.d2h(n, d_cnt)
    out = np.zeros(int(n[0]), ccs.JOB_SURVIVOR_DTYPE)
    ctx.d2h(out, d_out)
    got = sorted(zip(out["job_idx"].tolist(), out["u_off"].tolist()))
    assert got == sorted(want[0] + [(len(jobs) + ji, u) for ji, u in want[1]])
    assert len(want[0]) > 100 and len(want[1]) > 10
    gs.close()
    ctx.dfree(d_out)
    ctx.dfree(d_cnt)


def test_config3_shape_counters(ctx, oracle):
    s = WL.make_shape(ctx, "c3")
    o = oracle_shape(oracle, "c3")
    t0, _, _ = WL.c3_range(seed=0xC3)
    check_plain(ctx, oracle, "c3", t0, 12, shape=s, oshape=o)


# --- prove ------------------------------------------------------------------------

PROVE_FIELDS = ("chain_len", "back_steps", "flags", "root_lo", "root_hi")


def compare_prove(g, o):
    assert len(g) == len(o)
    for f, col in zip(("flags", "chain_len", "back_steps"), (0, 1, 2)):
        assert np.array_equal(g[f].astype(np.uint64), o[:, col]), f
    assert np.array_equal(g["root_lo"], o[:, 4])
    assert np.array_equal(g["root_hi"], o[:, 5])


def test_prove_reference_survivors(ctx, oracle, ref, oref):
    t0, _ = WL.reference_range(seed=0x99, bits=89)
    o_surv, _ = oref.filter(t0, 3000, cap=1 << 20)
    surv = np.zeros(len(o_surv), ccs.SURVIVOR_DTYPE)
    surv["base_idx"], surv["k_offset"], surv["tile_idx"] = o_surv[:, 0], o_surv[:, 1], o_surv[:, 2]
    say(f"[prove] {len(surv)} reference survivors, bits=89")
    for bits in (89, 0):
        g = ref.prove(t0, surv, bits=bits)
        o = oref.prove(t0, o_surv, bits=bits)
        compare_prove(g, o)
        assert np.any(g["flags"] & ccs.F_PRIME)


def test_prove_both_launch_forms_equal_oracle(ctx, oracle, ref, oref):
    """A prove launch over at most 4096 survivors runs as one kernel (a wave per survivor,
    every chain position at once), a larger one as the gate kernel and the chain kernel:
    the same device survivors through both forms give the oracle's results."""
    t0, _ = WL.reference_range(seed=0x9A, bits=89)
    o_surv, _ = oref.filter(t0, 3000, cap=1 << 20)
    n = len(o_surv)
    assert 100 < n <= 4096
    surv = np.zeros(n, ccs.SURVIVOR_DTYPE)
    surv["base_idx"], surv["k_offset"], surv["tile_idx"] = o_surv[:, 0], o_surv[:, 1], o_surv[:, 2]
    d_s, d_c = ctx.dalloc(n * 16), ctx.dalloc(64)
    ctx.h2d(d_s, surv)
    ctx.h2d(d_c, np.array([n], np.uint32))
    for bits in (89, 0):
        want = oref.prove(t0, o_surv, bits=bits)
        for cap in (n, 4096, 4097, 1 << 16):
            say(f"[prove] {n} survivors, launch cap {cap}, bits={bits}")
            d_r = ctx.dalloc(cap * 32)
            ref.prove_launch(t0, d_s, d_c, cap, d_r, bits=bits)
            ctx.synchronize()
            g = np.zeros(n, ccs.CHAIN_RESULT_DTYPE)
            ctx.d2h(g, d_r)
            ctx.dfree(d_r)
            assert list(g["survivor_idx"]) == list(range(n))
            compare_prove(g, want)
            assert np.any(g["flags"] & ccs.F_PRIME)
    ctx.dfree(d_s)
    ctx.dfree(d_c)


def test_prove_mr_base2_config1(ctx, oracle):
    s = WL.make_shape(ctx, "c1")
    o = oracle_shape(oracle, "c1")
    t0, tc, _ = WL.c1_range(seed=0x1C, t_count=1 << 10)
    o_surv, _ = o.filter(t0, tc, cap=1 << 20)
    surv = np.zeros(len(o_surv), ccs.SURVIVOR_DTYPE)
    surv["base_idx"], surv["k_offset"], surv["tile_idx"] = o_surv[:, 0], o_surv[:, 1], o_surv[:, 2]
    say(f"[prove] {len(surv)} config-1 survivors, MR base 2")
    g = s.prove(t0, surv, bits=60, test="mr", bases=(2,))
    compare_prove(g, o.prove(t0, o_surv, bits=60, kind=1, bases=(2,)))


def known_roots():
    import json
    import os
    from conftest import GOLDEN
    vec = json.load(open(os.path.join(GOLDEN, "reference_vectors.json")))
    roots = [(int(h, 16), L) for h, L in vec["v16_qvectors"]]
    roots += [(int(r), L) for r, L in vec["classic_chains"] if int(r) > 1000]
    roots.append((int(vec["cc13_hex"][0], 16), 13))
    with open(os.path.join(GOLDEN, "cc10plus_roots_sample_1000.txt")) as f:
        for i, line in enumerate(f):
            if i % 10 == 0:
                tag, hx, _ = line.split()
                roots.append((int(hx, 16), int(tag[2:])))
    return roots


def test_prove_known_chains_as_jobs(ctx, oracle):
    # class-form prove of arbitrary n: job (a = n mod W, u0 = n div W, count 1)
    s = WL.make_shape(ctx, "c4", classes=False)
    W = s.W
    roots = known_roots()
    root17 = 2759832934171386593519
    root18 = 0xb149165114f43bf4f72249
    members = [((root17 + 1) * 32 - 1, 17),   # chain position 5 of the CC17: walks back
               ((root18 + 1) * 2 ** 17 - 1, 18)]   # last member of a CC18: 17 steps back,
    cases = roots + members                      # past the 16 backward lanes
    jobs = [(n % W, n // W, 1) for n, _ in cases]
    surv = np.zeros(len(jobs), ccs.JOB_SURVIVOR_DTYPE)
    surv["job_idx"] = np.arange(len(jobs))
    say(f"[prove] {len(jobs)} known chains")
    g = s.prove_jobs(jobs, surv)
    for (n, L), r in zip(roots, g[: len(roots)]):
        assert r["chain_len"] == L and r["back_steps"] == 0, hex(n)
        assert (int(r["root_hi"]) << 64 | int(r["root_lo"])) == n
    m = g[len(roots)]
    assert m["chain_len"] == 17 and m["back_steps"] == 5 and m["flags"] & ccs.F_NONROOT
    assert (int(m["root_hi"]) << 64 | int(m["root_lo"])) == root17
    m = g[len(roots) + 1]
    assert m["chain_len"] == 18 and m["back_steps"] == 17 and m["flags"] & ccs.F_NONROOT
    assert (int(m["root_hi"]) << 64 | int(m["root_lo"])) == root18
    # the same numbers through the oracle
    for (n, _), r in zip(cases, g):
        f, root, ln, back = oracle.prove_one(n)
        assert (int(r["flags"]), int(r["chain_len"]), int(r["back_steps"])) == (f, ln, back)
    # and through a device launch whose cap (> 4096) selects the gate kernel and the
    # one-lane-per-chain walk
    js = ccs.JobSet(s, jobs)
    cap = 8192
    d_s, d_c, d_r = ctx.dalloc(cap * 8), ctx.dalloc(64), ctx.dalloc(cap * 32)
    ctx.h2d(d_s, surv)
    ctx.h2d(d_c, np.array([len(surv)], np.uint32))
    js.prove_launch(d_s, d_c, cap, d_r)
    ctx.synchronize()
    g2 = np.zeros(len(surv), ccs.CHAIN_RESULT_DTYPE)
    ctx.d2h(g2, d_r)
    for f in PROVE_FIELDS + ("survivor_idx",):
        assert np.array_equal(g2[f], g[f]), f
    js.close()
    for ptr in (d_s, d_c, d_r):
        ctx.dfree(ptr)


def test_prove_random_large_values_vs_oracle(ctx, oracle):
    # random odd values of 60..126 bits (primes among them), BPSW and MR-{2,3,5}
    rng = np.random.default_rng(3)
    s = WL.make_shape(ctx, "c4", classes=False)
    W = s.W
    vals = []
    import sympy
    for bits in (62, 64, 66, 90, 100, 120, 126):
        for _ in range(30):
            vals.append(int(rng.integers(1 << 62)) << (bits - 62) | 1 | (1 << (bits - 1)))
        for _ in range(10):
            vals.append(int(sympy.randprime(1 << (bits - 1), 1 << bits)))
    jobs = [(n % W, n // W, 1) for n in vals]
    surv = np.zeros(len(jobs), ccs.JOB_SURVIVOR_DTYPE)
    surv["job_idx"] = np.arange(len(jobs))
    for test, bases, kind in (("bpsw", (2,), 0), ("mr", (2, 3, 5), 1)):
        say(f"[prove] random values, {test}")
        g = s.prove_jobs(jobs, surv, test=test, bases=bases)
        for n, r in zip(vals, g):
            f, root, ln, back = oracle.prove_one(n, kind=kind, bases=bases)
            assert (int(r["flags"]), int(r["chain_len"]), int(r["back_steps"])) == (f, ln, back), hex(n)
            assert (int(r["root_hi"]) << 64 | int(r["root_lo"])) == root


# --- batched form ------------------------------------------------------------------

def test_batched_jobs_vs_oracle(ctx, oracle):
    jobs, W = WL.batched_jobs(n_jobs=24, seed=0xB4, lo_exp=10, hi_exp=20)
    s = WL.make_shape(ctx, "c4", classes=False, depth=10, target=18)
    o = oracle_shape(oracle, "c4", depth=10, target=18)
    assert s.W == W
    say(f"[batch] {len(jobs)} jobs, {sum(j[2] for j in jobs)} candidates")
    g, st = s.batch(jobs, cap=1 << 20)
    want = []
    alive = [0, 0, 0, 0]
    for ji, (a, u0, count) in enumerate(jobs):
        js, al = o.filter_class(a, W, u0, count, cuts=[st["dense_end"], st["tierb_end"], len(s.filter_primes)])
        want += [(ji, int(x)) for x in js]
        alive = [x + y for x, y in zip(alive, al)]
    got = sorted(zip(g["job_idx"].tolist(), g["u_off"].tolist()))
    assert st["candidates"] == alive[0]
    assert (st["alive_dense"], st["alive_tierb"], st["survivors"]) == tuple(alive[1:])
    assert got == sorted(want)
    assert len(got) > 50


# --- full-size properties (config 2, 2^36 candidates per launch) ---------------------

def test_config2_full_launch_properties(ctx, oracle):
    s = WL.make_shape(ctx, "c2")
    t0, tc, _ = WL.c2_range()
    say(f"[full] config 2 tiles={tc}")
    g1, st1 = s.filter(t0, tc)
    g2, st2 = s.filter(t0, tc)
    assert st1["candidates"] == s.count_candidates(t0, tc) >= 1 << 36
    # deterministic survivor set and counters
    assert sorted_keys(g1) == sorted_keys(g2)
    assert {k: st1[k] for k in ("alive_dense", "alive_tierb", "survivors")} == \
        {k: st2[k] for k in ("alive_dense", "alive_tierb", "survivors")}
    # every survivor satisfies the predicate (direct divisibility check on the CPU)
    o = oracle_shape(oracle, "c2")
    for rec in g1:
        n = s.n_of(t0, rec)
        assert o.first_kill_brute(n) == len(s.filter_primes)
        assert o.on_lattice(n)
    # survivor count against the analytic rate (position 0 excluded above 97)
    p = 1.0
    for q in s.filter_primes:
        lo = 0 if q <= 97 else 1
        p *= 1 - len({(pow(2, -i, q) - 1) % q for i in range(lo, 18)}) / q
    exp = st1["candidates"] * p
    assert abs(len(g1) - exp) < 6 * math.sqrt(exp) + 3
    # the dense-stage survivors also follow the analytic rate
    pd = 1.0
    for q in s.filter_primes[: st1["dense_end"]]:
        lo = 0 if q <= 97 else 1
        pd *= 1 - len({(pow(2, -i, q) - 1) % q for i in range(lo, 18)}) / q
    assert abs(st1["alive_dense"] / st1["candidates"] - pd) < 0.01 * pd


def test_pipelined_streams_match_sequential(ctx, ref):
    """Filter of batch k+1 on one stream overlapping the prove of batch k on another
    (context streams and events) gives each batch the results of a sequential run."""
    say("[plain] two-stream pipeline")
    cap, pcap = 1 << 18, 1 << 14
    ranges = [WL.reference_range(seed=0x700 + i, bits=89)[0] for i in range(4)]
    d_s, d_c, d_r = ctx.dalloc(cap * 16), ctx.dalloc(64), ctx.dalloc(pcap * 32)

    def read(d_surv, d_count, d_res):
        n = np.zeros(1, np.uint32)
        ctx.d2h(n, d_count)
        surv = np.zeros(int(n[0]), ccs.SURVIVOR_DTYPE)
        res = np.zeros(min(int(n[0]), pcap), ccs.CHAIN_RESULT_DTYPE)
        if len(surv):
            ctx.d2h(surv, d_surv)
            ctx.d2h(res, d_res)
        order = np.lexsort((surv["k_offset"], surv["base_idx"], surv["tile_idx"]))
        by_idx = {int(r["survivor_idx"]): r for r in res}
        return [(int(surv[i]["tile_idx"]), int(surv[i]["base_idx"]), int(surv[i]["k_offset"]),
                 int(by_idx[i]["flags"]), int(by_idx[i]["chain_len"])) for i in order]

    want = []
    for t0 in ranges:
        ref.filter_launch(t0, 1800, d_s, cap, d_c)
        ref.prove_launch(t0, d_s, d_c, pcap, d_r, bits=89)
        ctx.synchronize()
        want.append(read(d_s, d_c, d_r))
    sf, sp = ctx.stream(), ctx.stream()
    bufs = [(ctx.dalloc(cap * 16), ctx.dalloc(64), ctx.dalloc(pcap * 32)) for _ in range(2)]
    ev_f, ev_p = [ctx.event() for _ in range(2)], [ctx.event() for _ in range(2)]
    got = [None] * len(ranges)
    for k, t0 in enumerate(ranges):
        b = k % 2
        if k >= 2:
            ctx.wait(sf, ev_p[b])
            ctx.synchronize()
            got[k - 2] = read(*bufs[b])
        ref.filter_launch(t0, 1800, bufs[b][0], cap, bufs[b][1], stream=sf)
        ctx.record(ev_f[b], sf)
        ctx.wait(sp, ev_f[b])
        ref.prove_launch(t0, bufs[b][0], bufs[b][1], pcap, bufs[b][2], bits=89, stream=sp)
        ctx.record(ev_p[b], sp)
    ctx.synchronize()
    for k in range(len(ranges) - 2, len(ranges)):
        got[k] = read(*bufs[k % 2])
    assert got == want and all(len(w) > 0 for w in want)
    for e in ev_f + ev_p:
        ctx.event_destroy(e)
    ctx.stream_destroy(sf)
    ctx.stream_destroy(sp)


def test_concurrent_streams_keep_their_own_scratch(ctx, ref):
    """Filter and prove launches queued on four streams of one context with no ordering
    between them (each stream has its own stage-B buffer, counters and prove queue in the
    library): every batch equals its sequential run, and the synchronous host form on the
    context's own stream still reports its statistics."""
    say("[plain] four concurrent streams")
    cap, pcap = 1 << 18, 1 << 14
    ranges = [WL.reference_range(seed=0x900 + i, bits=89)[0] for i in range(4)]
    want = []
    for t0 in ranges:
        surv, st = ref.filter(t0, 30000, cap=cap)
        want.append(sorted(zip(surv["tile_idx"].tolist(), surv["base_idx"].tolist(), surv["k_offset"].tolist())))
    streams = [ctx.stream() for _ in ranges]
    bufs = [(ctx.dalloc(cap * 16), ctx.dalloc(64), ctx.dalloc(pcap * 32)) for _ in ranges]
    for t0, st, (d_s, d_c, d_r) in zip(ranges, streams, bufs):
        ref.filter_launch(t0, 30000, d_s, cap, d_c, stream=st)
        ref.prove_launch(t0, d_s, d_c, pcap, d_r, bits=89, stream=st)
    ctx.synchronize()
    for k, (d_s, d_c, d_r) in enumerate(bufs):
        n = np.zeros(1, np.uint32)
        ctx.d2h(n, d_c)
        surv = np.zeros(int(n[0]), ccs.SURVIVOR_DTYPE)
        res = np.zeros(int(n[0]), ccs.CHAIN_RESULT_DTYPE)
        if len(surv):
            ctx.d2h(surv, d_s)
            ctx.d2h(res, d_r)
        got = sorted(zip(surv["tile_idx"].tolist(), surv["base_idx"].tolist(), surv["k_offset"].tolist()))
        assert got == want[k] and len(got) > 50
        assert sorted(res["survivor_idx"].tolist()) == list(range(len(surv)))
    _, st = ref.filter(ranges[0], 30000, cap=cap)
    assert st["survivors"] == len(want[0]) and st["alive_dense"] > st["alive_tierb"] > 0
    for st in streams:
        ctx.stream_destroy(st)
    for b in bufs:
        for p in b:
            ctx.dfree(p)


def test_host_confirm_equals_gpu_prove(ctx, ref):
    """The host confirmation engine (CPU threads) and the GPU prove kernels give the same
    result for every survivor of a reference-shape range, and the same stats."""
    say("[prove] host confirm vs GPU prove")
    t0, _ = WL.reference_range(seed=0x3C, bits=89)
    surv, _ = ref.filter(t0, 6000)
    assert len(surv) > 500
    for bits in (89, 0):
        g = ref.prove(t0, surv, bits=bits)
        h, st = ref.confirm_host(t0, surv, bits=bits)
        assert np.array_equal(g, h)
        assert st == ccs.chain_stats(g) and st["primes"] > 0


def test_chain_roots_are_the_v13_k2_output(ctx, oracle, ref, oref):
    """ccs_chain_roots_launch gives the v13 K2 records (v13.cu:1462-1538): one
    {base_idx, k_offset, tile_idx, chain_len} per prime survivor whose (n-1)/2 is not prime,
    and the count of prime survivors, from the device survivors and prove results"""
    t0, _ = WL.reference_range(seed=0x9B, bits=89)
    cap, pcap = 1 << 16, 4096
    d_s, d_c, d_r = ctx.dalloc(cap * 16), ctx.dalloc(64), ctx.dalloc(pcap * 32)
    d_o, d_n = ctx.dalloc(pcap * 16), ctx.dalloc(64)
    try:
        ref.filter_launch(t0, 6000, d_s, cap, d_c)
        ref.prove_launch(t0, d_s, d_c, pcap, d_r, bits=89)
        ref.chain_roots_launch(d_s, d_r, d_c, pcap, d_o, pcap, d_n, ctypes_ptr(d_n, 4))
        ctx.synchronize()
        cnt = np.zeros(2, np.uint32)
        ctx.d2h(cnt, d_n)
        n = np.zeros(1, np.uint32)
        ctx.d2h(n, d_c)
        surv = np.zeros(n[0], ccs.SURVIVOR_DTYPE)
        ctx.d2h(surv, d_s)
        roots = np.zeros(cnt[0], ccs._lib.CHAIN_ROOT_DTYPE)
        ctx.d2h(roots, d_o)
        # a launch whose output holds fewer roots still reports their true count
        ref.chain_roots_launch(d_s, d_r, d_c, pcap, d_o, 2, d_n, None)
        ctx.synchronize()
        small = np.zeros(1, np.uint32)
        ctx.d2h(small, d_n)
    finally:
        for ptr in (d_s, d_c, d_r, d_o, d_n):
            ctx.dfree(ptr)
    o_surv = np.stack([surv["base_idx"], surv["k_offset"], surv["tile_idx"], surv["_pad"]], 1)
    want_res = oref.prove(t0, o_surv, bits=89)
    prime = (want_res[:, 0] & oracle.F_PRIME) != 0
    root = prime & (want_res[:, 2] == 0)
    want = sorted((int(s[0]), int(s[1]), int(s[2]), int(r[1])) for s, r, ok in zip(o_surv, want_res, root) if ok)
    got = sorted(zip(roots["base_idx"].tolist(), roots["k_offset"].tolist(), roots["tile_idx"].tolist(),
                     roots["chain_len"].tolist()))
    say(f"[roots] {len(surv)} survivors, {int(prime.sum())} prime, {len(want)} roots")
    assert len(want) > 10 and got == want
    assert cnt[1] == int(prime.sum())
    assert small[0] == len(want)


def ctypes_ptr(p, offset):
    """device pointer p + offset bytes"""
    import ctypes
    return ctypes.c_void_p(p.value + offset)
