#!/usr/bin/env python3
"""Benchmark: filtered candidates/s of the plain GPU form at CC18 depth (config 2:
M = 31#, depth 18, 123 552 classes, filter stages 37..2^10 / 2^10..2^14 / 2^14..2^16,
2^36 candidates per launch from a seeded start in [2^86, 2^89), Baillie-PSW prove of the
survivors).  One step = one filter launch over the whole batch + the prove launch over its
survivors, inputs resident on the GPU.  Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "cunningham-chain-search_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

# published campaign baseline (RTX 4090/5090, v13 3-stage filter, depth 18-20): 57-65 G/s
PUBLISHED_GPU_CAND_PER_S = 65e9

# HBM bytes per sieve_kernel launch of the config-2 step at the default tuning, from two
# rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE, KiB per dispatch), each scaled by what the
# counter reads for a 1 GiB device copy (bench.py --copy-calibration): FETCH_SIZE reads
# 524308 KiB for 1 GiB fetched (x2.0), WRITE_SIZE 1048576 KiB for 1 GiB written (x1.0).
SIEVE_TRAFFIC = {"c2": {"fetch_kib": 2065.25, "write_kib": 1879.53, "fetch_scale": 2.0, "write_scale": 1.0,
                        "profile": "profiles/sieve_kernel_cc18_31primorial_pmc.md"}}
# hardware utilization of the sieve kernel from the same counter profile (SQ counters over
# 256 CUs x kernel cycles): wave64 VALU issue, LDS port bytes against 128 B/clk per CU,
# LDS bank-conflict cycles
SIEVE_ISSUE = {"c2": {"valu_issue_frac": 0.57, "lds_port_frac": 0.76, "lds_bank_conflict_frac": 0.068,
                      "profile": "profiles/sieve_kernel_cc18_31primorial_pmc.md"}}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(name, t0, seconds):
    """The oracle's CPU restatement of the reference CPU path (the v34-style 64-candidate
    bit-vector filter, per-candidate for the larger primes), all host threads (<= 16), on a
    bounded sample of the same workload: consecutive tiles from the same start."""
    from oracle import pyoracle
    from ccs import workloads as WL
    p = WL.shape_params(name)
    o = pyoracle.Shape(p["crt"], p["wheel"], p["target"], p["depth"], p["filter_primes"], p["pos0_max_q"])
    nt = cpu_threads()
    o.filter_bv(t0, 6, cap=1 << 16, nthreads=1)        # one-time table build (startup work)
    tiles = 6 * nt * 50
    while True:                                         # calibrate on >= 1 s of work
        t = time.perf_counter()
        o.filter_bv(t0, tiles, cap=1 << 22, nthreads=nt)
        dt = time.perf_counter() - t
        if dt >= 1.0 or tiles >= 6 * 10 ** 8:
            break
        tiles *= 4
    tiles = int(min(6 * 10 ** 8, tiles * seconds / max(dt, 1e-3)))
    tiles -= tiles % 6
    t = time.perf_counter()
    surv, cand = o.filter_bv(t0, tiles, cap=1 << 22, nthreads=nt)
    dt = time.perf_counter() - t
    model = "cpu"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": cand / dt, "unit": "candidates/s", "cores": nt, "kind": "port",
            "sample": f"{tiles} consecutive tiles = {cand} candidates of the same shape and start, "
                      f"{dt:.1f} s on {nt} threads ({model}); v34-style bit-vector filter restated "
                      f"in C, sieve only (on the reference shape it runs at 2.86 G/s on 8 Xeon "
                      f"threads where the reference v34 binary ran at 4.07 G/s)"}


def measure_batched(ccs, WL, ctx, n_jobs=4096, tune_trials=40):
    """Config 4: n_jobs random admissible classes mod 37# at depth 18, log-uniform lengths
    in [2^20, 2^32]; one persistent work-stealing launch (and its HIP graph) against the
    plain per-shape launch loop over the same jobs and tuning."""
    s = WL.make_shape(ctx, "c4", classes=False)
    jobs, W = WL.batched_jobs(n_jobs=n_jobs)
    sample = [(a, u0, min(c, 1 << 24)) for a, u0, c in jobs[:512]]
    tuned, _ = s.autotune_jobs(sample, max_trials=tune_trials)
    js = ccs.JobSet(s, jobs)
    cap = 1 << 20
    d_out, d_cnt = ctx.dalloc(cap * 8), ctx.dalloc(64)

    def best_of(fn, reps=2):
        fn()
        ctx.synchronize()
        best = 1e30
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ctx.synchronize()
            best = min(best, time.perf_counter() - t)
        return best

    tb = best_of(lambda: js.launch(d_out, cap, d_cnt))
    nb = js.stats(d_cnt)["survivors"]
    tp = best_of(lambda: js.launch_each(d_out, cap, d_cnt))
    n_p = js.stats(d_cnt)["survivors"]
    g = js.graph(d_out, cap, d_cnt)
    tg = best_of(lambda: js.graph_launch(g))
    ng = js.stats(d_cnt)["survivors"]
    js.close()
    ctx.dfree(d_out)
    ctx.dfree(d_cnt)
    return {"workload": "config 4: 4096 admissible classes mod 37#, depth 18, lengths log-uniform in "
                        "[2^20, 2^32], candidates in [2^84, 2^90), filter primes 41..65521",
            "jobs": n_jobs, "candidates": js.candidates, "batched_s": tb, "graph_s": tg, "plain_loop_s": tp,
            "batched_candidates_per_s": js.candidates / tb, "graph_candidates_per_s": js.candidates / tg,
            "plain_loop_candidates_per_s": js.candidates / tp, "speedup_batched_vs_plain": tp / tb,
            "speedup_graph_vs_plain": tp / tg, "survivors_equal": nb == n_p == ng, "survivors": nb,
            "tuning": tuned}


def measure_served(ccs, WL, ctx, n_jobs=10000, load=0.8, tune_trials=16, slice_candidates=1 << 37,
                   max_slice_survivors=1 << 18):
    """Config 5: a Poisson stream of search jobs (depth U{12..20}, width U{64,80,96}, random
    admissible class mod 31#, length log-uniform in [2^24, 2^34]) at `load` times the server's
    measured capacity; p50 / p99 latency from submission to completion.  Each depth's shape
    is autotuned on its own sample jobs (the stage cutoffs that suit depth 12 and depth 20
    differ)."""
    import numpy as np
    shapes = {d: WL.make_shape(ctx, f"c5d{d}", classes=False) for d in range(12, 21)}
    tune_pool = WL.served_requests(n_jobs=2000, seed=0x7E)
    tuned = {}
    for d, sh in shapes.items():
        sample = [(r["a"], r["u0"], min(r["count"], 1 << 27)) for r in tune_pool if r["depth"] == d][:192]
        tuned[d], _ = sh.autotune_jobs(sample, max_trials=tune_trials)
        if tuned[d]["rows_per_group"] != 2:   # one row layout lets the server launch all depths at once
            try:
                sh.set_tuning(rows_per_group=2)
                tuned[d] = sh.tuning
            except ccs.CcsError:
                pass
    cal = WL.served_requests(n_jobs=400, seed=0xCA1)
    # capacity: the calibration jobs drained twice by one server, the second pass timed (the
    # plain loop below is timed on its second pass as well)
    srv = ccs.Server(ctx, shapes, slice_candidates=slice_candidates, max_slice_survivors=max_slice_survivors)
    for timed in (False, True):
        t = time.perf_counter()
        for r in cal:
            srv.submit(r["depth"], r["a"], r["u0"], r["count"])
        srv.drain()
        if timed:
            capacity = sum(r["count"] for r in cal) / (time.perf_counter() - t)
        while srv.poll():
            pass
    srv.shutdown()
    # the plain comparator: the same calibration jobs with one filter launch per job (each
    # with the unit size it would have alone; jobs above 2^32 candidates, the most one
    # launch takes per job, in pieces of 2^32), then one prove launch per depth
    cap = 1 << 22
    d_out, d_cnt, d_res = ctx.dalloc(cap * 8), ctx.dalloc(64), ctx.dalloc(cap * 32)
    by_depth = {}
    for r in cal:
        u0, left = r["u0"], r["count"]
        while left > 0:
            n = min(left, 1 << 32)
            by_depth.setdefault(r["depth"], []).append((r["a"], u0, n))
            u0 += n
            left -= n
    plain_s = 0.0
    for d, jl in by_depth.items():
        js = ccs.JobSet(shapes[d], jl)
        js.launch_each(d_out, cap, d_cnt)      # warm-up (and the per-job unit sizes)
        ctx.synchronize()
        t = time.perf_counter()
        js.launch_each(d_out, cap, d_cnt)
        js.prove_launch(d_out, d_cnt, cap, d_res)
        ctx.synchronize()
        plain_s += time.perf_counter() - t
        js.close()
    for ptr in (d_out, d_cnt, d_res):
        ctx.dfree(ptr)
    plain = sum(r["count"] for r in cal) / plain_s
    srv = ccs.Server(ctx, shapes, slice_candidates=slice_candidates, max_slice_survivors=max_slice_survivors)
    reqs = WL.served_requests(n_jobs=n_jobs)
    mean = sum(r["count"] for r in reqs) / len(reqs)
    lam = load * capacity / mean
    gaps = np.random.default_rng(0xC5).exponential(1.0 / lam, size=len(reqs))
    t0 = time.perf_counter()
    tn = t0
    for r, gap in zip(reqs, gaps):
        tn += gap
        d = tn - time.perf_counter()
        if d > 0:
            time.sleep(d)
        srv.submit(r["depth"], r["a"], r["u0"], r["count"], tag=r["id"], bits=r["bits"])
    srv.drain()
    wall = time.perf_counter() - t0
    done = []
    while True:
        got = srv.poll()
        if not got:
            break
        done += got
    st = srv.stats()
    srv.shutdown()
    lat = np.array([d["t_done"] - d["t_submit"] for d in done])
    total = sum(r["count"] for r in reqs)
    return {"workload": "config 5: Poisson stream, depth U{12..20}, width U{64,80,96}, classes mod 31#, "
                        "lengths log-uniform in [2^24, 2^34]",
            "jobs": len(done), "load": load, "arrival_rate_per_s": lam, "capacity_candidates_per_s": capacity,
            "plain_loop_candidates_per_s": plain, "speedup_capacity_vs_plain": capacity / plain,
            "aggregate_candidates_per_s": total / wall, "wall_s": wall,
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "slices": st["slices"], "launches": st["launches"],
            "slice_candidates": slice_candidates, "tuning": {str(d): t for d, t in tuned.items()}}


def measure_families(ccs, ctx, depth=20, q_per_family=1 << 28, exp_start=1, bits=(90, 127)):
    """The v16 Q-iteration workload: the reference's pool of the 100 most frequent immune
    fingerprints (known_fingerprints.txt), p_0 = Q*D*2^E - 1, each family's Q window for
    90..127-bit p_0, sieved at depth 20 by the primes 3..1009 (v16's GPU primes to 503 and
    its CPU ext sieve 509..1009).  One group launch over all families against one launch
    per family (v16's per-variant dispatch); the unit is (Q, V) pairs per second."""
    import random
    pool = ccs.pool_load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "pools",
                                      "known_fingerprints.txt"))
    primes = [q for q in range(3, 1010) if all(q % d for d in range(2, int(q ** 0.5) + 1))]
    rng = random.Random(0x16)
    shapes, jobs = [], []
    for k, e in enumerate(pool):
        sh = ccs.FamilyShape(ctx, e["D"], exp_start, depth, depth, primes)
        qmin, qmax = sh.window(*bits)
        shapes.append(sh)
        left, q0 = q_per_family, rng.randrange(qmin, qmax - q_per_family)
        while left:
            c = min(left, 1 << 32)
            jobs.append((k, (0, q0, c)))
            q0 += c
            left -= c
    total = q_per_family * len(shapes)
    cap = 1 << 22
    d_out, d_cnt = ctx.dalloc(cap * 8), ctx.dalloc(64)

    def best_of(fn, reps=3):
        fn()
        ctx.synchronize()
        best = 1e30
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ctx.synchronize()
            best = min(best, time.perf_counter() - t)
        return best

    gs = ccs.GroupSet(shapes, jobs)
    tg = best_of(lambda: gs.launch(d_out, cap, d_cnt))
    survivors = gs.stats(d_cnt)["survivors"]
    sets = [ccs.JobSet(shapes[k], [j for kk, j in jobs if kk == k]) for k in range(len(shapes))]
    tl = best_of(lambda: [js.launch(d_out, cap, d_cnt) for js in sets])
    for js in sets:
        js.close()
    gs.close()
    ctx.dfree(d_out)
    ctx.dfree(d_cnt)
    for sh in shapes:
        sh.close()
    return {"workload": f"v16 Q-iteration pool: {len(pool)} fingerprint families (known_fingerprints.txt), "
                        f"p_0 = Q*D*2^{exp_start} - 1 in {bits[0]}..{bits[1]} bits, depth {depth}, primes 3..1009, "
                        f"2^{q_per_family.bit_length() - 1} Q per family",
            "families": len(pool), "pairs": total, "group_s": tg, "per_family_loop_s": tl,
            "group_pairs_per_s": total / tg, "per_family_loop_pairs_per_s": total / tl,
            "survivors": survivors}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--config", default="c2", choices=["c2", "reference", "c3"])
    ap.add_argument("--tune", action="store_true",
                    help="run the autotuner on the same range first (default: the shape's default tuning)")
    ap.add_argument("--no-tune", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-extras", action="store_true", help="skip the batched and served measurements")
    ap.add_argument("--served-jobs", type=int, default=10000)
    ap.add_argument("--tune-trials", type=int, default=40)
    ap.add_argument("--tuning", default=None, help="JSON tuning to use instead of the autotuner")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device (scales rocprofv3's FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 28, device="cuda")      # 1 GiB of float32
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 4}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this pipeline shards by disjoint tile ranges with no "
              "collective, and multi-GPU runs are not part of this build (use --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import ccs
    from ccs import workloads as WL

    ctx = ccs.Context(0)
    name = args.config
    shape = WL.make_shape(ctx, name)
    if name == "c2":
        t0, tc, n_start = WL.c2_range()
        workload = ("config 2: M = 31# (CRT 5..19 x wheel 23,29,31 x mod 6), depth 18, 123552 classes, "
                    "filter primes 37..65521 in stages 37..1021 / 1031..16381 / 16411..65521, "
                    "2^36 candidates per launch from a seeded start in [2^86, 2^89), BPSW prove")
    elif name == "c3":
        t0, tc, n_start = WL.c3_range()
        workload = "config 3: M = 37#, depth 19, filter primes to 2^18, 2^38 candidates per launch"
    else:
        t0, tc = WL.reference_range(bits=89, n_tiles=3337446)
        n_start = t0 * shape.WP * shape.M
        workload = "v13 campaign shape (primes 37..863), 89-bit, 3337446 tiles"
    bits = n_start.bit_length()
    cand = shape.count_candidates(t0, tc)
    # algorithmic work: E[residue tests] per candidate under ascending-prime early exit
    # (also the analytic survival, which sizes the prove)
    fs = []
    for q in shape.filter_primes:
        lo = 0 if q <= shape.pos0_max_q else 1
        fs.append((q, len(WL.forbidden(q, lo, shape.depth))))
    e_tests, surv_frac = WL.expected_tests(fs)
    # prove cap 4096 when the step's survivors fit (config 2: ~190), so the prove of a step
    # is one kernel; the v13 campaign shape (~8·10^5 per 2^36) gets room for all of them
    expect = cand * surv_frac
    cap, pcap = 1 << 20, 1 << 12
    if 3 * expect > pcap:
        pcap = min(1 << 21, 1 << int(3 * expect).bit_length())
    # The prove of step k (stream B) overlaps the filters of the next steps (stream A), as
    # the reference's prover threads consume one batch while the GPU filters the next.
    # Three survivor buffers: filter k reuses the buffer of step k-3 once that step's
    # prove has read it, so a prove whose tail outlasts the next filter stalls nothing.
    nbuf = 3
    d_surv = [ctx.dalloc(cap * 16) for _ in range(nbuf)]
    d_count = [ctx.dalloc(64) for _ in range(nbuf)]
    d_res = [ctx.dalloc(pcap * 32) for _ in range(nbuf)]
    s_filter, s_prove = ctx.stream(), ctx.stream()
    ev_filtered = [ctx.event() for _ in range(nbuf)]
    ev_proved = [ctx.event() for _ in range(nbuf)]
    n_step = [0]

    def step():
        k = n_step[0]
        b, sf = k % nbuf, s_filter
        if k >= nbuf:
            ctx.wait(sf, ev_proved[b])            # prove k-3 has read buffer b
        shape.filter_launch(t0, tc, d_surv[b], cap, d_count[b], stream=sf)
        ctx.record(ev_filtered[b], sf)
        ctx.wait(s_prove, ev_filtered[b])
        shape.prove_launch(t0, d_surv[b], d_count[b], pcap, d_res[b], bits=bits, stream=s_prove)
        ctx.record(ev_proved[b], s_prove)
        n_step[0] = k + 1

    if args.tuning:
        shape.set_tuning(**json.loads(args.tuning))
    elif args.tune:
        # autotuner on the same range (setup work, not timed): the unit size it picks
        # depends on the run length per class, so a shorter sample would mislead it
        t_tune = time.perf_counter()
        tuned, ms = shape.autotune(t0, tc, max_trials=args.tune_trials)
        log(f"[bench] autotune {time.perf_counter() - t_tune:.1f}s -> {tuned} ({ms:.2f} ms on the sample)")
    log(f"[bench] {name}: tiles={tc} candidates/step={cand} tuning={shape.tuning}")
    # ctx.synchronize() is a device-wide hipDeviceSynchronize: both streams are drained
    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    ctx.profile_start()
    t = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    dt = time.perf_counter() - t
    kt = ctx.profile_stop()
    last = (n_step[0] - 1) % nbuf
    n_surv = np.zeros(1, np.uint32)
    ctx.d2h(n_surv, d_count[last])
    res = np.zeros(min(int(n_surv[0]), pcap), ccs.CHAIN_RESULT_DTYPE)
    if len(res):
        ctx.d2h(res, d_res[last])
    value = cand * args.steps / dt
    ms_sieve = kt["ms_sieve"] / max(kt["n_sieve"], 1)
    # 4 int32 VALU ops per residue test, against the int32 VALU peak
    tr = SIEVE_TRAFFIC.get(name)
    traffic = None
    if tr and tr["fetch_kib"] is not None:
        traffic = (tr["fetch_kib"] * tr["fetch_scale"] + tr["write_kib"] * tr["write_scale"]) * 1024
    ops = cand * e_tests * WL.INT32_OPS_PER_TEST
    achieved = ops / (ms_sieve / 1e3) / 1e12
    out = {
        "metric": "filtered candidates/s at CC18 depth (bit-exact)",
        "value": value,
        "unit": "candidates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / PUBLISHED_GPU_CAND_PER_S,
        "dtype": "int32 (filter), u128 (prove)",
        "data": "synthetic: seeded tile range (xorshift64), no dataset",
        "config": {"workload": workload, "tile_start": str(t0), "tile_count": tc,
                   "candidates_per_step": cand, "tuning": shape.tuning},
        "roofline": {"bound": "valu_int32", "achieved": achieved, "peak": WL.INT32_PEAK_TOPS,
                     "unit": "Tops/s (int32 VALU)", "frac": achieved / WL.INT32_PEAK_TOPS,
                     "traffic": traffic, "kernel": "sieve_kernel",
                     "kernel_ms": ms_sieve, "e_tests_per_candidate": e_tests,
                     "ops_per_launch": ops, "hardware": SIEVE_ISSUE.get(name)},
        "kernels_ms": {"sieve": ms_sieve, "deep": kt["ms_deep"] / max(kt["n_deep"], 1),
                       "prove": kt["ms_prove"] / max(kt["n_prove"], 1)},
        "survivors_per_step": int(n_surv[0]),
        "primes_per_step": int(((res["flags"] & ccs.F_PRIME) != 0).sum()),
        "analytic_survival": surv_frac,
    }
    if not args.no_extras:
        log("[bench] batched vs plain loop ...")
        out["batched"] = measure_batched(ccs, WL, ctx)
        log("[bench] served ...")
        out["served"] = measure_served(ccs, WL, ctx, n_jobs=args.served_jobs)
        log("[bench] fingerprint families ...")
        out["families"] = measure_families(ccs, ctx)
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(name, t0, args.cpu_seconds)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
