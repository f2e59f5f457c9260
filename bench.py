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

# Counter profile of the config-2 step's sieve_kernel at the default tuning (rocprofv3
# --pmc, one run per counter set, values per dispatch; the summary with its commands is
# profiles/sieve_kernel_cc18_31primorial_pmc.md).  HBM traffic: FETCH_SIZE and WRITE_SIZE
# (KiB per dispatch), each scaled by what the counter reads for a 1 GiB device copy
# (bench.py --copy-calibration): FETCH_SIZE reads 524308 KiB for 1 GiB fetched (x2.0),
# WRITE_SIZE 1048576 KiB for 1 GiB written (x1.0).  SQ counters: wave64 VALU and LDS
# instructions and LDS bank-conflict cycles.
SIEVE_PROFILE = {"c2": {"tuning": {"n_dense": 41, "n_tierb": 320, "rows_per_group": 4, "unit_j": 1048576},
                        "fetch_kib": 7776.08, "write_kib": 13803.1, "fetch_scale": 2.0, "write_scale": 1.0,
                        "sq_insts_valu": 1.71303e9, "sq_insts_lds": 3.63146e8, "sq_lds_bank_conflict": 1.38853e8,
                        "profile": "profiles/sieve_kernel_cc18_31primorial_pmc.md"}}
LDS_PROBE = os.path.join(PKG, "lds_probe")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """The host CPU as this process sees it: nproc, the affinity set, the thread share a
    GPU job is given (OMP_NUM_THREADS, set on the GPU host to the per-GPU share), and lscpu's
    model, sockets, cores per socket and threads per core."""
    import subprocess
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    info = {"nproc": os.cpu_count(), "affinity": aff, "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except (OSError, subprocess.TimeoutExpired):
        pass
    share = aff
    if info["omp_num_threads"] and info["omp_num_threads"].isdigit():
        share = min(aff, int(info["omp_num_threads"]))
    info["threads_used"] = max(1, share)
    return info


def time_port(o, t0, tiles_min, seconds, nthreads, engine, full, kind, bits):
    """One row of the port matrix: the oracle's v33-structured per-candidate filter
    ('v33') or v34-structured bit-vector filter ('v34') over consecutive tiles from t0,
    sized to about `seconds` by a short calibration; full = the prove of every survivor
    (the GMP-engine semantics, BPSW or the config's MR) included in the time."""
    filt = o.filter if engine == "v33" else o.filter_bv

    def run(tiles):
        t = time.perf_counter()
        surv, c = filt(t0, tiles, cap=1 << 22, nthreads=nthreads)
        cand = c[0] if engine == "v33" else c
        primes = 0
        if full and len(surv):
            res = o.prove(t0, surv, bits=bits, kind=kind, nthreads=nthreads)
            primes = int(((res[:, 0] & 4) != 0).sum())
        return time.perf_counter() - t, cand, len(surv), primes

    tiles = tiles_min
    while True:
        dt, *_ = run(tiles)
        if dt >= 0.25 or tiles >= 6 * 10 ** 8:
            break
        tiles *= 4
    tiles = max(tiles_min, int(min(6 * 10 ** 8, tiles * seconds / max(dt, 1e-3))))
    tiles -= tiles % 6
    dt, cand, ns, primes = run(tiles)
    return {"value": cand / dt, "candidates": cand, "seconds": dt, "tiles": tiles, "survivors": ns,
            "primes": primes}


def time_reference(binary, seconds, nthreads, full, bits=89, target=18):
    """One row of the reference matrix: the reference engine itself (oracle/_ref, built from
    the reference's C sources) on its own shape, a sequential sweep from the start of the
    --bits range, stopped by SIGINT after `seconds` (it finishes the batch in hand and
    prints its totals); the rate is its own wheel candidates over its own run time."""
    import re
    import signal
    import subprocess
    cmd = [binary, "--prefix", "0b1", "--sequential", "--bits", str(bits), "--target", str(target),
           "--threads", str(nthreads), "--report", "60", "--log", str(target)]
    if not full:
        cmd.append("--sieve-only")
    p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    time.sleep(seconds)
    p.send_signal(signal.SIGINT)
    out, _ = p.communicate(timeout=120)
    w = re.search(r"Wheel candidates \(L1\):\s+(\d+)", out)
    t = re.search(r"Time:\s+([\d.]+) seconds", out)
    pr = re.search(r"Primes found:\s+(\d+)", out)
    if not (w and t) or float(t.group(1)) <= 0:
        return None
    return {"value": int(w.group(1)) / float(t.group(1)), "candidates": int(w.group(1)),
            "seconds": float(t.group(1)), "primes": int(pr.group(1)) if pr else 0}


def cpu_baseline(seconds=2.0, configs=("reference", "c1", "c2", "c3")):
    """The CPU path on this host, every thread of the job's share: the reference's own v33
    and v34 engines (oracle/_ref, when built) on the reference shape, and the oracle's
    v33- and v34-structured restatements on each config; sieve-only and full (with the
    prove of every survivor).  The headline is the fastest sieve-only row."""
    from oracle import pyoracle
    from ccs import workloads as WL
    host = host_cpu()
    nt = host["threads_used"]
    rows = []
    ref_dir = os.path.join(ROOT, "oracle", "_ref")
    for eng, binary in (("v34", "cc_gmp_v34"), ("v33", "cc_gmp_v33")):
        path = os.path.join(ref_dir, binary)
        if not os.access(path, os.X_OK):
            continue
        for full in (False, True):
            r = time_reference(path, seconds + 1.0, nt, full)
            if r:
                rows.append(dict(r, kind="reference", engine=eng, config="reference", mode="full" if full else "sieve",
                                 threads=nt))
    starts = {"reference": WL.reference_range(bits=89)[0], "c1": WL.c1_range()[0], "c2": WL.c2_range()[0],
              "c3": WL.c3_range()[0]}
    for name in configs:
        p = WL.shape_params(name)
        o = pyoracle.Shape(p["crt"], p["wheel"], p["target"], p["depth"], p["filter_primes"], p["pos0_max_q"])
        o.filter_bv(starts[name], 6, cap=1 << 16, nthreads=1)      # table build (startup work)
        kind, bits = (pyoracle.MR, 60) if name == "c1" else (pyoracle.BPSW, 0)
        for eng in ("v34", "v33"):
            for full in (False, True):
                r = time_port(o, starts[name], 6 * nt, seconds, nt, eng, full, kind, bits)
                rows.append(dict(r, kind="port", engine=eng, config=name, mode="full" if full else "sieve",
                                 threads=nt))
    best = max((r for r in rows if r["mode"] == "sieve"), key=lambda r: r["value"])
    gap = None
    pr = [r for r in rows if r["config"] == "reference" and r["mode"] == "sieve" and r["engine"] == "v34"]
    if len(pr) == 2:
        gap = [r for r in pr if r["kind"] == "port"][0]["value"] / [r for r in pr if r["kind"] == "reference"][0]["value"]
    return {"value": best["value"], "unit": "candidates/s", "cores": nt, "kind": best["kind"],
            "sample": f"fastest sieve-only row: {best['engine']} {best['kind']} on {best['config']}, "
                      f"{best['candidates']} candidates in {best['seconds']:.1f} s on {nt} threads "
                      f"({host.get('model', 'cpu')}); every row below ~{seconds:.0f} s",
            "host": host, "port_over_reference_v34": gap, "matrix": rows}


def measure_batched(ccs, WL, ctx, n_jobs=4096, tune_trials=40):
    """Config 4: n_jobs random admissible classes mod 37# at depth 18, log-uniform lengths
    in [2^20, 2^32]; one persistent work-stealing launch (and its HIP graph) against the
    plain per-shape launch loop over the same jobs and tuning."""
    s = WL.make_shape(ctx, "c4", classes=False)
    jobs, W = WL.batched_jobs(n_jobs=n_jobs)
    sample = jobs[:512]   # an eighth of the set, lengths as drawn: the unit size suits the set
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


def tune_rows_fixed(ccs, ctx, sh, sample, rows=2):
    """A shape's tuning with its row layout fixed (the server launches the depth profiles of a
    launch group together, so they share one): coordinate descent from the autotuner's
    pick, timed on the job-set launch over the sample jobs.  The dense cutoff that suits 4
    rows does not suit 2 (the patterns' LDS extension halves)."""
    cap = 1 << 22
    d_out, d_cnt = ctx.dalloc(cap * 8), ctx.dalloc(64)

    def timed(t):
        sh.set_tuning(**t)
        js = ccs.JobSet(sh, sample)
        js.launch(d_out, cap, d_cnt)
        ctx.synchronize()
        best = 1e30
        for _ in range(2):
            t0 = time.perf_counter()
            js.launch(d_out, cap, d_cnt)
            ctx.synchronize()
            best = min(best, time.perf_counter() - t0)
        js.close()
        return best

    best = dict(sh.tuning, rows_per_group=rows)
    try:
        tb = timed(best)
    except ccs.CcsError:   # the autotuner's pick does not fit at this row layout
        best = dict(best, pair_max=16384)
        tb = timed(best)
    fp = sh.filter_primes
    knobs = (("n_dense", [sum(1 for q in fp if q < c) for c in (128, 200, 256, 320, 400, 512)]),
             ("pair_max", [0, 4096, 16384, 65536]), ("n_dense", None),
             ("n_tierb", [256, 512, 768, 1024]), ("unit_j", [65536, 131072, 262144, 1048576]))
    last_dense = None
    for knob, values in knobs:
        if values is None:            # the dense cutoff again, after the pattern period
            knob, values = last_dense
        if knob == "n_dense":
            last_dense = (knob, values)
        for v in values:
            if v == best[knob]:
                continue
            t = dict(best, **{knob: v})
            try:
                tt = timed(t)
            except ccs.CcsError:
                continue
            if tt < tb:
                best, tb = t, tt
    sh.set_tuning(**best)
    ctx.dfree(d_out)
    ctx.dfree(d_cnt)
    return sh.tuning


def measure_served(ccs, WL, ctx, n_jobs=10000, load=0.8, tune_trials=16, slice_candidates=1 << 38,
                   max_slice_survivors=1 << 19, reps=3):
    """Config 5: a Poisson stream of search jobs (depth U{12..20}, width U{64,80,96}, random
    admissible class mod 31#, length log-uniform in [2^24, 2^34]).  Capacity is measured on
    the stream's own jobs: all of them submitted at once and drained by one server, against
    the plain loop over the same jobs (one filter launch per job, one prove launch per
    depth), `reps` times each, alternating, for the spread of the ratio.  Then the same jobs
    arrive as a Poisson stream at `load` times the median capacity: p50 / p99 latency from
    submission to completion.  Each depth's shape is autotuned on its own sample jobs (the
    stage cutoffs that suit depth 12 and depth 20 differ)."""
    import numpy as np
    shapes = {d: WL.make_shape(ctx, f"c5d{d}", classes=False) for d in range(12, 21)}
    tune_pool = WL.served_requests(n_jobs=2000, seed=0x7E)
    tuned = {}
    for d, sh in shapes.items():
        sample = [(r["a"], r["u0"], min(r["count"], 1 << 27)) for r in tune_pool if r["depth"] == d][:192]
        sh.autotune_jobs(sample, max_trials=tune_trials)
        tuned[d] = tune_rows_fixed(ccs, ctx, sh, sample, rows=2)
    reqs = WL.served_requests(n_jobs=n_jobs)
    total = sum(r["count"] for r in reqs)
    cal = WL.served_requests(n_jobs=400, seed=0xCA1)     # warm-up jobs (not timed)
    srv = ccs.Server(ctx, shapes, slice_candidates=slice_candidates, max_slice_survivors=max_slice_survivors)

    def served_pass(jobs):
        t = time.perf_counter()
        for r in jobs:
            srv.submit(r["depth"], r["a"], r["u0"], r["count"])
        srv.drain()
        dt = time.perf_counter() - t
        while srv.poll():
            pass
        return dt

    # the plain comparator: one filter launch per job (each with the unit size it would have
    # alone; jobs above 2^32 candidates, the most one launch takes per job, in pieces of
    # 2^32), then one prove launch per depth over that depth's survivors
    cap = 1 << 22
    d_out, d_cnt, d_res = ctx.dalloc(cap * 8), ctx.dalloc(64), ctx.dalloc(cap * 32)

    def job_sets(jobs):
        by_depth = {}
        for r in jobs:
            u0, left = r["u0"], r["count"]
            while left > 0:
                n = min(left, 1 << 32)
                by_depth.setdefault(r["depth"], []).append((r["a"], u0, n))
                u0 += n
                left -= n
        return [ccs.JobSet(shapes[d], jl) for d, jl in sorted(by_depth.items())]

    def plain_pass(sets):
        ctx.synchronize()
        t = time.perf_counter()
        for js in sets:
            js.launch_each(d_out, cap, d_cnt)
            js.prove_launch(d_out, d_cnt, cap, d_res)
        ctx.synchronize()
        return time.perf_counter() - t

    cal_sets, full_sets = job_sets(cal), job_sets(reqs)
    served_pass(cal)
    plain_pass(cal_sets)
    caps, plains = [], []
    for _ in range(reps):
        caps.append(total / served_pass(reqs))
        plains.append(total / plain_pass(full_sets))
    for js in cal_sets + full_sets:
        js.close()
    # one launch per depth over long jobs (64 jobs of 2^30 candidates each, the same tuned
    # shapes): a rate the served form cannot pass, so its ratio to the plain loop caps the
    # served speed-up for this job mix
    whole_ms, whole_cand = {}, 0
    for d, sh in shapes.items():
        js = ccs.JobSet(sh, [(r["a"], r["u0"], 1 << 30) for r in tune_pool if r["depth"] == d][:64])
        js.launch(d_out, cap, d_cnt)
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(3):
            js.launch(d_out, cap, d_cnt)
        ctx.synchronize()
        whole_ms[str(d)] = (time.perf_counter() - t) / 3 * 1e3
        whole_cand += js.candidates
        js.close()
    whole = whole_cand / (sum(whole_ms.values()) / 1e3)
    for ptr in (d_out, d_cnt, d_res):
        ctx.dfree(ptr)
    ratios = sorted(c / p for c, p in zip(caps, plains))
    capacity = float(np.median(caps))
    # the Poisson stream at `load` times the capacity
    lam = load * capacity / (total / len(reqs))
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
    return {"workload": "config 5: Poisson stream, depth U{12..20}, width U{64,80,96}, classes mod 31#, "
                        "lengths log-uniform in [2^24, 2^34]",
            "jobs": len(done), "candidates": total, "load": load, "arrival_rate_per_s": lam,
            "capacity_candidates_per_s": capacity, "capacity_runs": caps,
            "plain_loop_candidates_per_s": float(np.median(plains)), "plain_loop_runs": plains,
            "speedup_capacity_vs_plain": ratios[len(ratios) // 2],
            "speedup_runs": {"min": ratios[0], "median": ratios[len(ratios) // 2], "max": ratios[-1], "n": reps},
            "whole_set_candidates_per_s": whole, "whole_set_ms_per_depth": whole_ms,
            "speedup_cap_whole_set_vs_plain": whole / float(np.median(plains)),
            "aggregate_candidates_per_s": total / wall, "wall_s": wall,
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "slices": st["slices"], "launches": st["launches"],
            "incomplete_jobs": st["incomplete"],
            "slice_candidates": slice_candidates, "tuning": {str(d): t for d, t in tuned.items()}}


def measure_families(ccs, ctx, depth=20, q_per_family=1 << 28, exp_start=1, bits=(90, 127)):
    """The v16 Q-iteration workload: the reference's pool of the 100 most frequent immune
    fingerprints (known_fingerprints.txt), p_0 = Q*D*2^E - 1, each family's Q window for
    90..127-bit p_0, sieved at depth 20 by the primes 3..1009 (v16's GPU primes to 503 and
    its CPU ext sieve 509..1009).  One group launch over all families against one launch
    per family (v16's per-variant dispatch), each form at the best tuning a coordinate
    descent finds for it on the launch itself; the unit is (Q, V) pairs per second."""
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

    defaults = [sh.tuning for sh in shapes]

    def apply(tun):
        """one tuning for every family (None: each family's own default); CcsError when it
        does not fit some family's tables"""
        for sh, own in zip(shapes, defaults):
            if tun is None:
                sh.set_tuning(**own)
                continue
            nf = int(sh.info.n_filter)   # a family's immune primes are not among its filter primes
            nd = min(tun["n_dense"], nf)
            sh.set_tuning(**dict(tun, n_dense=nd, n_tierb=min(tun["n_tierb"], nf - nd)))

    def group_s(tun):
        apply(tun)
        gs = ccs.GroupSet(shapes, jobs)
        t = best_of(lambda: gs.launch(d_out, cap, d_cnt))
        gs.close()
        return t

    def loop_s(tun):
        apply(tun)
        sets = [ccs.JobSet(shapes[k], [j for kk, j in jobs if kk == k]) for k in range(len(shapes))]
        t = best_of(lambda: [js.launch(d_out, cap, d_cnt) for js in sets])
        for js in sets:
            js.close()
        return t

    def tune(timer):
        """coordinate descent on the launch form itself: a group launch reloads each family's
        LDS image as its blocks reach that family, so the single-shape autotuner (which never
        pays that) misjudges pattern sizes there"""
        best, tb = None, timer(None)
        start = defaults[0]
        knobs = (("n_dense", [sum(1 for q in primes if q < c) for c in (64, 100, 128, 200, 256)]),
                 ("pair_max", [0, 4096, 16384, 65536]), ("unit_j", [65536, 262144, 1048576]),
                 ("n_tierb", [64, 128, 192]))
        for knob, values in knobs:
            for v in values:
                base = best or start
                if v == base[knob]:
                    continue
                t = dict(base, **{knob: v})
                try:
                    tt = timer(t)
                except ccs.CcsError:
                    continue
                if tt < tb:
                    best, tb = t, tt
        return best, tb

    tuned_g, tg = tune(group_s)
    tuned_l, tl = tune(loop_s)
    apply(tuned_g)
    gs = ccs.GroupSet(shapes, jobs)
    gs.launch(d_out, cap, d_cnt)
    survivors = gs.stats(d_cnt)["survivors"]
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
            "survivors": survivors, "tuning_group": tuned_g, "tuning_per_family_loop": tuned_l}


def lds_probe():
    """LDS read throughput of this GPU (lds_probe: every CU reading with ds_read_b128)."""
    import subprocess
    try:
        r = subprocess.run([LDS_PROBE], capture_output=True, text=True, timeout=120)
        return json.loads(r.stdout.strip().splitlines()[-1]) if r.returncode == 0 else None
    except (OSError, ValueError, IndexError, subprocess.TimeoutExpired):
        return None


def roofline(shape, name, cand, ms_sieve, e_tests, value, probe=None):
    """The sieve kernel against the two resources it runs on: int32 VALU lane-ops (the
    bound: its algorithmic lower limit on time is the larger one) and LDS bytes, both from
    the stage work model per candidate and the kernel's HIP-event time; the counter profile
    beside them (same tuning only); HBM traffic from the profile's FETCH/WRITE counters."""
    from ccs import workloads as WL
    info = shape.info
    m = WL.sieve_work_model(shape.filter_primes, shape.depth, shape.pos0_max_q, info.n_dense, info.n_slots,
                            info.n_tierb, info.n_tierb_lds)
    t = ms_sieve / 1e3
    valu = cand * m["valu_lane_ops"] / t / 1e12
    lds = cand * m["lds_bytes"] / t / 1e9
    out = {"bound": "valu", "achieved": valu, "peak": WL.INT32_PEAK_TOPS, "unit": "Tops/s (int32 lane-ops)",
           "frac": valu / WL.INT32_PEAK_TOPS, "traffic": None, "kernel": "sieve_kernel", "kernel_ms": ms_sieve,
           "model_per_candidate": {"valu_lane_ops": m["valu_lane_ops"], "lds_bytes": m["lds_bytes"],
                                   "n_slots": int(info.n_slots), "n_dense": int(info.n_dense),
                                   "n_tierb": int(info.n_tierb), "p_dense": m["p_dense"],
                                   "stage_b_probes": m["stage_b_probes"]},
           "lds": {"achieved": lds, "peak": WL.LDS_PEAK_GBS, "unit": "GB/s", "frac": lds / WL.LDS_PEAK_GBS,
                   "probe": probe},
           "equivalent_mod_ops": {"value": value * e_tests, "unit": "residue tests/s",
                                  "e_tests_per_candidate": e_tests,
                                  "note": "the reference predicate's per-candidate tests (4 int32 ops each); "
                                          "a dense pattern OR tests 32 candidates at once, so this is not a "
                                          "fraction of any peak"}}
    prof = SIEVE_PROFILE.get(name)
    tun = shape.tuning
    if prof and all(tun.get(k) == v for k, v in prof["tuning"].items()):
        out["traffic"] = (prof["fetch_kib"] * prof["fetch_scale"] + prof["write_kib"] * prof["write_scale"]) * 1024
        cu_clk = WL.MI355X_CUS * t * WL.MI355X_CLK_GHZ * 1e9
        out["hardware"] = {
            "valu_issue_frac": prof["sq_insts_valu"] / cu_clk,
            "valu_lane_ops_per_candidate": prof["sq_insts_valu"] * 64 / cand,
            "lds_insts": prof["sq_insts_lds"],
            "lds_insts_model": cand * (m["dense_lds_reads_b128"] + m["stage_b_lds_reads"]),
            "lds_bank_conflict_frac": prof["sq_lds_bank_conflict"] / cu_clk,
            "profile": prof["profile"]}
    return out



def served_depth(args):
    """One served depth profile on its own: 64 of the stream's jobs of that depth, 2^30
    candidates each (2^36 in all), as one job-set launch (filter: sieve + deep stage) per
    step, at the profile's default tuning or --tuning; prints one JSON line (a profiling
    target for the served kernels, not the headline)."""
    import ccs
    from ccs import workloads as WL
    ctx = ccs.Context(0)
    sh = WL.make_shape(ctx, f"c5d{args.served_depth}", classes=False)
    if args.tuning:
        sh.set_tuning(**json.loads(args.tuning))
    reqs = [r for r in WL.served_requests(n_jobs=2000, seed=0xC5) if r["depth"] == args.served_depth][:64]
    js = ccs.JobSet(sh, [(r["a"], r["u0"], 1 << 30) for r in reqs])
    cap = 1 << 22
    d_out, d_cnt = ctx.dalloc(cap * 8), ctx.dalloc(64)
    for _ in range(args.warmup):
        js.launch(d_out, cap, d_cnt)
    ctx.synchronize()
    ctx.profile_start()
    t = time.perf_counter()
    for _ in range(args.steps):
        js.launch(d_out, cap, d_cnt)
    ctx.synchronize()
    dt = (time.perf_counter() - t) / args.steps
    kt = ctx.profile_stop()
    st = js.stats(d_cnt)
    print(json.dumps({"workload": f"served depth {args.served_depth}: 64 jobs of 2^30 candidates, classes mod 31#",
                      "candidates_per_step": js.candidates, "ms_per_step": dt * 1e3,
                      "candidates_per_s": js.candidates / dt, "tuning": sh.tuning,
                      "kernels_ms": {"sieve": kt["ms_sieve"] / max(kt["n_sieve"], 1),
                                     "deep": kt["ms_deep"] / max(kt["n_deep"], 1)},
                      "alive_dense": st["alive_dense"], "alive_tierb": st["alive_tierb"],
                      "survivors": st["survivors"]}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=2.0, help="seconds per CPU-baseline matrix row")
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
    ap.add_argument("--served-depth", type=int, default=0,
                    help="only time one served depth profile's job-set launch (64 jobs of 2^30 candidates)")
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
    if args.served_depth:
        served_depth(args)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this pipeline shards by disjoint tile ranges with no "
              "collective, and multi-GPU runs are not part of this build (use --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import ccs
    from ccs import workloads as WL

    # the LDS ceiling of this box, measured in a child process before this one opens the GPU
    probe = lds_probe()
    if probe:
        probe = {"b128_gbs": probe["b128_bytes_per_s"] / 1e9, "b32_pair_gbs": probe["b32_pair_bytes_per_s"] / 1e9,
                 "clock_khz": probe["clock_khz"], "cus": probe["cus"]}
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
    if not args.tuning:
        # rows per lane step chosen on whole steps (setup work, not timed): with 4 rows the
        # sieve takes a CU's whole register file, so a prove-heavy shape (the v13 campaign
        # shape proves ~8·10^5 survivors a step) can no longer run its prove beside it
        best_rg, best_t = None, 1e30
        for rg in (4, 2):
            try:
                shape.set_tuning(rows_per_group=rg)
            except ccs.CcsError:
                continue
            for _ in range(3):
                step()
            ctx.synchronize()
            t = time.perf_counter()
            for _ in range(8):
                step()
            ctx.synchronize()
            if time.perf_counter() - t < best_t:
                best_rg, best_t = rg, time.perf_counter() - t
        if best_rg is not None:
            shape.set_tuning(rows_per_group=best_rg)
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
    roof = roofline(shape, name, cand, ms_sieve, e_tests, value, probe)
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
        "roofline": roof,
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
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
