"""Host-side mirror of the reference's filter/prove interface over libccs.so.

The reference (nmicic/cunningham-chain-search) exposes its pipeline through one engine
per file: tables + CRT bases (v13.cu:793-1059), the K1 filter launch over a tile batch
(v13.cu:3764-3796), and the prover over the survivors (v13.cu:1794-2068).  This module
gives the same operations as Python calls on top of the C ABI (include/ccs.h); all
compute runs in the HIP kernels.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import (CHAIN_RESULT_DTYPE, JOB_SURVIVOR_DTYPE, SURVIVOR_DTYPE, CcsError,  # noqa: F401
                   F_BITS_OK, F_NONROOT, F_OVERFLOW, F_PRIME, F_TRIAL_OK, F_UNSUPPORTED)

__all__ = ["Context", "Shape", "JobSet", "Server", "CcsError", "device_count", "format_results", "u128_split"]


def u128_split(v):
    return v & ((1 << 64) - 1), v >> 64


def device_count():
    n = C.c_int(0)
    L.check(L.load().ccs_device_count(C.byref(n)), "ccs_device_count")
    return n.value


class Context:
    """One GPU: stream, scratch buffers and timing events (ccs_ctx)."""

    def __init__(self, device=0):
        self.lib = L.load()
        self.h = C.c_void_p()
        L.check(self.lib.ccs_ctx_create(device, C.byref(self.h)), "ccs_ctx_create")

    def close(self):
        if self.h:
            self.lib.ccs_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        """wait for all work of the device (every stream)"""
        L.check(self.lib.ccs_synchronize(self.h), "ccs_synchronize")

    # streams and events for overlapping launches (default priority)
    def stream(self):
        p = C.c_void_p()
        L.check(self.lib.ccs_stream_create(self.h, C.byref(p)), "ccs_stream_create")
        return p

    def stream_destroy(self, st):
        L.check(self.lib.ccs_stream_destroy(self.h, st), "ccs_stream_destroy")

    def event(self):
        p = C.c_void_p()
        L.check(self.lib.ccs_event_create(self.h, C.byref(p)), "ccs_event_create")
        return p

    def event_destroy(self, ev):
        L.check(self.lib.ccs_event_destroy(self.h, ev), "ccs_event_destroy")

    def record(self, ev, stream=None):
        L.check(self.lib.ccs_event_record(self.h, ev, stream), "ccs_event_record")

    def wait(self, stream, ev):
        L.check(self.lib.ccs_stream_wait_event(self.h, stream, ev), "ccs_stream_wait_event")

    def status(self, stream=None):
        """stage-B overflow status of the stream's launches since the last call
        (ccs_stream_status): a dict, with "overflow" True when some launch overflowed"""
        st = L.LaunchStatus()
        rc = self.lib.ccs_stream_status(self.h, stream, C.byref(st))
        if rc < 0:
            L.check(rc, "ccs_stream_status")
        return dict(st.as_dict(), overflow=rc == L.CCS_OVERFLOW)

    def limit_stage_b(self, entries):
        """test knob: launches use at most `entries` of their stage-B buffer (0 = all)"""
        L.check(self.lib.ccs_debug_limit_stage_b(self.h, entries), "ccs_debug_limit_stage_b")

    def deep_batch(self, per_wave):
        """test knob: deep-stage records per wave batch (1: one at a time, 0: by count)"""
        L.check(self.lib.ccs_debug_deep_batch(self.h, per_wave), "ccs_debug_deep_batch")

    def profile_start(self):
        L.check(self.lib.ccs_profile_start(self.h), "ccs_profile_start")

    def profile_stop(self):
        t = L.KernelTimes()
        L.check(self.lib.ccs_profile_stop(self.h, C.byref(t)), "ccs_profile_stop")
        return t.as_dict()

    # device buffers for the asynchronous entry points
    def dalloc(self, nbytes):
        p = C.c_void_p()
        L.check(self.lib.ccs_dmalloc(self.h, C.byref(p), nbytes), "ccs_dmalloc")
        return p

    def dfree(self, p):
        L.check(self.lib.ccs_dfree(self.h, p), "ccs_dfree")

    def h2d(self, dptr, arr):
        arr = np.ascontiguousarray(arr)
        L.check(self.lib.ccs_memcpy_h2d(self.h, dptr, arr.ctypes.data, arr.nbytes), "h2d")

    def d2h(self, arr, dptr):
        L.check(self.lib.ccs_memcpy_d2h(self.h, arr.ctypes.data, dptr, arr.nbytes), "d2h")
        return arr


def _u32arr(xs):
    a = np.ascontiguousarray(np.asarray(list(xs), dtype=np.uint32))
    return a, a.ctypes.data_as(C.POINTER(C.c_uint32))


class Shape:
    """CRT shape + filter profile resident on the GPU (ccs_shape).

    n = (tile*WP + k)*M + base, M = prod(crt), WP = prod(wheel), n = 5 (mod 6);
    filter primes tested at positions [0, depth) (q <= pos0_max_q) or [1, depth)."""

    def __init__(self, ctx, crt, wheel, target, depth, filter_primes, stage_end=(),
                 pos0_max_q=97, classes=True, tuning=None):
        self.ctx = ctx
        self.lib = ctx.lib
        self._keep = []
        crt_a, crt_p = _u32arr(crt)
        wh_a, wh_p = _u32arr(wheel)
        fp_a, fp_p = _u32arr(filter_primes)
        st_a, st_p = _u32arr(stage_end)
        self._keep += [crt_a, wh_a, fp_a, st_a]
        self.crt, self.wheel, self.filter_primes = list(crt), list(wheel), [int(x) for x in fp_a]
        self.target, self.depth, self.pos0_max_q = target, depth, pos0_max_q
        self.stage_end = list(stage_end)
        p = L.ShapeParams(len(crt_a), crt_p, len(wh_a), wh_p, target, depth, len(fp_a), fp_p,
                          len(st_a), st_p, pos0_max_q, 0 if classes else L.CCS_SHAPE_NO_CLASSES)
        self.h = C.c_void_p()
        tp = None
        if tuning is not None:
            tp = C.byref(tuning if isinstance(tuning, L.Tuning) else L.Tuning(**tuning))
        L.check(self.lib.ccs_shape_build(ctx.h, C.byref(p), tp, C.byref(self.h)), "ccs_shape_build")
        self._refresh_info()

    def _refresh_info(self):
        info = L.ShapeInfo()
        L.check(self.lib.ccs_shape_query(self.h, C.byref(info)), "ccs_shape_query")
        self.info = info
        self.M, self.WP, self.W = int(info.M), int(info.WP), int(info.W)
        self.n_bases, self.n_classes = int(info.n_bases), int(info.n_classes)

    def close(self):
        if self.h:
            self.lib.ccs_shape_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def tuning(self):
        t = L.Tuning()
        L.check(self.lib.ccs_shape_get_tuning(self.h, C.byref(t)), "get_tuning")
        return t.as_dict()

    def set_tuning(self, **kw):
        t = L.Tuning(**{**self.tuning, **kw})
        L.check(self.lib.ccs_shape_set_tuning(self.h, C.byref(t)), "ccs_shape_set_tuning")
        self._refresh_info()

    def autotune(self, tile_start, tile_count, max_trials=40):
        """Pick the fastest tuning on a sample range (ccs_autotune); returns (tuning, ms)."""
        t = L.Tuning()
        ms = C.c_float(0)
        L.check(self.lib.ccs_autotune(self.ctx.h, self.h, L.U128.of(tile_start), tile_count, max_trials,
                                      C.byref(t), C.byref(ms)), "ccs_autotune")
        self._refresh_info()
        return t.as_dict(), float(ms.value)

    def autotune_jobs(self, jobs, max_trials=40):
        """Pick the fastest tuning on the batched launch over sample jobs (ccs_autotune_jobs)."""
        jarr = jobs_array(jobs)
        t = L.Tuning()
        ms = C.c_float(0)
        L.check(self.lib.ccs_autotune_jobs(self.ctx.h, self.h, jarr.ctypes.data, len(jarr), max_trials,
                                           C.byref(t), C.byref(ms)), "ccs_autotune_jobs")
        self._refresh_info()
        return t.as_dict(), float(ms.value)

    def bases(self):
        out = np.zeros(self.n_bases, np.uint64)
        L.check(self.lib.ccs_shape_bases(self.h, out.ctypes.data, self.n_bases), "bases")
        return out

    def classes(self):
        n = self.n_classes
        a = np.zeros(n, np.uint64)
        b = np.zeros(n, np.uint32)
        k = np.zeros(n, np.uint32)
        r = np.zeros(n, np.uint32)
        L.check(self.lib.ccs_shape_classes(self.h, a.ctypes.data, b.ctypes.data, k.ctypes.data,
                                           r.ctypes.data, n), "classes")
        return a, b, k, r

    def tile_range(self, bits, prefix=None, lanes=0, lane_id=0):
        """[tile_min, tile_max) for --bits, --prefix 0b..., --prefix-lanes (ccs_tile_range)."""
        pv, pb = 0, 0
        if prefix:
            if prefix[:2] not in ("0b", "0B") or not set(prefix[2:]) <= {"0", "1"}:
                raise ValueError("prefix must be binary (0b...)")
            pv, pb = int(prefix[2:], 2), len(prefix) - 2
        lo, hi = L.U128(), L.U128()
        L.check(self.lib.ccs_tile_range(self.h, bits, L.U128.of(pv), pb, lanes, lane_id, C.byref(lo), C.byref(hi)),
                "ccs_tile_range")
        return lo.lo | (lo.hi << 64), hi.lo | (hi.hi << 64)

    def decompose(self, n):
        """ccs_shape_decompose: where n sits on the lattice (the --verify-root diagnostic)."""
        d = L.Decomposition()
        L.check(self.lib.ccs_shape_decompose(self.h, L.U128.of(n), C.byref(d)), "ccs_shape_decompose")
        return d.as_dict()

    def count_candidates(self, tile_start, tile_count):
        return int(self.lib.ccs_count_candidates(self.h, L.U128.of(tile_start), tile_count))

    def filter(self, tile_start, tile_count, cap=1 << 20):
        """Plain form: survivors of tiles [tile_start, tile_start + tile_count) as a
        structured array of v13 Survivor records (unordered), and the stage counters."""
        out = np.zeros(cap, SURVIVOR_DTYPE)
        n = C.c_uint32(0)
        st = L.FilterStats()
        rc = self.lib.ccs_filter(self.ctx.h, self.h, L.U128.of(tile_start), tile_count,
                                 out.ctypes.data, cap, C.byref(n), C.byref(st))
        L.check(rc, "ccs_filter")
        if rc == L.CCS_OVERFLOW:
            raise CcsError(f"survivor overflow: {n.value} > cap {cap}")
        return out[: n.value].copy(), st.as_dict()

    def filter_launch_noalloc(self, tile_start, tile_count, d_out, cap, d_count, stream=None):
        """ccs_filter_launch with no device allocation (ccs_filter_launch_noalloc): the
        library's fixed per-device scratch; *d_count = 0xFFFFFFFF on a stage-B overflow"""
        L.check(self.lib.ccs_filter_launch_noalloc(self.ctx.h, self.h, L.U128.of(tile_start), tile_count, d_out,
                                                   cap, d_count, stream), "ccs_filter_launch_noalloc")

    def filter_launch(self, tile_start, tile_count, d_out, cap, d_count, stream=None):
        """Asynchronous plain filter into device buffers (ccs_filter_launch)."""
        L.check(self.lib.ccs_filter_launch(self.ctx.h, self.h, L.U128.of(tile_start), tile_count,
                                           d_out, cap, d_count, stream), "ccs_filter_launch")

    def prove_launch(self, tile_start, d_surv, d_count, cap, d_out, bits=0, test="bpsw", bases=(2,),
                     stream=None):
        """Asynchronous prove of device survivors (ccs_prove_launch)."""
        pp = _prove_params(bits, test, bases)
        L.check(self.lib.ccs_prove_launch(self.ctx.h, self.h, L.U128.of(tile_start), d_surv, d_count, cap,
                                          C.byref(pp), d_out, stream), "ccs_prove_launch")

    def chain_roots_launch(self, d_surv, d_res, d_count, cap, d_out, out_cap, d_root_count, d_prime_count=None,
                           stream=None):
        """The v13 K2 output (ChainResult records of the chain roots) from prove results
        (ccs_chain_roots_launch)."""
        L.check(self.lib.ccs_chain_roots_launch(self.ctx.h, d_surv, d_res, d_count, cap, d_out, out_cap,
                                                d_root_count, d_prime_count, stream), "chain_roots_launch")

    def prove(self, tile_start, surv, bits=0, test="bpsw", bases=(2,)):
        surv = np.ascontiguousarray(surv, SURVIVOR_DTYPE)
        out = np.zeros(len(surv), CHAIN_RESULT_DTYPE)
        pp = _prove_params(bits, test, bases)
        L.check(self.lib.ccs_prove(self.ctx.h, self.h, L.U128.of(tile_start), surv.ctypes.data,
                                   len(surv), C.byref(pp), out.ctypes.data), "ccs_prove")
        return out

    def confirm_host(self, tile_start, surv, bits=0, test="bpsw", bases=(2,), nthreads=0):
        """Host confirmation engine over plain survivors (ccs_confirm_host): the GPU prove's
        semantics on CPU threads.  Returns (results, stats dict)."""
        surv = np.ascontiguousarray(surv, SURVIVOR_DTYPE)
        out = np.zeros(len(surv), CHAIN_RESULT_DTYPE)
        pp = _prove_params(bits, test, bases)
        st = L.ConfirmStats()
        L.check(self.lib.ccs_confirm_host(self.h, L.U128.of(tile_start), surv.ctypes.data, len(surv),
                                          C.byref(pp), nthreads, out.ctypes.data, C.byref(st)),
                "ccs_confirm_host")
        return out, st.as_dict()

    def batch(self, jobs, cap=1 << 20):
        """Batched form: one persistent launch over class-form jobs (a, u0, count)."""
        jarr = jobs_array(jobs)
        out = np.zeros(cap, JOB_SURVIVOR_DTYPE)
        n = C.c_uint32(0)
        st = L.FilterStats()
        rc = self.lib.ccs_batch(self.ctx.h, self.h, jarr.ctypes.data, len(jarr), out.ctypes.data, cap,
                                C.byref(n), C.byref(st))
        L.check(rc, "ccs_batch")
        if rc == L.CCS_OVERFLOW:
            raise CcsError(f"survivor overflow: {n.value} > cap {cap}")
        return out[: n.value].copy(), st.as_dict()

    def prove_jobs(self, jobs, surv, bits=0, test="bpsw", bases=(2,)):
        jarr = jobs_array(jobs)
        surv = np.ascontiguousarray(surv, JOB_SURVIVOR_DTYPE)
        out = np.zeros(len(surv), CHAIN_RESULT_DTYPE)
        pp = _prove_params(bits, test, bases)
        L.check(self.lib.ccs_prove_jobs(self.ctx.h, self.h, jarr.ctypes.data, len(jarr), surv.ctypes.data,
                                        len(surv), C.byref(pp), out.ctypes.data), "ccs_prove_jobs")
        return out

    def n_of(self, tile_start, s):
        """n of a plain survivor record (v13.cu:1819-1826)."""
        return ((tile_start + int(s["tile_idx"])) * self.WP + int(s["k_offset"])) * self.M + \
            int(self.bases()[int(s["base_idx"])])


class FamilyShape(Shape):
    """Fingerprint family of the v16 Q-iteration front end (ccs_family_shape_build):
    p_0 = Q*D*2^E - 1, sieved at positions [0, depth) by the filter primes that do not
    divide D.  Jobs are Q windows (0, Q0, count) for JobSet / batch / prove_jobs."""

    def __init__(self, ctx, D, exp_start, target, depth, filter_primes, tuning=None):
        self.ctx = ctx
        self.lib = ctx.lib
        fp_a, fp_p = _u32arr(filter_primes)
        self._keep = [fp_a]
        self.crt, self.wheel = [], []
        self.filter_primes = [int(x) for x in fp_a if D % int(x)]
        self.target, self.depth, self.pos0_max_q, self.stage_end = target, depth, None, []
        self.D, self.exp_start = D, exp_start
        p = L.FamilyParams(L.U128.of(D), exp_start, target, depth, len(fp_a), fp_p)
        self.h = C.c_void_p()
        tp = None
        if tuning is not None:
            tp = C.byref(tuning if isinstance(tuning, L.Tuning) else L.Tuning(**tuning))
        L.check(self.lib.ccs_family_shape_build(ctx.h, C.byref(p), tp, C.byref(self.h)), "ccs_family_shape_build")
        self._refresh_info()
        self.W = D << exp_start

    def window(self, bits_lo, bits_hi):
        return family_window(self.D, self.exp_start, bits_lo, bits_hi)

    def n_of_q(self, q):
        return q * self.W - 1


def family_window(D, exp_start, bits_lo, bits_hi):
    """(q_min, q_max) of candidates with bits_lo..bits_hi bits, or None when empty."""
    lo, hi = L.U128(), L.U128()
    rc = L.check(L.load().ccs_family_window(L.U128.of(D), exp_start, bits_lo, bits_hi, C.byref(lo), C.byref(hi)),
                 "ccs_family_window")
    if rc == L.CCS_DONE:
        return None
    return lo.lo | (lo.hi << 64), hi.lo | (hi.hi << 64)


def _pool(fn, arg, keep, cap):
    k = np.ascontiguousarray(np.asarray(list(keep), np.uint32))
    while True:
        out = (L.PoolEntry * max(cap, 1))()
        n = C.c_uint32(0)
        rc = L.check(fn(arg, k.ctypes.data if len(k) else None, len(k), out, cap, C.byref(n)), "pool")
        if rc == L.CCS_OVERFLOW:
            cap = n.value
            continue
        return [out[i].as_dict() for i in range(n.value)]


def pool_parse(text, keep=(), cap=1024):
    """Pool list "Z#/X,P*Q^k,auto:Z#:K" -> [{name, D, primes}] (ccs_pool_parse)."""
    return _pool(L.load().ccs_pool_parse, text.encode(), keep, cap)


def pool_load(path, keep=(), cap=4096):
    """Pool file (one list per line, '#' comments) -> [{name, D, primes}] (ccs_pool_load)."""
    return _pool(L.load().ccs_pool_load, str(path).encode(), keep, cap)


def dedup(values):
    """First-occurrence flags of integer values < 2^128 (ccs_dedup_u128)."""
    arr = np.zeros((len(values), 2), np.uint64)
    for i, v in enumerate(values):
        arr[i] = (v & ((1 << 64) - 1), v >> 64)
    first = np.zeros(len(values), np.uint8)
    L.check(L.load().ccs_dedup_u128(arr.ctypes.data, len(values), first.ctypes.data), "ccs_dedup_u128")
    return first.astype(bool)


class JobSet:
    """Class-form jobs uploaded once (ccs_jobset): batched launches, the per-job loop, graphs."""

    def __init__(self, shape, jobs):
        self.shape, self.ctx, self.lib = shape, shape.ctx, shape.lib
        self.jobs = jobs_array(jobs)
        self.h = C.c_void_p()
        L.check(self.lib.ccs_jobset_create(self.ctx.h, shape.h, self.jobs.ctypes.data, len(self.jobs),
                                           C.byref(self.h)), "ccs_jobset_create")
        self.candidates = int(self.jobs["count"].sum())
        self._graphs = []

    def close(self):
        for g in self._graphs:
            self.lib.ccs_graph_free(g)
        self._graphs = []
        if self.h:
            self.lib.ccs_jobset_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update(self, jobs, stream=None):
        """Replace the jobs (at most as many as the set was built with): ccs_jobset_update."""
        arr = jobs_array(jobs)
        L.check(self.lib.ccs_jobset_update(self.ctx.h, self.h, arr.ctypes.data, len(arr), stream), "jobset_update")
        self.jobs = arr
        self.candidates = int(arr["count"].sum())

    def launch(self, d_out, cap, d_count, stream=None):
        L.check(self.lib.ccs_jobset_launch(self.ctx.h, self.h, d_out, cap, d_count, stream), "jobset_launch")

    def launch_each(self, d_out, cap, d_count, stream=None):
        L.check(self.lib.ccs_jobset_launch_each(self.ctx.h, self.h, d_out, cap, d_count, stream),
                "jobset_launch_each")

    def stats(self, d_count):
        st = L.FilterStats()
        L.check(self.lib.ccs_jobset_stats(self.ctx.h, self.h, d_count, C.byref(st)), "jobset_stats")
        return st.as_dict()

    def status(self, stream=None):
        """stage-B overflow status of the set's launches since the last call (ccs_jobset_status)"""
        st = L.LaunchStatus()
        rc = self.lib.ccs_jobset_status(self.ctx.h, self.h, stream, C.byref(st))
        if rc < 0:
            L.check(rc, "ccs_jobset_status")
        return dict(st.as_dict(), overflow=rc == L.CCS_OVERFLOW)

    def graph(self, d_out, cap, d_count):
        g = C.c_void_p()
        L.check(self.lib.ccs_graph_create(self.ctx.h, self.h, d_out, cap, d_count, C.byref(g)), "graph_create")
        self._graphs.append(g)
        return g

    def graph_launch(self, g, stream=None):
        L.check(self.lib.ccs_graph_launch(self.ctx.h, g, stream), "graph_launch")

    def top_gate(self, d_in, d_in_count, in_cap, top_pos, reps, d_out, out_cap, d_out_count, stream=None):
        """K1.5 top-position MR gate of a family set (ccs_family_top_gate_launch)."""
        L.check(self.lib.ccs_family_top_gate_launch(self.ctx.h, self.h, d_in, d_in_count, in_cap, top_pos, reps,
                                                    d_out, out_cap, d_out_count, stream), "top_gate")

    def prove_launch(self, d_surv, d_count, cap, d_out, bits=0, test="bpsw", bases=(2,), stream=None):
        pp = _prove_params(bits, test, bases)
        L.check(self.lib.ccs_prove_jobset_launch(self.ctx.h, self.h, d_surv, d_count, cap, C.byref(pp),
                                                 d_out, stream), "prove_jobset_launch")


class GroupSet(JobSet):
    """Jobs of several shapes in one persistent launch (ccs_groupset_create): jobs is a list
    of (shape_index, (a, u0, count)) grouped by shape index."""

    def __init__(self, shapes, jobs):
        self.shapes = list(shapes)
        self.shape, self.ctx, self.lib = self.shapes[0], self.shapes[0].ctx, self.shapes[0].lib
        self.job_shape = np.ascontiguousarray(np.asarray([k for k, _ in jobs], np.uint32))
        self.jobs = jobs_array([j for _, j in jobs])
        arr = (C.c_void_p * len(self.shapes))(*[sh.h.value for sh in self.shapes])
        self._arr = arr
        self.h = C.c_void_p()
        L.check(self.lib.ccs_groupset_create(self.ctx.h, arr, len(self.shapes), self.jobs.ctypes.data,
                                             self.job_shape.ctypes.data, len(self.jobs), C.byref(self.h)),
                "ccs_groupset_create")
        self.candidates = int(self.jobs["count"].sum())
        self._graphs = []

    def update(self, jobs, stream=None):
        job_shape = np.ascontiguousarray(np.asarray([k for k, _ in jobs], np.uint32))
        arr = jobs_array([j for _, j in jobs])
        L.check(self.lib.ccs_groupset_update(self.ctx.h, self.h, arr.ctypes.data, job_shape.ctypes.data, len(arr),
                                             stream), "ccs_groupset_update")
        self.jobs, self.job_shape = arr, job_shape
        self.candidates = int(arr["count"].sum())


class Server:
    """Continuous-batching request loop (ccs_server): submit jobs of any depth / width /
    length, poll completions with their timestamps, fetch per-job results.  Each admitted
    job's sieve phases stay resident on the device (recompute=True turns that off: the
    comparison mode)."""

    def __init__(self, ctx, shapes_by_depth, mem_budget_bytes=256 << 20, max_active=1024,
                 max_slice_survivors=1 << 16, slice_candidates=1 << 35, min_chunk=1 << 18,
                 max_chunk=1 << 32, bits=0, test="bpsw", bases=(2,), recompute=False, group_width_ratio=0):
        self.ctx, self.lib = ctx, ctx.lib
        self.depths = sorted(shapes_by_depth)
        self.shapes = [shapes_by_depth[d] for d in self.depths]
        arr = (C.c_void_p * len(self.shapes))(*[sh.h.value for sh in self.shapes])
        dep = (C.c_int32 * len(self.depths))(*self.depths)
        flags = L.SERVER_RECOMPUTE_PHASES if recompute else 0
        p = L.ServerParams(mem_budget_bytes, max_active, flags, max_slice_survivors, group_width_ratio,
                           slice_candidates, min_chunk, max_chunk, _prove_params(bits, test, bases))
        self.h = C.c_void_p()
        self.incomplete = False
        L.check(self.lib.ccs_server_create(ctx.h, arr, dep, len(self.shapes), C.byref(p), C.byref(self.h)),
                "ccs_server_create")

    def submit(self, depth, a, u0, count, tag=0, bits=0):
        r = L.Request(tag, depth, bits, a, L.U128.of(u0), count)
        jid = C.c_uint64(0)
        L.check(self.lib.ccs_server_submit(self.h, C.byref(r), C.byref(jid)), "ccs_server_submit")
        return jid.value

    def poll(self, cap=4096):
        """finished jobs since the last poll (each with its flags; JOB_INCOMPLETE marks a job
        whose survivors were cut by a full buffer)"""
        buf = (L.JobDone * cap)()
        n = C.c_uint32(0)
        rc = self.lib.ccs_server_poll(self.h, buf, cap, C.byref(n))
        if rc < 0:
            L.check(rc, "ccs_server_poll")
        self.incomplete |= rc == L.CCS_OVERFLOW
        return [buf[i].as_dict() for i in range(n.value)]

    def results(self, job_id):
        """(u_off uint64 array, chain results) of a finished job"""
        n = C.c_uint32(0)
        L.check(self.lib.ccs_server_job_results(self.h, job_id, None, None, 0, C.byref(n)), "results")
        u_off = np.zeros(n.value, np.uint64)
        res = np.zeros(n.value, CHAIN_RESULT_DTYPE)
        L.check(self.lib.ccs_server_job_results(self.h, job_id, u_off.ctypes.data, res.ctypes.data, n.value,
                                                C.byref(n)), "results")
        return u_off, res

    def drain(self):
        """wait for every submitted job; True when every job so far finished complete"""
        rc = self.lib.ccs_server_drain(self.h)
        if rc < 0:
            L.check(rc, "ccs_server_drain")
        return rc == L.CCS_OK

    def stats(self):
        c = L.ServerCounters()
        L.check(self.lib.ccs_server_stats(self.h, C.byref(c)), "ccs_server_stats")
        return c.as_dict()

    def shutdown(self):
        if self.h:
            rc = self.lib.ccs_server_shutdown(self.h)
            self.h = C.c_void_p()
            L.check(rc, "ccs_server_shutdown")

    def __del__(self):
        try:
            self.shutdown()
        except Exception:
            pass


def jobs_array(jobs):
    if isinstance(jobs, np.ndarray) and jobs.dtype == L.JOB_DTYPE:
        return np.ascontiguousarray(jobs)
    arr = np.zeros(len(jobs), L.JOB_DTYPE)
    for i, (a, u0, count) in enumerate(jobs):
        arr[i] = (a, u0 & ((1 << 64) - 1), u0 >> 64, count)
    return arr


def _prove_params(bits, test, bases):
    pp = L.ProveParams()
    pp.bits = bits
    pp.test_kind = L.CCS_TEST_BPSW if test == "bpsw" else L.CCS_TEST_MR
    bl = list(bases)
    pp.n_bases = len(bl)
    for i, b in enumerate(bl):
        pp.bases[i] = b
    return pp


class Campaign:
    """Campaign driver state (ccs_campaign_*): the v13 main loop's tile walk, stop rules,
    statistics and CC18_CUDA_v10 checkpoint.  Built from a shape (range from --bits /
    --prefix / --prefix-lanes) or from an explicit tile range."""

    WALKS = {"inherit": L.WALK_INHERIT, "sequential": L.WALK_SEQUENTIAL, "random": L.WALK_RANDOM}

    def __init__(self, shape=None, tile_range=None, bits=89, target=18, depth=18, prefix=None,
                 walk="inherit", lanes=0, lane_id=0, start=None, seed=None, batch_tiles=16384, max_tiles=0):
        self.lib = L.load()
        p = L.CampaignParams()
        p.bits, p.target, p.depth = bits, target, depth
        self._prefix = prefix.encode() if prefix else None
        p.prefix = self._prefix
        p.walk = self.WALKS[walk]
        p.lanes, p.lane_id = lanes, lane_id
        p.has_start = start is not None
        p.start = L.U128.of(start or 0)
        p.has_seed = seed is not None
        p.seed = seed or 0
        p.batch_tiles, p.max_tiles = batch_tiles, max_tiles
        h = C.c_void_p()
        if shape is not None:
            L.check(self.lib.ccs_campaign_create(shape.h, C.byref(p), C.byref(h)), "ccs_campaign_create")
        else:
            lo, hi = tile_range
            L.check(self.lib.ccs_campaign_create_range(L.U128.of(lo), L.U128.of(hi), C.byref(p), C.byref(h)),
                    "ccs_campaign_create_range")
        self.h = h

    def next(self):
        """(tile_start, tile_count), or None when the campaign is done."""
        ts, n = L.U128(), C.c_uint32(0)
        rc = self.lib.ccs_campaign_next(self.h, C.byref(ts), C.byref(n))
        if rc == L.CCS_DONE:
            return None
        L.check(rc, "ccs_campaign_next")
        return ts.lo | (ts.hi << 64), n.value

    def record(self, tile_start, tile_count, candidates=0, survivors=0, res=None):
        res = np.zeros(0, CHAIN_RESULT_DTYPE) if res is None else np.ascontiguousarray(res, CHAIN_RESULT_DTYPE)
        L.check(self.lib.ccs_campaign_record(self.h, L.U128.of(tile_start), tile_count, candidates, survivors,
                                             res.ctypes.data, len(res)), "ccs_campaign_record")

    def state(self):
        st = L.CampaignState()
        L.check(self.lib.ccs_campaign_get_state(self.h, C.byref(st)), "ccs_campaign_get_state")
        return st.as_dict()

    def save(self, path):
        L.check(self.lib.ccs_campaign_save(self.h, str(path).encode()), "ccs_campaign_save")

    def resume(self, path):
        L.check(self.lib.ccs_campaign_resume(self.h, str(path).encode()), "ccs_campaign_resume")

    def close(self):
        if self.h:
            self.lib.ccs_campaign_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def parse_prefix(text):
    """ccs_parse_prefix: '0b101' -> (5, 3); raises CcsError on the reference's errors."""
    lib = L.load()
    v, b = L.U128(), C.c_int32(0)
    L.check(lib.ccs_parse_prefix(text.encode(), C.byref(v), C.byref(b)), "ccs_parse_prefix")
    return v.lo | (v.hi << 64), b.value


def confirm_values(values, bits=0, test="bpsw", bases=(2,), nthreads=0):
    """Host confirmation engine over arbitrary n < 2^128 (ccs_confirm_values); needs no GPU.
    Returns (results, stats dict)."""
    lib = L.load()
    vals = np.zeros(len(values), np.dtype([("lo", "<u8"), ("hi", "<u8")]))
    for i, v in enumerate(values):
        vals[i] = (int(v) & ((1 << 64) - 1), int(v) >> 64)
    out = np.zeros(len(vals), CHAIN_RESULT_DTYPE)
    pp = _prove_params(bits, test, bases)
    st = L.ConfirmStats()
    L.check(lib.ccs_confirm_values(vals.ctypes.data, len(vals), C.byref(pp), nthreads, out.ctypes.data,
                                   C.byref(st)), "ccs_confirm_values")
    return out, st.as_dict()


def chain_stats(res):
    """Stats record of prove results (ccs_confirm_stats_add)."""
    lib = L.load()
    res = np.ascontiguousarray(res, CHAIN_RESULT_DTYPE)
    st = L.ConfirmStats()
    L.check(lib.ccs_confirm_stats_add(res.ctypes.data, len(res), C.byref(st)), "ccs_confirm_stats_add")
    return st.as_dict()


def format_results(res, log_threshold=8, target=18):
    """Results-file lines in the v13 format (v13.cu:2030-2033)."""
    lib = L.load()
    res = np.ascontiguousarray(res, CHAIN_RESULT_DTYPE)
    need = C.c_uint64(0)
    L.check(lib.ccs_format_results(res.ctypes.data, len(res), log_threshold, target, None, 0,
                                   C.byref(need)), "format")
    buf = C.create_string_buffer(need.value + 1)
    L.check(lib.ccs_format_results(res.ctypes.data, len(res), log_threshold, target, buf,
                                   need.value + 1, C.byref(need)), "format")
    return buf.value.decode()
