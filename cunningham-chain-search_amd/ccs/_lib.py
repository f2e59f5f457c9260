"""ctypes binding of libccs.so (the C ABI in include/ccs.h).

The library must exist (built by cunningham-chain-search_amd/build.py); there is no
fallback: every compute call goes to the HIP kernels.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libccs.so")

CCS_OK, CCS_OVERFLOW = 0, 1
CCS_TEST_BPSW, CCS_TEST_MR = 0, 1
CCS_DONE = 2
WALK_INHERIT, WALK_SEQUENTIAL, WALK_RANDOM = 0, 1, 2
CCS_SHAPE_NO_CLASSES = 1
F_BITS_OK, F_TRIAL_OK, F_PRIME, F_NONROOT, F_OVERFLOW, F_UNSUPPORTED = 1, 2, 4, 8, 16, 32

EXPORTS = [
    "ccs_ctx_create", "ccs_ctx_destroy", "ccs_device_count", "ccs_last_error",
    "ccs_dmalloc", "ccs_dfree", "ccs_memcpy_h2d", "ccs_memcpy_d2h", "ccs_memset_d",
    "ccs_synchronize", "ccs_shape_build", "ccs_shape_free", "ccs_shape_query",
    "ccs_shape_bases", "ccs_shape_classes", "ccs_shape_set_tuning", "ccs_shape_get_tuning",
    "ccs_count_candidates", "ccs_filter_launch", "ccs_filter", "ccs_last_stats",
    "ccs_prove_launch", "ccs_prove", "ccs_format_results", "ccs_batch_launch", "ccs_batch",
    "ccs_prove_jobs", "ccs_profile_start", "ccs_profile_stop",
    "ccs_jobset_create", "ccs_jobset_free", "ccs_jobset_launch", "ccs_jobset_launch_each",
    "ccs_jobset_stats", "ccs_graph_create", "ccs_graph_launch", "ccs_graph_free",
    "ccs_prove_jobset_launch", "ccs_autotune", "ccs_jobset_update",
    "ccs_server_create", "ccs_server_submit", "ccs_server_poll", "ccs_server_job_results",
    "ccs_server_drain", "ccs_server_stats", "ccs_server_shutdown", "ccs_autotune_jobs",
    "ccs_tile_range", "ccs_stream_create", "ccs_stream_destroy", "ccs_event_create",
    "ccs_event_destroy", "ccs_event_record", "ccs_stream_wait_event",
    "ccs_confirm_values", "ccs_confirm_host", "ccs_confirm_stats_add",
    "ccs_event_synchronize", "ccs_parse_prefix", "ccs_shape_decompose", "ccs_campaign_create", "ccs_campaign_create_range",
    "ccs_campaign_free", "ccs_campaign_next", "ccs_campaign_record", "ccs_campaign_get_state",
    "ccs_campaign_save", "ccs_campaign_resume", "ccs_family_shape_build", "ccs_family_window",
    "ccs_pool_parse", "ccs_pool_load", "ccs_dedup_u128", "ccs_family_top_gate_launch",
    "ccs_groupset_create", "ccs_groupset_update", "ccs_stream_status", "ccs_jobset_status",
    "ccs_debug_limit_stage_b", "ccs_chain_roots_launch", "ccs_filter_launch_noalloc",
    "ccs_debug_deep_batch",
]


class U128(C.Structure):
    _fields_ = [("lo", C.c_uint64), ("hi", C.c_uint64)]

    @staticmethod
    def of(v):
        return U128(v & ((1 << 64) - 1), v >> 64)


class ShapeParams(C.Structure):
    _fields_ = [
        ("n_crt", C.c_uint32), ("crt_primes", C.POINTER(C.c_uint32)),
        ("n_wheel", C.c_uint32), ("wheel_primes", C.POINTER(C.c_uint32)),
        ("target", C.c_int32), ("depth", C.c_int32),
        ("n_filter", C.c_uint32), ("filter_primes", C.POINTER(C.c_uint32)),
        ("n_stages", C.c_uint32), ("stage_end", C.POINTER(C.c_uint32)),
        ("pos0_max_q", C.c_uint32), ("flags", C.c_uint32),
    ]


class ShapeInfo(C.Structure):
    _fields_ = [
        ("M", C.c_uint64), ("WP", C.c_uint64), ("W", C.c_uint64),
        ("n_bases", C.c_uint32), ("n_filter", C.c_uint32), ("n_classes", C.c_uint64),
        ("n_dense", C.c_uint32), ("n_tierb", C.c_uint32), ("n_deep", C.c_uint32),
        ("rows_per_group", C.c_uint32), ("unit_j", C.c_uint32),
        ("W_hi", C.c_uint64), ("family", C.c_int32), ("exp_start", C.c_int32), ("D", U128),
        ("n_slots", C.c_uint32), ("n_tierb_lds", C.c_uint32),
    ]


class FamilyParams(C.Structure):
    _fields_ = [("D", U128), ("exp_start", C.c_int32), ("target", C.c_int32), ("depth", C.c_int32),
                ("n_filter", C.c_uint32), ("filter_primes", C.POINTER(C.c_uint32))]


class PoolEntry(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("D", U128), ("primes", C.c_uint32 * 64), ("n_primes", C.c_uint32),
                ("_pad", C.c_uint32)]

    def as_dict(self):
        return {"name": self.name.decode(), "D": self.D.lo | (self.D.hi << 64),
                "primes": [int(self.primes[i]) for i in range(self.n_primes)]}


class Tuning(C.Structure):
    _fields_ = [
        ("n_dense", C.c_uint32), ("n_tierb", C.c_uint32), ("rows_per_group", C.c_uint32),
        ("unit_j", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("pair_max", C.c_uint32),
        ("waves_per_block", C.c_uint32),
    ]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


class FilterStats(C.Structure):
    _fields_ = [
        ("candidates", C.c_uint64), ("alive_dense", C.c_uint64), ("alive_tierb", C.c_uint64),
        ("survivors", C.c_uint64), ("dense_end", C.c_uint32), ("tierb_end", C.c_uint32),
        ("ms_sieve", C.c_float), ("ms_deep", C.c_float), ("ms_prove", C.c_float),
    ]

    def as_dict(self):
        return {f: (float(getattr(self, f)) if t is C.c_float else int(getattr(self, f)))
                for f, t in self._fields_}


class LaunchStatus(C.Structure):
    _fields_ = [("overflowed_launches", C.c_uint32), ("worst_stage_b", C.c_uint32), ("dropped", C.c_uint64),
                ("stage_b_cap", C.c_uint32), ("_pad", C.c_uint32)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_ if f != "_pad"}


class ProveParams(C.Structure):
    _fields_ = [("bits", C.c_int32), ("test_kind", C.c_int32), ("n_bases", C.c_uint32),
                ("bases", C.c_uint32 * 16)]


class ConfirmStats(C.Structure):
    _fields_ = [("primes", C.c_uint64), ("chains", C.c_uint64), ("non_roots", C.c_uint64),
                ("best_chain", C.c_uint32), ("_pad", C.c_uint32), ("chain_dist", C.c_uint64 * 65)]

    def as_dict(self):
        return {"primes": int(self.primes), "chains": int(self.chains), "non_roots": int(self.non_roots),
                "best_chain": int(self.best_chain), "chain_dist": [int(x) for x in self.chain_dist]}


class Decomposition(C.Structure):
    _fields_ = [("on_lattice", C.c_uint32), ("base_found", C.c_uint32), ("class_found", C.c_uint32),
                ("bucket_ok", C.c_uint32), ("n_mod6", C.c_uint32), ("base_idx", C.c_uint32),
                ("k_offset", C.c_uint32), ("_pad", C.c_uint32), ("base", C.c_uint64), ("tile", U128)]

    def as_dict(self):
        d = {f: int(getattr(self, f)) for f, _ in self._fields_ if f not in ("_pad", "tile")}
        d["tile"] = self.tile.lo | (self.tile.hi << 64)
        return d


class CampaignParams(C.Structure):
    _fields_ = [("bits", C.c_int32), ("target", C.c_int32), ("depth", C.c_int32), ("prefix", C.c_char_p),
                ("walk", C.c_int32), ("lanes", C.c_uint32), ("lane_id", C.c_uint32), ("has_start", C.c_int32),
                ("has_seed", C.c_int32), ("seed", C.c_uint64), ("start", U128), ("batch_tiles", C.c_uint32),
                ("_pad", C.c_uint32), ("max_tiles", C.c_uint64)]


class CampaignState(C.Structure):
    _fields_ = [("tile_min", U128), ("tile_max", U128), ("current", U128), ("seed", C.c_uint64),
                ("batches", C.c_uint64), ("total_tiles", C.c_uint64), ("total_candidates", C.c_uint64),
                ("total_survivors", C.c_uint64), ("random_mode", C.c_int32), ("done", C.c_int32),
                ("chains", ConfirmStats)]

    def as_dict(self):
        d = {}
        for f, t in self._fields_:
            v = getattr(self, f)
            if t is U128:
                d[f] = v.lo | (v.hi << 64)
            elif t is ConfirmStats:
                d[f] = v.as_dict()
            else:
                d[f] = int(v)
        return d


class KernelTimes(C.Structure):
    _fields_ = [("n_sieve", C.c_uint32), ("n_deep", C.c_uint32), ("n_prove", C.c_uint32),
                ("ms_sieve", C.c_double), ("ms_deep", C.c_double), ("ms_prove", C.c_double)]

    def as_dict(self):
        return {f: (float(getattr(self, f)) if t is C.c_double else int(getattr(self, f)))
                for f, t in self._fields_}


class ServerParams(C.Structure):
    _fields_ = [("mem_budget_bytes", C.c_uint64), ("max_active", C.c_uint32), ("flags", C.c_uint32),
                ("max_slice_survivors", C.c_uint32), ("group_width_ratio", C.c_uint32), ("slice_candidates", C.c_uint64),
                ("min_chunk", C.c_uint64), ("max_chunk", C.c_uint64), ("prove", ProveParams)]


SERVER_RECOMPUTE_PHASES = 1
JOB_INCOMPLETE = 1


class Request(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("depth", C.c_int32), ("bits", C.c_int32), ("a", C.c_uint64),
                ("u0", U128), ("count", C.c_uint64)]


class JobDone(C.Structure):
    _fields_ = [("job_id", C.c_uint64), ("tag", C.c_uint64), ("candidates", C.c_uint64),
                ("n_survivors", C.c_uint32), ("n_primes", C.c_uint32), ("best_len", C.c_uint32),
                ("flags", C.c_uint32), ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServerCounters(C.Structure):
    _fields_ = [("candidates", C.c_uint64), ("slices", C.c_uint64), ("launches", C.c_uint64),
                ("resident_bytes", C.c_uint64), ("active", C.c_uint32), ("pending", C.c_uint32),
                ("job_bytes", C.c_uint64), ("fixed_bytes", C.c_uint64), ("slots", C.c_uint32),
                ("incomplete", C.c_uint32), ("busy_s", C.c_double)]

    def as_dict(self):
        return {f: (float(getattr(self, f)) if f == "busy_s" else int(getattr(self, f))) for f, _ in self._fields_}


class Job(C.Structure):
    _fields_ = [("a", C.c_uint64), ("u0", U128), ("count", C.c_uint64)]


SURVIVOR_DTYPE = np.dtype([("base_idx", "<u4"), ("k_offset", "<u4"), ("tile_idx", "<u4"), ("_pad", "<u4")])
JOB_SURVIVOR_DTYPE = np.dtype([("job_idx", "<u4"), ("u_off", "<u4")])
CHAIN_RESULT_DTYPE = np.dtype([("survivor_idx", "<u4"), ("chain_len", "<u4"), ("back_steps", "<u4"),
                               ("flags", "<u4"), ("root_lo", "<u8"), ("root_hi", "<u8")])
JOB_DTYPE = np.dtype([("a", "<u8"), ("u0_lo", "<u8"), ("u0_hi", "<u8"), ("count", "<u8")])
CHAIN_ROOT_DTYPE = np.dtype([("base_idx", "<u4"), ("k_offset", "<u4"), ("tile_idx", "<u4"), ("chain_len", "<u4")])

_lib = None


class CcsError(RuntimeError):
    pass


def load():
    """Load libccs.so; raise when it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CcsError(f"{LIB_PATH} is not built: run cunningham-chain-search_amd/build.py")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.ccs_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.ccs_ctx_destroy.argtypes = [vp]
    L.ccs_ctx_destroy.restype = None
    L.ccs_device_count.argtypes = [C.POINTER(C.c_int)]
    L.ccs_last_error.restype = C.c_char_p
    L.ccs_dmalloc.argtypes = [vp, C.POINTER(vp), C.c_uint64]
    L.ccs_dfree.argtypes = [vp, vp]
    L.ccs_memcpy_h2d.argtypes = [vp, vp, vp, C.c_uint64]
    L.ccs_memcpy_d2h.argtypes = [vp, vp, vp, C.c_uint64]
    L.ccs_memset_d.argtypes = [vp, vp, C.c_int, C.c_uint64]
    L.ccs_synchronize.argtypes = [vp]
    L.ccs_stream_create.argtypes = [vp, C.POINTER(vp)]
    L.ccs_stream_destroy.argtypes = [vp, vp]
    L.ccs_event_create.argtypes = [vp, C.POINTER(vp)]
    L.ccs_event_destroy.argtypes = [vp, vp]
    L.ccs_event_record.argtypes = [vp, vp, vp]
    L.ccs_stream_wait_event.argtypes = [vp, vp, vp]
    L.ccs_event_synchronize.argtypes = [vp, vp]
    L.ccs_shape_build.argtypes = [vp, C.POINTER(ShapeParams), C.POINTER(Tuning), C.POINTER(vp)]
    L.ccs_shape_free.argtypes = [vp]
    L.ccs_shape_free.restype = None
    L.ccs_shape_query.argtypes = [vp, C.POINTER(ShapeInfo)]
    L.ccs_shape_bases.argtypes = [vp, vp, C.c_uint32]
    L.ccs_shape_classes.argtypes = [vp, vp, vp, vp, vp, C.c_uint64]
    L.ccs_shape_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.ccs_shape_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.ccs_count_candidates.argtypes = [vp, U128, C.c_uint32]
    L.ccs_count_candidates.restype = C.c_uint64
    L.ccs_filter_launch.argtypes = [vp, vp, U128, C.c_uint32, vp, C.c_uint32, vp, vp]
    L.ccs_filter.argtypes = [vp, vp, U128, C.c_uint32, vp, C.c_uint32, C.POINTER(C.c_uint32),
                             C.POINTER(FilterStats)]
    L.ccs_last_stats.argtypes = [vp, C.POINTER(FilterStats)]
    L.ccs_stream_status.argtypes = [vp, vp, C.POINTER(LaunchStatus)]
    L.ccs_chain_roots_launch.argtypes = [vp, vp, vp, vp, C.c_uint32, vp, C.c_uint32, vp, vp, vp]
    L.ccs_filter_launch_noalloc.argtypes = [vp, vp, U128, C.c_uint32, vp, C.c_uint32, vp, vp]
    L.ccs_debug_limit_stage_b.argtypes = [vp, C.c_uint32]
    L.ccs_debug_deep_batch.argtypes = [vp, C.c_uint32]
    L.ccs_prove_launch.argtypes = [vp, vp, U128, vp, vp, C.c_uint32, C.POINTER(ProveParams), vp, vp]
    L.ccs_prove.argtypes = [vp, vp, U128, vp, C.c_uint32, C.POINTER(ProveParams), vp]
    L.ccs_confirm_values.argtypes = [vp, C.c_uint32, C.POINTER(ProveParams), C.c_uint32, vp,
                                     C.POINTER(ConfirmStats)]
    L.ccs_confirm_host.argtypes = [vp, U128, vp, C.c_uint32, C.POINTER(ProveParams), C.c_uint32, vp,
                                   C.POINTER(ConfirmStats)]
    L.ccs_confirm_stats_add.argtypes = [vp, C.c_uint32, C.POINTER(ConfirmStats)]
    L.ccs_parse_prefix.argtypes = [C.c_char_p, C.POINTER(U128), C.POINTER(C.c_int32)]
    L.ccs_shape_decompose.argtypes = [vp, U128, C.POINTER(Decomposition)]
    L.ccs_campaign_create.argtypes = [vp, C.POINTER(CampaignParams), C.POINTER(vp)]
    L.ccs_campaign_create_range.argtypes = [U128, U128, C.POINTER(CampaignParams), C.POINTER(vp)]
    L.ccs_campaign_free.argtypes = [vp]
    L.ccs_campaign_free.restype = None
    L.ccs_campaign_next.argtypes = [vp, C.POINTER(U128), C.POINTER(C.c_uint32)]
    L.ccs_campaign_record.argtypes = [vp, U128, C.c_uint32, C.c_uint64, C.c_uint64, vp, C.c_uint32]
    L.ccs_campaign_get_state.argtypes = [vp, C.POINTER(CampaignState)]
    L.ccs_campaign_save.argtypes = [vp, C.c_char_p]
    L.ccs_campaign_resume.argtypes = [vp, C.c_char_p]
    L.ccs_format_results.argtypes = [vp, C.c_uint32, C.c_int32, C.c_int32, C.c_char_p, C.c_uint64,
                                     C.POINTER(C.c_uint64)]
    L.ccs_batch_launch.argtypes = [vp, vp, vp, C.c_uint32, vp, C.c_uint32, vp, vp]
    L.ccs_batch.argtypes = [vp, vp, vp, C.c_uint32, vp, C.c_uint32, C.POINTER(C.c_uint32),
                            C.POINTER(FilterStats)]
    L.ccs_prove_jobs.argtypes = [vp, vp, vp, C.c_uint32, vp, C.c_uint32, C.POINTER(ProveParams), vp]
    L.ccs_jobset_create.argtypes = [vp, vp, vp, C.c_uint32, C.POINTER(vp)]
    L.ccs_jobset_free.argtypes = [vp]
    L.ccs_jobset_free.restype = None
    L.ccs_jobset_launch.argtypes = [vp, vp, vp, C.c_uint32, vp, vp]
    L.ccs_jobset_launch_each.argtypes = [vp, vp, vp, C.c_uint32, vp, vp]
    L.ccs_jobset_stats.argtypes = [vp, vp, vp, C.POINTER(FilterStats)]
    L.ccs_jobset_status.argtypes = [vp, vp, vp, C.POINTER(LaunchStatus)]
    L.ccs_graph_create.argtypes = [vp, vp, vp, C.c_uint32, vp, C.POINTER(vp)]
    L.ccs_graph_launch.argtypes = [vp, vp, vp]
    L.ccs_graph_free.argtypes = [vp]
    L.ccs_graph_free.restype = None
    L.ccs_prove_jobset_launch.argtypes = [vp, vp, vp, vp, C.c_uint32, C.POINTER(ProveParams), vp, vp]
    L.ccs_autotune.argtypes = [vp, vp, U128, C.c_uint32, C.c_uint32, C.POINTER(Tuning), C.POINTER(C.c_float)]
    L.ccs_tile_range.argtypes = [vp, C.c_int32, U128, C.c_int32, C.c_uint32, C.c_uint32, C.POINTER(U128),
                                 C.POINTER(U128)]
    L.ccs_autotune_jobs.argtypes = [vp, vp, vp, C.c_uint32, C.c_uint32, C.POINTER(Tuning), C.POINTER(C.c_float)]
    L.ccs_jobset_update.argtypes = [vp, vp, vp, C.c_uint32, vp]
    L.ccs_server_create.argtypes = [vp, vp, vp, C.c_uint32, C.POINTER(ServerParams), C.POINTER(vp)]
    L.ccs_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_uint64)]
    L.ccs_server_poll.argtypes = [vp, vp, C.c_uint32, C.POINTER(C.c_uint32)]
    L.ccs_server_job_results.argtypes = [vp, C.c_uint64, vp, vp, C.c_uint32, C.POINTER(C.c_uint32)]
    L.ccs_server_drain.argtypes = [vp]
    L.ccs_server_stats.argtypes = [vp, C.POINTER(ServerCounters)]
    L.ccs_server_shutdown.argtypes = [vp]
    L.ccs_family_shape_build.argtypes = [vp, C.POINTER(FamilyParams), C.POINTER(Tuning), C.POINTER(vp)]
    L.ccs_family_window.argtypes = [U128, C.c_int32, C.c_int32, C.c_int32, C.POINTER(U128), C.POINTER(U128)]
    L.ccs_pool_parse.argtypes = [C.c_char_p, vp, C.c_uint32, vp, C.c_uint32, C.POINTER(C.c_uint32)]
    L.ccs_pool_load.argtypes = [C.c_char_p, vp, C.c_uint32, vp, C.c_uint32, C.POINTER(C.c_uint32)]
    L.ccs_dedup_u128.argtypes = [vp, C.c_uint32, vp]
    L.ccs_family_top_gate_launch.argtypes = [vp, vp, vp, vp, C.c_uint32, C.c_int32, C.c_int32, vp,
                                             C.c_uint32, vp, vp]
    L.ccs_groupset_create.argtypes = [vp, vp, C.c_uint32, vp, vp, C.c_uint32, C.POINTER(vp)]
    L.ccs_groupset_update.argtypes = [vp, vp, vp, vp, C.c_uint32, vp]
    L.ccs_profile_start.argtypes = [vp]
    L.ccs_profile_stop.argtypes = [vp, C.POINTER(KernelTimes)]
    _lib = L
    return L


def check(rc, what=""):
    if rc < 0:
        msg = load().ccs_last_error().decode(errors="replace")
        raise CcsError(f"{what} failed ({rc}): {msg}")
    return rc
