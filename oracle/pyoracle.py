"""ctypes binding of the CPU oracle (oracle/ccs_oracle.c).

Test infrastructure only: imported by tests/, the smoke check and the CPU-baseline
leg of bench.py, never by the product.  Each wrapper names the C function it calls;
those name the reference file:line they restate.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libccs_oracle.so")
_lib = None

u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

BPSW, MR = 0, 1


def default_threads():
    """worker threads: the process's CPU affinity, at most 16 (one GPU's CPU share)"""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


F_BITS_OK, F_TRIAL_OK, F_PRIME, F_NONROOT, F_OVERFLOW = 1, 2, 4, 8, 16


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    L.orc_forbidden.argtypes = [C.c_uint32, C.c_int, C.c_int, u32p]
    L.orc_forbidden.restype = C.c_int
    L.orc_primes_between.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_int]
    L.orc_primes_between.restype = C.c_int
    L.orc_shape_new.argtypes = [u32p, C.c_int, u32p, C.c_int, C.c_int, C.c_int, u32p, C.c_int, C.c_uint32]
    L.orc_shape_new.restype = C.c_void_p
    L.orc_shape_free.argtypes = [C.c_void_p]
    for name in ("orc_shape_nbases",):
        getattr(L, name).argtypes = [C.c_void_p]
        getattr(L, name).restype = C.c_int
    for name in ("orc_shape_nclasses", "orc_shape_M", "orc_shape_WP"):
        getattr(L, name).argtypes = [C.c_void_p]
        getattr(L, name).restype = C.c_uint64
    L.orc_shape_bases.argtypes = [C.c_void_p, u64p]
    L.orc_shape_wheel_count.argtypes = [C.c_void_p, C.c_int]
    L.orc_shape_wheel_count.restype = C.c_uint32
    L.orc_shape_bucket.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    L.orc_shape_bucket.restype = C.c_uint32
    L.orc_shape_classes.argtypes = [C.c_void_p, u32p, u32p, u8p]
    for name in ("orc_first_kill", "orc_first_kill_brute", "orc_on_lattice"):
        getattr(L, name).argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
        getattr(L, name).restype = C.c_int
    L.orc_first_kill_brute_class.argtypes = [C.c_void_p] + [C.c_uint64] * 7 + [i32p]
    L.orc_decompose.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint32),
                                C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.orc_decompose.restype = C.c_int
    L.orc_filter.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, u32p, C.c_int64,
                             i32p, C.c_int, u64p, C.c_int]
    L.orc_filter.restype = C.c_int64
    L.orc_count_candidates.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64]
    L.orc_count_candidates.restype = C.c_uint64
    L.orc_is_prime.argtypes = [C.c_uint64, C.c_uint64]
    L.orc_is_prime.restype = C.c_int
    L.orc_mr.argtypes = [C.c_uint64, C.c_uint64, u32p, C.c_int]
    L.orc_mr.restype = C.c_int
    L.orc_powmod.argtypes = [C.c_uint64] * 6 + [C.POINTER(C.c_uint64)] * 2
    L.orc_prove_one.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_int, u32p, C.c_int,
                                C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.orc_prove_one.restype = C.c_uint32
    L.orc_prove.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, u32p, C.c_int64, C.c_int, C.c_int,
                            u32p, C.c_int, u64p, C.c_int]
    L.orc_filter_class.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                   C.c_uint64, u32p, C.c_int64, C.c_void_p, C.c_int, C.c_void_p]
    L.orc_filter_class.restype = C.c_int64
    L.orc_filter_family.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_int, u32p, C.c_int, C.c_uint64,
                                    C.c_uint64, C.c_uint64, u32p, C.c_int64]
    L.orc_filter_family.restype = C.c_int64
    L.orc_filter_bv.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, u32p, C.c_int64,
                                C.c_int, C.POINTER(C.c_uint64), C.c_int]
    L.orc_filter_bv.restype = C.c_int64
    L.orc_bv_prepare.argtypes = [C.c_void_p, C.c_int]
    _lib = L
    return L


def split(n):
    return n & ((1 << 64) - 1), n >> 64


def forbidden(q, pos_lo, pos_hi):
    out = np.zeros(64, np.uint32)
    k = lib().orc_forbidden(q, pos_lo, pos_hi, out)
    return [int(x) for x in out[:k]]


def primes_between(lo, hi):
    n = lib().orc_primes_between(lo, hi, None, 0)
    out = np.zeros(max(n, 1), np.uint32)
    lib().orc_primes_between(lo, hi, out.ctypes.data, n)
    return out[:n].copy()


def is_prime(n):
    return bool(lib().orc_is_prime(*split(n)))


def mr(n, bases):
    b = np.asarray(bases, np.uint32)
    return bool(lib().orc_mr(*split(n), b, len(b)))


def powmod(b, e, n):
    lo, hi = C.c_uint64(), C.c_uint64()
    lib().orc_powmod(*split(b), *split(e), *split(n), C.byref(lo), C.byref(hi))
    return lo.value | (hi.value << 64)


def prove_one(n, bits=0, kind=BPSW, bases=(2,)):
    b = np.asarray(bases, np.uint32)
    rl, rh, ln, bk = C.c_uint64(), C.c_uint64(), C.c_uint32(), C.c_uint32()
    f = lib().orc_prove_one(*split(n), bits, kind, b, len(b), C.byref(rl), C.byref(rh),
                            C.byref(ln), C.byref(bk))
    return int(f), rl.value | (rh.value << 64), ln.value, bk.value


class Shape:
    """Oracle shape: n = (tile*WP + k)*M + base with the mod-6 rule (v13.cu:907-1059),
    filter primes fp tested at positions [0|1, depth) (v13.cu:793-901)."""

    def __init__(self, crt, wheel, target, depth, fp, pos0_max_q=97):
        self.crt = np.asarray(crt, np.uint32)
        self.wheel = np.asarray(wheel, np.uint32)
        self.fp = np.asarray(fp, np.uint32)
        self.target, self.depth, self.pos0_max_q = target, depth, pos0_max_q
        self.h = lib().orc_shape_new(self.crt, len(self.crt), self.wheel, len(self.wheel),
                                     target, depth, self.fp, len(self.fp), pos0_max_q)
        self.M = int(lib().orc_shape_M(self.h))
        self.WP = int(lib().orc_shape_WP(self.h))
        self.nbases = lib().orc_shape_nbases(self.h)
        self.nclasses = int(lib().orc_shape_nclasses(self.h))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_shape_free(self.h)
            self.h = None

    def bases(self):
        out = np.zeros(self.nbases, np.uint64)
        lib().orc_shape_bases(self.h, out)
        return out

    def wheel_count(self, b):
        return int(lib().orc_shape_wheel_count(self.h, b))

    def bucket(self, b, m):
        n = lib().orc_shape_bucket(self.h, b, m, None)
        out = np.zeros(max(n, 1), np.uint32)
        lib().orc_shape_bucket(self.h, b, m, out.ctypes.data)
        return out[:n].copy()

    def classes(self):
        b = np.zeros(self.nclasses, np.uint32)
        k = np.zeros(self.nclasses, np.uint32)
        r = np.zeros(self.nclasses, np.uint8)
        lib().orc_shape_classes(self.h, b, k, r)
        return b, k, r

    def first_kill(self, n):
        return lib().orc_first_kill(self.h, *split(n))

    def first_kill_brute(self, n):
        return lib().orc_first_kill_brute(self.h, *split(n))

    def first_kill_brute_class(self, a, W, u0, count):
        """first_kill_brute of n = a + W*(u0 + j) for j in [0, count) (int32 array)"""
        out = np.zeros(max(count, 1), np.int32)
        lib().orc_first_kill_brute_class(self.h, *split(a), *split(W), *split(u0), count, out)
        return out[:count]

    def accept(self, n):
        return self.first_kill(n) == len(self.fp)

    def on_lattice(self, n):
        return bool(lib().orc_on_lattice(self.h, *split(n)))

    def decompose(self, n):
        k, tl, th = C.c_uint32(), C.c_uint64(), C.c_uint64()
        b = lib().orc_decompose(self.h, *split(n), C.byref(k), C.byref(tl), C.byref(th))
        if b < 0:
            return None
        return b, k.value, tl.value | (th.value << 64)

    def count_candidates(self, tile_start, tile_count):
        return int(lib().orc_count_candidates(self.h, *split(tile_start), tile_count))

    def filter(self, tile_start, tile_count, cap=1 << 20, cuts=(), nthreads=None):
        """Survivors as an (n, 4) uint32 array of {base_idx, k_offset, tile_idx, 0}
        in (tile, base, k) order, plus the stage counters."""
        if nthreads is None:
            nthreads = default_threads()
        out = np.zeros((cap, 4), np.uint32)
        cut = np.asarray(list(cuts) or [0], np.int32)
        counts = np.zeros(len(cut) + 1, np.uint64)
        n = lib().orc_filter(self.h, *split(tile_start), tile_count, out.reshape(-1), cap, cut,
                             len(cuts), counts, nthreads)
        if n > cap:
            raise RuntimeError(f"oracle survivor overflow: {n} > {cap}")
        return out[:n].copy(), [int(x) for x in counts[: len(cuts) + 1]]

    def filter_bv(self, tile_start, tile_count, cap=1 << 20, nbv=None, nthreads=None):
        """v34-style 64-candidate bit-vector filter over the primes <= 97 (or the first nbv),
        per-candidate for the rest: the same survivors as filter(), and the candidate count."""
        if nthreads is None:
            nthreads = default_threads()
        if nbv is None:
            nbv = int(sum(1 for q in self.fp if q <= 97))
        lib().orc_bv_prepare(self.h, nbv)
        out = np.zeros((cap, 4), np.uint32)
        cand = C.c_uint64(0)
        n = lib().orc_filter_bv(self.h, *split(tile_start), tile_count, out.reshape(-1), cap, nbv,
                                C.byref(cand), nthreads)
        if n > cap:
            raise RuntimeError(f"oracle survivor overflow: {n} > {cap}")
        return out[:n].copy(), cand.value

    def filter_class(self, a, W, u0, count, cap=1 << 20, cuts=()):
        out = np.zeros(cap, np.uint32)
        cut = np.asarray(list(cuts) or [0], np.int32)
        alive = np.zeros(len(cut) + 1, np.uint64)
        n = lib().orc_filter_class(self.h, a, W, *split(u0), count, out, cap, alive.ctypes.data,
                                   len(cuts), cut.ctypes.data)
        if n > cap:
            raise RuntimeError("oracle survivor overflow")
        return out[:n].copy(), [int(x) for x in alive[: len(cuts) + 1]]

    def prove(self, tile_start, surv, bits=0, kind=BPSW, bases=(2,), nthreads=None):
        """Per survivor (flags, len, back, 0, root_lo, root_hi), restating prover_worker."""
        if nthreads is None:
            nthreads = default_threads()
        surv = np.ascontiguousarray(surv, np.uint32)
        out = np.zeros((len(surv), 6), np.uint64)
        b = np.asarray(bases, np.uint32)
        lib().orc_prove(self.h, *split(tile_start), surv.reshape(-1), len(surv), bits, kind, b, len(b),
                        out.reshape(-1), nthreads)
        return out

    def n_of(self, tile_start, surv_row):
        b, k, t = int(surv_row[0]), int(surv_row[1]), int(surv_row[2])
        return ((tile_start + t) * self.WP + k) * self.M + int(self.bases()[b])


def tile_range(M, WP, min_base, max_base, bits, prefix=None, lanes=0, lane_id=0):
    """[tile_min, tile_max) of a campaign: the --bits range with floor k bounds
    (v13.cu:3480-3491), narrowed to a binary prefix "0b..." by compute_prefix_tile_range
    (v13.cu:450-518) and clamped to the full range (v13.cu:3524-3530), then split into
    prefix lanes, the first (total mod lanes) lanes one tile longer (v13.cu:3532-3550).
    Raises ValueError where the reference prints an error and exits."""
    if not 21 <= bits <= 127:
        raise ValueError("bits")
    n_min, n_max = 1 << (bits - 1), (1 << bits) - 1
    if n_min <= max_base:    # the reference's unsigned subtraction would wrap here
        raise ValueError("bits too small for the lattice bases")
    tmin = ((n_min - max_base) // M) // WP
    tmax = ((n_max - min_base) // M) // WP + 1
    if prefix is None:
        if lanes:
            raise ValueError("--prefix-lanes requires --prefix")
        return tmin, tmax
    if prefix[:2] not in ("0b", "0B") or len(prefix) < 3 or set(prefix[2:]) - {"0", "1"} or prefix[2] != "1":
        raise ValueError("binary prefix")          # parse_prefix (v13.cu:408-448)
    pv, pb = int(prefix[2:], 2), len(prefix) - 2
    if pb >= bits:
        raise ValueError("prefix_bits must be < bits")   # v13.cu:3181-3185
    shift = bits - pb
    cmin, cmax = pv << shift, ((pv + 1) << shift) - 1
    if cmax < n_min or cmin > n_max:
        raise ValueError("prefix does not overlap the bit range")
    cmin, cmax = max(cmin, n_min), min(cmax, n_max)
    kmin = -(-(cmin - max_base) // M) if cmin > max_base else 0
    kmax = (cmax - min_base) // M
    if kmin > kmax:
        raise ValueError("empty k range")
    pmin, pmax = max(kmin // WP, tmin), min(kmax // WP + 1, tmax)
    if pmin >= pmax:
        raise ValueError("empty tile range")
    if lanes:
        if not 0 <= lane_id < lanes:
            raise ValueError("lane id")
        size, rem = divmod(pmax - pmin, lanes)
        start = pmin + lane_id * size + min(lane_id, rem)
        return start, start + size + (1 if lane_id < rem else 0)
    return pmin, pmax


def filter_family(D, E, depth, primes, q0, count, cap=1 << 22):
    """Survivor offsets Q - q0 of the family p_j = Q*D*2^(E+j) - 1 (orc_filter_family)."""
    pr = np.ascontiguousarray(np.asarray(list(primes), np.uint32))
    out = np.zeros(cap, np.uint32)
    n = lib().orc_filter_family(*split(D), E, depth, pr, len(pr), *split(q0), count, out, cap)
    if n > cap:
        raise RuntimeError("oracle survivor overflow")
    return out[:n].copy()
