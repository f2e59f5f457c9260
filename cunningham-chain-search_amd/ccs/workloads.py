"""Workload definitions: the reference campaign shape and the plain / batched / served
search configurations, with seeded synthetic ranges.

Shapes follow the reference lattice n = (tile*WP + k)*M + base (v13.cu:155-201):
CRT primes 5..19 (M = 1 616 615), wheel primes above them, mod-6 rule.  Random ranges
use xorshift64 (the reference's tile RNG, v13.cu:710-723) so a seed fixes the input.
"""
import math

MASK64 = (1 << 64) - 1
CRT = [5, 7, 11, 13, 17, 19]


def primes_between(lo, hi):
    """primes in [lo, hi]"""
    if hi < 2:
        return []
    sieve = bytearray([1]) * (hi + 1)
    sieve[0:2] = b"\x00\x00"
    for i in range(2, int(hi ** 0.5) + 1):
        if sieve[i]:
            sieve[i * i:: i] = bytearray(len(range(i * i, hi + 1, i)))
    return [i for i in range(max(lo, 2), hi + 1) if sieve[i]]


class XorShift64:
    """xorshift64 as in v13.cu:710-723 (rng_next)."""

    def __init__(self, seed):
        self.s = (seed & MASK64) or 1

    def next(self):
        x = self.s
        x ^= (x << 13) & MASK64
        x ^= x >> 7
        x ^= (x << 17) & MASK64
        self.s = x
        return x

    def below(self, n):
        """uniform integer in [0, n) for any n (rejection on a wide draw)"""
        bits = max(n.bit_length(), 1)
        words = (bits + 63) // 64 + 1
        while True:
            v = 0
            for _ in range(words):
                v = (v << 64) | self.next()
            lim = (1 << (64 * words)) - ((1 << (64 * words)) % n)
            if v < lim:
                return v % n

    def uniform(self):
        return (self.next() >> 11) / float(1 << 53)


def forbidden(q, lo, hi):
    """(2^-i - 1) mod q for i in [lo, min(hi, q-1)) (v13.cu:771-787)"""
    out = []
    for i in range(lo, min(hi, q - 1)):
        r = (pow(pow(2, i, q), -1, q) - 1) % q
        if r not in out:
            out.append(r)
    return out


def shape_params(name):
    """Parameters of a named shape: crt, wheel, target, depth, filter primes, stage ends."""
    if name == "reference":
        # v13 campaign shape: L2 37..61, ext-L2 67..97, line sieve 101..863 (v13.cu:155-199)
        fp = primes_between(37, 863)
        return dict(crt=CRT, wheel=[23, 29, 31], target=18, depth=18, filter_primes=fp,
                    stage_end=[7, 14, len(fp)], pos0_max_q=97)
    if name == "c1":
        # M = 23#, depth 10, the next 500 primes after 23
        fp = primes_between(29, 4000)[:500]
        return dict(crt=CRT, wheel=[23], target=10, depth=10, filter_primes=fp,
                    stage_end=[sum(1 for q in fp if q <= 1021), len(fp)], pos0_max_q=97)
    if name == "c2":
        # M = 31#, depth 18, stages 37..2^10, 2^10..2^14, 2^14..2^16
        fp = primes_between(37, 65535)
        return dict(crt=CRT, wheel=[23, 29, 31], target=18, depth=18, filter_primes=fp,
                    stage_end=[sum(1 for q in fp if q < 1024), sum(1 for q in fp if q < 16384), len(fp)],
                    pos0_max_q=97)
    if name == "c3":
        # M = 37#, depth 19, filter primes to 2^18
        fp = primes_between(41, (1 << 18) - 1)
        return dict(crt=CRT, wheel=[23, 29, 31, 37], target=19, depth=19, filter_primes=fp,
                    stage_end=[sum(1 for q in fp if q < 1024), sum(1 for q in fp if q < 16384), len(fp)],
                    pos0_max_q=97)
    if name == "c4":
        # batched: classes mod 37# at depth 18, C2 stages from 41
        fp = primes_between(41, 65535)
        return dict(crt=CRT, wheel=[23, 29, 31, 37], target=18, depth=18, filter_primes=fp,
                    stage_end=[sum(1 for q in fp if q < 1024), sum(1 for q in fp if q < 16384), len(fp)],
                    pos0_max_q=97)
    if name.startswith("c5d"):
        # served: classes mod 31#, depth from the name ("c5d12".."c5d20")
        d = int(name[3:])
        fp = primes_between(37, 65535)
        return dict(crt=CRT, wheel=[23, 29, 31], target=d, depth=d, filter_primes=fp,
                    stage_end=[sum(1 for q in fp if q < 1024), sum(1 for q in fp if q < 16384), len(fp)],
                    pos0_max_q=97)
    raise KeyError(name)


def make_shape(ctx, name, classes=True, tuning=None, **over):
    from . import Shape
    p = shape_params(name)
    p.update(over)
    return Shape(ctx, p["crt"], p["wheel"], p["target"], p["depth"], p["filter_primes"],
                 p["stage_end"], p["pos0_max_q"], classes=classes, tuning=tuning)


def tile_range_for_bits(M, WP, bits_lo, bits_hi, n_candidates, cand_per_tile, seed):
    """Seeded tile range: n_start uniform in [2^bits_lo, 2^bits_hi), tile_start =
    floor(n_start / (WP*M)), tile_count covering ~n_candidates (a multiple of 6)."""
    rng = XorShift64(seed)
    n_start = (1 << bits_lo) + rng.below((1 << bits_hi) - (1 << bits_lo))
    tile_start = n_start // (WP * M)
    tile_count = 6 * math.ceil(n_candidates / (6 * cand_per_tile))
    return tile_start, tile_count, n_start


def c2_range(seed=0xC2, n_candidates=1 << 36):
    """Config 2: 2^36 candidates per launch from a seeded start in [2^86, 2^89)."""
    M, WP = 1616615, 23 * 29 * 31
    return tile_range_for_bits(M, WP, 86, 89, n_candidates, 123552 / 6, seed)


def c1_range(seed=0xC1, t_count=1 << 24):
    """Config 1: t covers t_count values from a seeded start with p in [2^59, 2^60);
    tiles [6 t0, 6 (t0 + t_count))."""
    W = 223092870
    rng = XorShift64(seed)
    lo = -(-(1 << 59) // W)
    hi = ((1 << 60) - W) // W - t_count
    t0 = lo + rng.below(hi - lo)
    return 6 * t0, 6 * t_count, t0


def c3_range(seed=0xC3, n_candidates=1 << 38):
    M, WP = 1616615, 23 * 29 * 31 * 37
    return tile_range_for_bits(M, WP, 88, 92, n_candidates, 2021760 / 6, seed)


def reference_range(seed=0x13, bits=89, n_tiles=16384):
    """The v13 campaign batch: 16 384 tiles at --bits (v13.cu:3078-3082, :3480-3500)."""
    M, WP = 1616615, 23 * 29 * 31
    rng = XorShift64(seed)
    lo = (1 << (bits - 1)) // (WP * M) + 1
    hi = ((1 << bits) // (WP * M)) - n_tiles - 1
    return lo + rng.below(hi - lo), n_tiles


def random_admissible_class(rng, primes, depth):
    """A residue class mod 6*prod(primes) chosen uniformly per prime among the residues
    admissible at `depth` (no p_i = 0 mod q for i < depth), with n = 5 (mod 6); CRT-combined."""
    mod, a = 6, 5
    for q in primes:
        bad = set(forbidden(q, 0, depth))
        ok = [r for r in range(q) if r not in bad]
        r = ok[rng.below(len(ok))]
        # combine a (mod mod) with r (mod q)
        t = ((r - a) * pow(mod, -1, q)) % q
        a = a + mod * t
        mod *= q
    return a, mod


def batched_jobs(n_jobs=4096, seed=0xC4, lo_exp=20, hi_exp=32, bits_lo=84, bits_hi=90, depth=18):
    """Config 4: n_jobs admissible classes mod 37# at depth 18, log-uniform lengths in
    [2^lo_exp, 2^hi_exp], candidates in [2^bits_lo, 2^bits_hi).  Returns [(a, u0, count)]."""
    rng = XorShift64(seed)
    primes = [5, 7, 11, 13, 17, 19, 23, 29, 31, 37]
    jobs = []
    for _ in range(n_jobs):
        a, W = random_admissible_class(rng, primes, depth)
        count = int(round(2 ** (lo_exp + (hi_exp - lo_exp) * rng.uniform())))
        count = max(1, min(count, 1 << hi_exp))
        umin = -(-((1 << bits_lo) - a) // W)
        umax = ((1 << bits_hi) - a) // W - count
        u0 = umin + rng.below(max(umax - umin, 1))
        jobs.append((a, u0, count))
    return jobs, W


def served_requests(n_jobs=10000, seed=0xC5, lo_exp=24, hi_exp=34):
    """Config 5: depth U{12..20}, width U{64,80,96}, admissible class mod 31#,
    log-uniform length in [2^lo_exp, 2^hi_exp].  Returns dicts."""
    rng = XorShift64(seed)
    primes = [5, 7, 11, 13, 17, 19, 23, 29, 31]
    out = []
    for i in range(n_jobs):
        depth = 12 + rng.below(9)
        bits = (64, 80, 96)[rng.below(3)]
        a, W = random_admissible_class(rng, primes, depth)
        count = int(round(2 ** (lo_exp + (hi_exp - lo_exp) * rng.uniform())))
        umin = -(-((1 << (bits - 1)) - a) // W)
        umax = ((1 << bits) - a) // W - count
        u0 = umin + rng.below(max(umax - umin, 1))
        out.append(dict(id=i, depth=depth, bits=bits, a=a, W=W, u0=u0, count=count))
    return out


# One MI355X (gfx950): 256 CUs at 2.4 GHz.  A CU issues one wave64 VALU instruction per
# clock (four SIMD16s, four clocks each), i.e. 64 int32 lane-ops per clock, and its LDS
# returns 256 B per clock to 128-bit reads (the LDS throughput probe, lds_probe, reads
# 118 TB/s with ds_read_b128 over the whole chip, above the 78.6 TB/s a 128 B/clk port
# would allow).
MI355X_CUS = 256
MI355X_LANES_PER_CU = 64
MI355X_CLK_GHZ = 2.4
MI355X_LDS_BYTES_PER_CLK = 256
INT32_PEAK_TOPS = MI355X_CUS * MI355X_LANES_PER_CU * MI355X_CLK_GHZ * 1e9 / 1e12   # 39.3
LDS_PEAK_GBS = MI355X_CUS * MI355X_LDS_BYTES_PER_CLK * MI355X_CLK_GHZ                # 157 286

# The per-candidate figure of the reference's predicate (one residue test per prime until
# the first kill, 4 int32 ops each): kept as "equivalent mod-ops", not as a roofline.
INT32_OPS_PER_TEST = 4

# The sieve kernel's algorithmic work per candidate, by stage:
#  dense stage, per pattern (slot) and 128 candidates of a lane: two aligned 16-byte LDS
#    blocks (32 B), four v_alignbit (the 128-bit window) and two v_or3 (one 3-input OR
#    folds two patterns into a word), i.e. 0.25 B and 6/128 lane-ops per candidate;
#    the row's live words: a complement and a non-zero test per 32 candidates;
#  stage B, per candidate alive after the dense stage and group of 64 primes probed: every
#    lane tests its prime (phase add, f32-reciprocal quotient cvt/mul/cvt, v_mad_i32_i24
#    remainder, word index shift/and/add, bit extract, compare = 10 lane-ops) with one
#    4-byte bitset read (LDS for the groups held there, L2 for the rest).
DENSE_LDS_BYTES_PER_SLOT = 32 / 128
DENSE_VALU_PER_SLOT = 6 / 128
ROW_VALU_PER_CANDIDATE = 2 / 32
STAGE_B_VALU_PER_PROBE = 10 * 64
STAGE_B_BYTES_PER_PROBE = 4 * 64


def expected_tests(fp_sizes):
    """E[tests] under ascending-prime early exit: sum over primes of P(alive before q)."""
    alive, e = 1.0, 0.0
    for q, nf in fp_sizes:
        e += alive
        alive *= 1.0 - nf / q
    return e, alive


def sieve_work_model(filter_primes, depth, pos0_max_q, n_dense, n_slots, n_tierb, n_tierb_lds):
    """Algorithmic VALU lane-ops and LDS bytes per candidate of the sieve kernel for a
    shape split as (n_dense primes in n_slots patterns, n_tierb stage-B primes of which the
    first n_tierb_lds have their bitsets in LDS), with the survival fractions from the
    forbidden-residue sets (uniform residues)."""
    alive, p_dense, groups, lds_groups = 1.0, None, 0.0, 0.0
    for i, q in enumerate(filter_primes[: n_dense + n_tierb]):
        if i == n_dense:
            p_dense = alive
        if i >= n_dense and (i - n_dense) % 64 == 0:   # a stage-B group starts: probed if alive
            groups += alive
            if i - n_dense < n_tierb_lds:
                lds_groups += alive
        lo = 0 if q <= pos0_max_q else 1
        alive *= 1.0 - len(forbidden(q, lo, depth)) / q
    if p_dense is None:
        p_dense = alive
    valu = n_slots * DENSE_VALU_PER_SLOT + ROW_VALU_PER_CANDIDATE + groups * STAGE_B_VALU_PER_PROBE
    lds = n_slots * DENSE_LDS_BYTES_PER_SLOT + lds_groups * STAGE_B_BYTES_PER_PROBE
    return {"valu_lane_ops": valu, "lds_bytes": lds, "p_dense": p_dense,
            "stage_b_probes": groups, "stage_b_lds_probes": lds_groups, "p_stage_b": alive,
            "dense_lds_reads_b128": n_slots * 2 / 8192, "stage_b_lds_reads": lds_groups}
