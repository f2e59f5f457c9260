"""Pins the CPU oracle (oracle/ccs_oracle.c) before the GPU tests trust it as their model.

Every check below comes from outside the oracle: the reference's own self-test vectors
and data sample (tests/golden/), counts the reference prints, the mathematics of the
predicate (q | p_i evaluated directly), and sympy's primality test.
"""
import json
import math
import os
import random

import numpy as np
import pytest
import sympy

from conftest import GOLDEN

VEC = json.load(open(os.path.join(GOLDEN, "reference_vectors.json")))
CRT = [5, 7, 11, 13, 17, 19]


def sample_roots():
    out = []
    with open(os.path.join(GOLDEN, "cc10plus_roots_sample_1000.txt")) as f:
        for line in f:
            tag, hx, _ = line.split()
            out.append((int(tag[2:]), int(hx, 16)))
    return out


def p_i(n, i):
    return (n + 1) * (1 << i) - 1


@pytest.fixture(scope="module")
def ref_shape(oracle):
    fp = [int(q) for q in oracle.primes_between(37, 863)]
    return oracle.Shape(CRT, [23, 29, 31], 18, 18, fp, 97)


# --- forbidden residues -------------------------------------------------------

@pytest.mark.parametrize("q", [5, 7, 11, 31, 37, 61, 97, 101, 863, 65521])
def test_forbidden_is_divisibility_of_chain_positions(oracle, q):
    for lo, hi in [(0, 18), (1, 18), (0, 10), (0, 40)]:
        got = set(oracle.forbidden(q, lo, hi))
        want = {r for r in range(q) if q < 5000 and any(p_i(r, i) % q == 0 for i in range(lo, min(hi, q - 1)))} \
            if q < 5000 else None
        if want is not None:
            assert got == want
        # the immune class q-1 is never forbidden; 0 is forbidden iff some tested position
        # i has 2^i = 1 (mod q): i = 0, or a multiple of the order of 2
        assert (q - 1) not in got
        order = sympy.n_order(2, q)
        assert (0 in got) == any(i % order == 0 for i in range(lo, min(hi, q - 1)))


def test_primes_between(oracle):
    assert list(oracle.primes_between(37, 97)) == list(sympy.primerange(37, 98))
    assert len(oracle.primes_between(101, 863)) == 125          # v13 LINE_SIEVE_COUNT


# --- lattice ----------------------------------------------------------------

def test_reference_shape_counts(ref_shape):
    # v13 banner: 36 bases, 3432 wheel offsets per base, 20592 candidates per tile
    assert ref_shape.M == 1616615 and ref_shape.WP == 20677
    assert ref_shape.nbases == 36
    assert all(ref_shape.wheel_count(b) == 3432 for b in range(36))
    assert ref_shape.nclasses == 123552
    assert ref_shape.count_candidates(0, 6) == 123552


def test_config_shape_counts(oracle):
    s1 = oracle.Shape(CRT, [23], 10, 10, [29], 97)
    assert s1.nclasses == 12636                                # admissible classes mod 23#
    s3 = oracle.Shape(CRT, [23, 29, 31, 37], 19, 19, [41], 97)
    assert s3.nclasses == 2021760                              # mod 37# at depth 19


def test_lattice_brute_force_small(oracle):
    # every n in a window is on the lattice iff it is reached by (tile, base, k)
    s = oracle.Shape([5, 7], [11], 4, 4, [13, 17], 97)
    W = 6 * 5 * 7 * 11
    bases = s.bases()
    reached = set()
    for b in range(s.nbases):
        for m in range(6):
            for k in s.bucket(b, m):
                for tile in range(m, 60, 6):
                    reached.add((tile * 11 + int(k)) * 35 + int(bases[b]))
    limit = 60 * 11 * 35
    direct = {n for n in range(limit) if s.on_lattice(n)}
    assert reached == direct
    # and on_lattice is the chain condition itself
    for n in range(0, 3 * W):
        ok = n % 6 == 5 and all(p_i(n, i) % q for q in (5, 7, 11) for i in range(min(4, q - 1)))
        assert s.on_lattice(n) == ok


def test_mod6_buckets_and_cc17_base(ref_shape):
    bases = [int(b) for b in ref_shape.bases()]
    # v13 self-test 15: every bucket-selected candidate is 5 mod 6
    for tile in range(60):
        for b in range(36):
            for k in ref_shape.bucket(b, tile % 6)[:5]:
                assert ((tile * 20677 + int(k)) * 1616615 + bases[b]) % 6 == 5
    # v13 self-test 18: buckets partition the wheel
    for b in range(36):
        assert sum(len(ref_shape.bucket(b, m)) for m in range(6)) == 3432
    # v13 self-test 17: CC17 root base 190189 is a base
    assert VEC["cc17_base"] in bases
    assert int(VEC["classic_chains"][-1][0]) % 1616615 == VEC["cc17_base"]


def test_q7_immunized_roots_decompose(ref_shape):
    # v13 self-test 19: four 109-bit roots decompose into (base, k, tile) of the lattice
    for hx in VEC["q7_immunized_roots"]:
        n = int(hx, 16)
        d = ref_shape.decompose(n)
        assert d is not None
        b, k, tile = d
        assert int(ref_shape.bases()[b]) % 7 == 6
        assert ref_shape.n_of(tile, [b, k, 0]) == n


def test_record_and_sample_roots_on_lattice(ref_shape, oracle):
    # both record CC18s and every root of the reference's sample lie on the lattice
    for hx, _ in VEC["v16_qvectors"][:2]:
        assert ref_shape.on_lattice(int(hx, 16))
    roots = sample_roots()
    s10 = oracle.Shape(CRT, [23, 29, 31], 10, 10, [37], 97)
    assert all(s10.on_lattice(n) for _, n in roots)


# --- acceptance predicate ------------------------------------------------------

def brute_first_kill(n, fp, depth, pos0_max_q=97):
    for idx, q in enumerate(fp):
        lo = 0 if q <= pos0_max_q else 1
        if any(p_i(n, i) % q == 0 for i in range(lo, depth)):
            return idx
    return len(fp)


def test_predicate_matches_direct_divisibility(ref_shape):
    rng = random.Random(7)
    fp = [int(q) for q in ref_shape.fp]
    bases = [int(b) for b in ref_shape.bases()]
    for _ in range(3000):
        b = rng.randrange(36)
        m = rng.randrange(6)
        ks = ref_shape.bucket(b, m)
        k = int(ks[rng.randrange(len(ks))])
        tile = rng.randrange(1 << 50) * 6 + m
        n = (tile * 20677 + k) * 1616615 + bases[b]
        want = brute_first_kill(n, fp, 18)
        assert ref_shape.first_kill(n) == want
        assert ref_shape.first_kill_brute(n) == want


def test_filter_range_against_direct_predicate(ref_shape):
    fp = [int(q) for q in ref_shape.fp]
    t0 = (1 << 88) // (20677 * 1616615) + 12345
    surv, counts = ref_shape.filter(t0, 60, cap=1 << 16, cuts=[7, 14, len(fp)])
    assert counts[0] == ref_shape.count_candidates(t0, 60) == 60 * 20592
    assert counts[3] == len(surv)
    # every survivor passes the direct predicate, in (tile, base, k) order
    keys = [(int(s[2]), int(s[0]), int(s[1])) for s in surv]
    assert keys == sorted(keys)
    for s in surv:
        assert brute_first_kill(ref_shape.n_of(t0, s), fp, 18) == len(fp)
    # stage pass rates of the reference campaign: L2 3.88 %, ext 0.80 % cumulative
    assert 0.036 < counts[1] / counts[0] < 0.042
    assert 0.0070 < counts[2] / counts[0] < 0.0090


def test_filter_exhaustive_small_window(oracle):
    # tiny shape: enumerate the window by hand and apply the direct predicate
    fp = [int(q) for q in oracle.primes_between(13, 60)]
    s = oracle.Shape([5, 7], [11], 6, 6, fp, 31)
    t0, tc = 1000, 90
    surv, counts = s.filter(t0, tc, cap=1 << 16)
    bases = [int(b) for b in s.bases()]
    want = []
    for tile in range(t0, t0 + tc):
        for b in range(s.nbases):
            for k in s.bucket(b, tile % 6):
                n = (tile * 11 + int(k)) * 35 + bases[b]
                if brute_first_kill(n, fp, 6, 31) == len(fp):
                    want.append((tile - t0, b, int(k)))
    got = [(int(x[2]), int(x[0]), int(x[1])) for x in surv]
    assert got == want
    assert counts[0] == s.count_candidates(t0, tc)


def test_sample_roots_pass_predicates(ref_shape, oracle):
    roots = sample_roots()
    fp10 = [int(q) for q in oracle.primes_between(37, 863)]
    s10 = oracle.Shape(CRT, [23, 29, 31], 10, 10, fp10, 97)
    # a CC10 root has p_0..p_9 prime, so no filter prime divides them
    assert all(s10.accept(n) for _, n in roots)
    # at depth 18 a root passes iff its positions 10..17 avoid the filter primes
    fp = [int(q) for q in ref_shape.fp]
    for _, n in roots[:200]:
        assert ref_shape.accept(n) == (brute_first_kill(n, fp, 18) == len(fp))


# --- primality and chains -------------------------------------------------------

def strong_pseudoprimes_base2(limit):
    out = []
    for n in range(9, limit, 2):
        if sympy.isprime(n):
            continue
        d, s = n - 1, 0
        while d % 2 == 0:
            d //= 2
            s += 1
        x = pow(2, d, n)
        if x == 1 or x == n - 1:
            out.append(n)
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                out.append(n)
                break
    return out


def test_is_prime_small_exhaustive(oracle):
    for n in range(0, 20000):
        assert oracle.is_prime(n) == sympy.isprime(n), n


def test_bpsw_rejects_base2_strong_pseudoprimes(oracle):
    spsp = strong_pseudoprimes_base2(400000)
    assert spsp[:4] == [2047, 3277, 4033, 4681]
    for n in spsp:
        assert oracle.mr(n, [2])            # they do fool base-2 Miller-Rabin
        assert not oracle.is_prime(n)       # and the Lucas half rejects them


def test_known_pseudoprimes_and_carmichaels(oracle):
    for c in VEC["v33_carmichael"] + [3215031751, 3474749660383, 341550071728321,
                                      3825123056546413051, 318665857834031151167461]:
        assert not oracle.is_prime(c)
    # 3474749660383 is a strong pseudoprime to bases 2..13; 3825123056546413051 to 2..23
    assert oracle.mr(3474749660383, [2, 3, 5, 7, 11, 13])
    assert oracle.mr(3825123056546413051, [2, 3, 5, 7, 11, 13, 17, 19, 23])
    for p in VEC["v33_mr_primes"]:
        assert oracle.is_prime(p) and oracle.mr(p, [2, 3, 5])
    for b, e, m, r in VEC["v33_powm"]:
        assert oracle.powmod(b, e, m) == r


def test_is_prime_random_large(oracle):
    rng = random.Random(11)
    for bits in (62, 64, 65, 80, 89, 100, 110, 120, 126):
        for _ in range(150):
            n = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
            assert oracle.is_prime(n) == sympy.isprime(n), n
        for _ in range(20):
            p = sympy.randprime(1 << (bits - 1), 1 << bits)
            assert oracle.is_prime(p)


def test_large_strong_pseudoprimes_rejected(oracle):
    # n = p(2p-1) with 2 a quadratic residue mod 2p-1 is often a base-2 strong pseudoprime
    rng = random.Random(5)
    found = 0
    while found < 12:
        p = sympy.randprime(1 << 40, 1 << 58)
        q = 2 * p - 1
        if not sympy.isprime(q):
            continue
        n = p * q
        if oracle.mr(n, [2]) and pow(2, n - 1, n) == 1:
            assert not oracle.is_prime(n)
            found += 1


@pytest.mark.parametrize("root,length", [(int(r), L) for r, L in VEC["classic_chains"]])
def test_classic_chains(oracle, root, length):
    f, r, ln, back = oracle.prove_one(root)
    if root == 89:   # 89 is one of the prover's trial primes: v13.cu:1832-1838 skips it
        assert f == oracle.F_BITS_OK
        return
    assert f & oracle.F_PRIME and not f & oracle.F_NONROOT
    assert (r, ln, back) == (root, length, 0)


def test_hex_chain_vectors(oracle):
    n, L = VEC["cc13_hex"]
    assert oracle.prove_one(int(n, 16))[1:3] == (int(n, 16), L)
    for hx, L in VEC["v16_qvectors"]:
        f, r, ln, back = oracle.prove_one(int(hx, 16))
        assert (r, ln, back) == (int(hx, 16), L, 0)
    # 135-bit CC15: beyond the u128 prover — the oracle reports overflow, not a length
    n, _ = VEC["cc15_135bit_hex"]
    assert int(n, 16).bit_length() == 135


def test_backward_walk_cc17(oracle):
    # v13 self-test 6: start from chain position 5, walk back to the CC17 root
    root = int(VEC["classic_chains"][-1][0])
    member = p_i(root, 5)
    f, r, ln, back = oracle.prove_one(member)
    assert f & oracle.F_NONROOT
    assert (r, ln, back) == (root, 17, 5)


def test_sample_roots_chain_lengths(oracle):
    roots = sample_roots()
    for L, n in roots[::5]:
        f, r, ln, back = oracle.prove_one(n)
        assert (r, ln, back) == (n, L, 0)
    # sympy agrees on a few full chains (independent primality)
    for L, n in roots[:5]:
        assert all(sympy.isprime(p_i(n, i)) for i in range(L))
        assert not sympy.isprime(p_i(n, L))
        assert not sympy.isprime((n - 1) // 2)


def test_mr_base2_mode_chain(oracle):
    # config 1 confirms with base-2 Miller-Rabin: same chains on genuine primes
    for L, n in sample_roots()[:50]:
        f, r, ln, back = oracle.prove_one(n, kind=oracle.MR, bases=(2,))
        assert (r, ln) == (n, L)
    # a base-2 strong pseudoprime without small factors is "prime" under MR-2, not BPSW
    spsp = next(n for n in strong_pseudoprimes_base2(3000000) if min(sympy.factorint(n)) > 409)
    assert oracle.prove_one(spsp, kind=oracle.MR, bases=(2,))[0] & oracle.F_PRIME
    assert not oracle.prove_one(spsp)[0] & oracle.F_PRIME


def test_prove_gates(oracle):
    root = int(VEC["v16_qvectors"][0][0], 16)
    assert oracle.prove_one(root, bits=root.bit_length())[2] == 18
    assert oracle.prove_one(root, bits=root.bit_length() + 1) == (0, 0, 0, 0)
    assert oracle.prove_one(67 * 1000003)[0] == oracle.F_BITS_OK          # trial division
    big = (1 << 126) + 1
    assert oracle.prove_one(big)[0] & oracle.F_OVERFLOW or not oracle.prove_one(big)[0] & oracle.F_PRIME


def test_survival_rate_matches_analytic(ref_shape):
    fp = [int(q) for q in ref_shape.fp]
    t0 = (1 << 88) // (20677 * 1616615) + 99
    surv, counts = ref_shape.filter(t0, 1200, cap=1 << 16, cuts=[len(fp)])
    p = 1.0
    for q in fp:
        # distinct residues (2^-i - 1) mod q over the tested positions, counted directly
        lo = 0 if q <= 97 else 1
        nf = len({(pow(2, -i, q) - 1) % q for i in range(lo, 18)})
        p *= 1 - nf / q
    assert abs(p - 1.1505e-5) < 1e-8          # the analytic survival of the reference shape
    expected = counts[0] * p
    assert abs(len(surv) - expected) < 5 * math.sqrt(expected)


@pytest.mark.parametrize("wheel,depth", [([23, 29, 31], 10), ([23, 29, 31], 18),
                                         ([23, 29, 31, 37], 10), ([23, 29, 31, 37], 18)])
@pytest.mark.parametrize("primes", ["small", "large"])
def test_filter_class_against_direct_predicate(oracle, wheel, depth, primes):
    """orc_filter_class (the class-form checker of the batched, group and served tests)
    against first_kill_brute, which evaluates q | p_i directly (itself pinned above by
    test_predicate_matches_direct_divisibility): survivors and the alive counts after each
    cut, for classes mod 6*31# and 6*37# at depths 10 and 18, in windows of 2^18 where u
    crosses 2^64 and where n crosses 2^64.  "small": the primes from the first above the
    wheel to 1500 (dense- and stage-B-sized); "large": primes near 2^16 and near 2^18
    (deep-sized tables), alone so that they do the killing."""
    import sys
    from conftest import PKG
    sys.path.insert(0, PKG)
    from ccs import workloads as WL
    if primes == "small":   # enough primes that a few dozen of 2^18 candidates survive
        fp = [int(q) for q in oracle.primes_between(wheel[-1] + 1, 1500 if depth <= 10 else 400)]
        cuts = [sum(1 for q in fp if q < 100), sum(1 for q in fp if q < 200), len(fp)]
    else:
        fp = [int(q) for q in oracle.primes_between(64000, 65535)][:24]
        fp += [int(q) for q in oracle.primes_between(261000, (1 << 18) - 1)][:24]
        cuts = [12, 24, len(fp)]
    # the class-form predicate needs only the filter primes and the depth: the oracle
    # object's own lattice is kept small (wheel 23, target 18)
    s = oracle.Shape(CRT, [23], 18, depth, fp, 97)
    W = 6 * math.prod(CRT) * math.prod(wheel)
    nfp = len(fp)
    rng = WL.XorShift64(0xC1A55 + depth + len(wheel) + len(fp))
    count = 1 << 18
    for u0_kind in ("u_crosses_2^64", "n_crosses_2^64"):
        a, Wc = WL.random_admissible_class(rng, CRT + wheel, depth)
        assert Wc == W
        if u0_kind == "u_crosses_2^64":
            u0 = (1 << 64) - count // 2
        else:
            u0 = (1 << 64) // W - count // 2
            assert (a + W * u0) < (1 << 64) < (a + W * (u0 + count))
        got, alive = s.filter_class(a, W, u0, count, cap=count, cuts=cuts)
        kills = s.first_kill_brute_class(a, W, u0, count)
        # the window's n computed in Python's integers on a sample
        for j in random.Random(depth).sample(range(count), 256):
            assert kills[j] == s.first_kill_brute(a + W * (u0 + j))
        want = np.nonzero(kills == nfp)[0]
        assert np.array_equal(got.astype(np.int64), want)
        assert alive[0] == count
        for c, cut in enumerate(cuts):
            assert alive[c + 1] == int((kills >= cut).sum()), (u0_kind, cut)
        # the window exercises every cut: some candidates die between consecutive cuts
        assert count > alive[1] > alive[2] > alive[3] > 0


@pytest.mark.parametrize("name", ["reference", "c1", "c2"])
def test_bitvector_filter_equals_per_candidate(oracle, name):
    import sys
    from conftest import PKG
    sys.path.insert(0, PKG)
    from ccs import workloads as WL
    p = WL.shape_params(name)
    s = oracle.Shape(p["crt"], p["wheel"], p["target"], p["depth"], p["filter_primes"], p["pos0_max_q"])
    t0 = WL.c1_range()[0] if name == "c1" else (1 << 88) // (s.M * s.WP) + 3
    for tc in (1, 7, 600):
        a, c = s.filter(t0, tc, cap=1 << 18)
        b, cand = s.filter_bv(t0, tc, cap=1 << 18)
        assert cand == c[0]
        assert np.array_equal(a, b)


def _lattice_n(t, k, base, WP=20677, M=1616615):
    return (t * WP + k) * M + base


@pytest.mark.parametrize("seed", range(6))
def test_tile_range_bounds_every_prefixed_candidate(oracle, ref_shape, seed):
    """compute_prefix_tile_range is conservative and tight: no candidate with the bit
    width and prefix lies in a tile outside the range, and the range is at most two tiles
    wider than the prefix's interval."""
    rnd = random.Random(seed)
    b = ref_shape.bases()
    lo_b, hi_b = int(b.min()), int(b.max())
    M, WP = ref_shape.M, ref_shape.WP
    for _ in range(200):
        bits = rnd.randint(40, 127)
        pb = rnd.randint(1, min(bits - 1, 14))
        pv = rnd.randrange(1 << (pb - 1), 1 << pb)
        prefix = "0b" + format(pv, "b")
        tmin, tmax = oracle.tile_range(M, WP, lo_b, hi_b, bits, prefix)
        shift = bits - pb
        cmin = max(pv << shift, 1 << (bits - 1))
        cmax = min(((pv + 1) << shift) - 1, (1 << bits) - 1)
        assert _lattice_n(tmin, 0, hi_b) - M < cmin or tmin == 0       # last n of tile tmin-1
        assert _lattice_n(tmax, 0, lo_b) > cmax                         # first n of tile tmax
        assert tmax - tmin <= (cmax - cmin) // (M * WP) + 3
    full = oracle.tile_range(M, WP, lo_b, hi_b, 89)
    halves = [oracle.tile_range(M, WP, lo_b, hi_b, 89, p) for p in ("0b10", "0b11")]
    assert halves[0][0] == full[0] and halves[1][1] == full[1]
    assert halves[0][1] >= halves[1][0] >= halves[0][1] - 1     # adjacent, one shared tile at most
    assert oracle.tile_range(M, WP, lo_b, hi_b, 89, "0b1") == full


def test_tile_range_lanes_partition(oracle, ref_shape):
    b = ref_shape.bases()
    M, WP = ref_shape.M, ref_shape.WP
    lo, hi = oracle.tile_range(M, WP, int(b.min()), int(b.max()), 89, "0b101")
    for lanes in (1, 3, 4, 7, 1000):
        parts = [oracle.tile_range(M, WP, int(b.min()), int(b.max()), 89, "0b101", lanes, i)
                 for i in range(lanes)]
        assert parts[0][0] == lo and parts[-1][1] == hi
        assert all(parts[i][1] == parts[i + 1][0] for i in range(lanes - 1))
        sizes = [e - s for s, e in parts]
        assert max(sizes) - min(sizes) <= 1 and sizes == sorted(sizes, reverse=True)
    for bad in [dict(bits=89, prefix="0b0"), dict(bits=89, prefix="101"), dict(bits=10),
                dict(bits=89, lanes=2), dict(bits=89, prefix="0b1", lanes=2, lane_id=2),
                dict(bits=40, prefix="0b1" + "0" * 39), dict(bits=21)]:
        with pytest.raises(ValueError):
            oracle.tile_range(M, WP, int(b.min()), int(b.max()), **bad)
