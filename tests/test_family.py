"""Fingerprint-family front end on the host (no GPU): pool lists and pool files
(v16...cu:556-738, :828-880), Q windows (v16...cu:4405-4468), p_0 dedup (:2478-2541), the
ABI structs, and the oracle's family predicate.

The pool fixtures under tests/golden/pools/ are data files of the reference
(src/cuda/experimental/pools/known_fingerprints*.txt; the exp_d and primorial ones cut to
their first lines).  Each line's D is checked against the line itself (the product of the
P^k factors written on it) and against the log2(D) the file's generator wrote beside it."""
import ctypes as C
import itertools
import math
import os
import random
import re
import subprocess

import pytest

import ccs
from ccs import _lib
from conftest import GOLDEN, ROOT

POOLS = os.path.join(GOLDEN, "pools")


def primes_upto(n):
    return [q for q in range(2, n + 1) if all(q % d for d in range(2, int(q ** 0.5) + 1))]


def spec_value(spec):
    """D of a 'P^k*Q*...' line, evaluated here from the text."""
    D = 1
    for tok in spec.split("*"):
        p, _, k = tok.partition("^")
        D *= int(p) ** int(k or 1)
    return D


def pool_lines(name):
    for line in open(os.path.join(POOLS, name)):
        body, _, comment = line.partition("#")
        body = body.strip()
        if body:
            yield body, comment


@pytest.mark.parametrize("name", ["known_fingerprints_top10.txt", "known_fingerprints.txt",
                                  "known_fingerprints_exp_d_head160.txt",
                                  "known_fingerprints_primorial_head120.txt"])
def test_pool_files(name):
    got = ccs.pool_load(os.path.join(POOLS, name))
    lines = list(pool_lines(name))
    assert len(got) == len(lines) > 5
    for e, (body, comment) in zip(got, lines):
        assert e["name"] == body
        assert e["D"] == spec_value(body)
        assert math.prod(e["primes"]) == e["D"]
        assert e["primes"] == sorted(e["primes"])
        m = re.search(r"bits=([0-9.]+)", comment)
        if m and "." in m.group(1):       # generator's log2(D), two decimals
            assert abs(math.log2(e["D"]) - float(m.group(1))) < 0.006, (body, m.group(1))
        elif m:                            # primorial file: integer bits
            assert e["D"].bit_length() in (int(m.group(1)), int(m.group(1)) + 1)


def test_primorial_specs_and_omits():
    got = ccs.pool_parse("53#, 53#/7/11, 61# / 5, 10#")
    assert [e["name"] for e in got] == ["53#", "53#/7/11", "61# / 5", "10#"]
    assert got[0]["primes"] == primes_upto(53)
    assert got[1]["primes"] == [q for q in primes_upto(53) if q not in (7, 11)]
    assert got[2]["primes"] == [q for q in primes_upto(61) if q != 5]
    assert got[3]["D"] == 2 * 3 * 5 * 7                   # Z need not be prime
    assert got[1]["D"] == math.prod(got[1]["primes"])


def test_power_lists():
    e, = ccs.pool_parse("2*3^2*5*11^2*13*17*19*37")
    assert e["primes"] == [2, 3, 3, 5, 11, 11, 13, 17, 19, 37]
    assert e["D"] == 2 * 9 * 5 * 121 * 13 * 17 * 19 * 37
    e, = ccs.pool_parse("7^16")
    assert e["D"] == 7 ** 16 and len(e["primes"]) == 16


def test_auto_expansion():
    keep = [2, 3, 5]
    got = ccs.pool_parse("auto:31#:1-2", keep=keep)
    drop = [q for q in primes_upto(31) if q not in keep]
    want = [c for k in (1, 2) for c in itertools.combinations(drop, k)]
    assert len(got) == len(want) == 8 + 28
    for e, c in zip(got, want):
        assert e["name"] == "31#" + "".join(f"/{q}" for q in c)
        assert e["primes"] == [q for q in primes_upto(31) if q not in c]
    assert ccs.pool_parse("auto:13#:0", keep=keep)[0]["D"] == 30030
    # K above the droppable count is clamped (emit nothing beyond all-dropped)
    assert len(ccs.pool_parse("auto:7#:5", keep=[2])) == 0
    assert len(ccs.pool_parse("auto:7#:1-9", keep=[2])) == 3 + 3 + 1


def test_pool_errors():
    for bad in ["600#", "1#", "53#/4", "53#/x", "2*4*5", "2^17", "2^0", "3*^2", "auto:13:1", "auto:13#1",
                "auto:13#:2-1", "2^16*3^16*5^16*7^16*11"]:
        with pytest.raises(ccs.CcsError):
            ccs.pool_parse(bad)
    with pytest.raises(ccs.CcsError):
        ccs.pool_load(os.path.join(POOLS, "does_not_exist.txt"))
    # D above 2^128 overflows
    with pytest.raises(ccs.CcsError):
        ccs.pool_parse("*".join(str(q) for q in primes_upto(109)))


def test_pool_file_comment_rule(tmp_path):
    # from the first '#' of a line on is a comment, so a primorial spec cannot sit in a file:
    # "53#/7" reads as the one-prime list "53" (parse_seed_pool_file, v16...cu:697-738)
    p = tmp_path / "pool.txt"
    p.write_text("# header\n\n  2*3*5 , 7*11  # freq=1\n53#/7\n\t13\n")
    got = ccs.pool_load(str(p))
    assert [e["D"] for e in got] == [30, 77, 53, 13]


def test_windows_match_integer_formula():
    rng = random.Random(0xF16)
    for _ in range(400):
        D = rng.randrange(1, 1 << rng.randrange(1, 100))
        E = rng.randrange(0, 8)
        lo = rng.randrange(2, 128)
        hi = rng.randrange(lo, 128)
        if (D << E) >= 1 << 127:
            with pytest.raises(ccs.CcsError):
                ccs.family_window(D, E, lo, hi)
            continue
        DE = D << E
        qmin, qmax = -(-(1 << (lo - 1)) // DE), ((1 << hi) - 1) // DE
        got = ccs.family_window(D, E, lo, hi)
        if qmin > qmax:
            assert got is None
        else:
            assert got == (qmin, qmax)
            assert (qmin * DE - 1).bit_length() >= lo - 1 and (qmax * DE - 1).bit_length() <= hi
    for args in [(0, 0, 90, 127), (30, 0, 1, 90), (30, 0, 91, 90), (30, 0, 90, 128), (30, -1, 90, 100)]:
        with pytest.raises(ccs.CcsError):
            ccs.family_window(*args)


def test_dedup():
    rng = random.Random(5)
    vals = [rng.choice([1, 2, 3, 1 << 100, (1 << 100) + 1, 7]) for _ in range(200)]
    seen, want = set(), []
    for v in vals:
        want.append(v not in seen)
        seen.add(v)
    assert list(ccs.dedup(vals)) == want
    assert list(ccs.dedup([])) == []


def test_family_structs_match_header(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "ccs.h"\nint main(void){'
                   'printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(ccs_family_params), sizeof(ccs_pool_entry),'
                   ' sizeof(ccs_shape_info), offsetof(ccs_shape_info, D), offsetof(ccs_pool_entry, n_primes),'
                   ' offsetof(ccs_family_params, filter_primes)); return 0; }\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(_lib.FamilyParams), C.sizeof(_lib.PoolEntry), C.sizeof(_lib.ShapeInfo),
            _lib.ShapeInfo.D.offset, _lib.PoolEntry.n_primes.offset, _lib.FamilyParams.filter_primes.offset]
    assert got == want


def test_oracle_family_predicate_brute_force(oracle):
    """orc_filter_family against q | p_j evaluated on integers, including E > 0, a prime
    dividing D (immune) and Q windows across 2^64."""
    rng = random.Random(11)
    for D, E, depth, q0 in [(2 * 3 * 5 * 11 * 13 * 19, 0, 12, 10 ** 20), (2 * 3 * 5 * 7, 3, 9, (1 << 64) - 700),
                            (6, 1, 5, 123456789), (30030 * 17, 2, 8, rng.randrange(1 << 80))]:
        primes = [q for q in primes_upto(120) if q > 2]
        got = list(oracle.filter_family(D, E, depth, primes, q0, 3000))
        want = []
        for k in range(3000):
            n = (q0 + k) * D * 2 ** E - 1
            if all(((2 ** j) * (n + 1) - 1) % q for q in primes if D % q for j in range(depth)):
                want.append(k)
        assert got == want, (D, E)
        assert len(want) > 0


def test_oracle_family_keeps_the_reference_chains(oracle):
    """Every v16 q-iter vector (v16...cu:998-1009) survives its family's sieve at its own
    depth over primes 3..1009 (the GPU sieve plus the v16 CPU ext sieve)."""
    import json
    vecs = json.load(open(os.path.join(GOLDEN, "reference_vectors.json")))["v16_qiter_family"]
    primes = [q for q in primes_upto(1009) if q > 2]
    for v in vecs:
        D, Q = spec_value(v["D_spec"]), int(v["Q"], 16)
        assert Q * D - 1 == int(v["p0"], 16)
        got = oracle.filter_family(D, 0, v["len"], primes, Q - 100, 200)
        assert 100 in set(got.tolist())
        # one step deeper than the chain: p_len is composite with a factor that may be small
        assert oracle.is_prime(Q * D - 1)
