"""Arithmetic facts the kernels rely on, checked exhaustively on the CPU.

Stage B reduces x = j + s (x < 2^21) modulo a prime q < 16384 with one f32 multiply by a
reciprocal rounded up to the smallest float >= (1/q)(1 + 2^-23) and a truncation, with no
correction step.  numpy's float32 multiply rounds to nearest like the GPU's v_mul_f32.

The deep stage reduces y = l0 + l1*c32 + l2*c64 + l3*c96 < 2^52 (a u128 folded by
c_k = 2^(32k) mod q) modulo a prime q < 2^18 in f64: r = 1/q correctly rounded (the host's
1.0 / q), qt = trunc(y * r), rem = fma(-qt, q, y), then rem += q if negative and rem -= q
if >= q.  The test checks that rem lands in [-q, 2q) for primes across the range, at the
ends of y's range and next to multiples of q, so that the two corrections always give
y mod q.
"""
from fractions import Fraction

import numpy as np
import pytest


def rcp_up(q):
    d = (1.0 / q) * (1.0 + 2.0 ** -23)
    f = np.float32(d)
    while float(f) < d:
        f = np.nextafter(f, np.float32(np.inf))
    return f


def test_rounded_up_reciprocal_gives_exact_quotients(oracle):
    O = oracle
    x = np.arange(0, 1 << 21, dtype=np.uint32)
    xf = x.astype(np.float32)
    for q in (int(v) for v in O.primes_between(3, 16384)):
        qt = (xf * rcp_up(q)).astype(np.uint32)
        assert np.array_equal(qt, x // q), q


def test_plain_reciprocal_is_not_exact():
    # the rounding-up matters: 1/q rounded to nearest under-estimates some quotients
    x = np.arange(0, 1 << 21, dtype=np.uint32)
    bad = sum(not np.array_equal((x.astype(np.float32) * np.float32(1.0 / q)).astype(np.uint32), x // q)
              for q in (37, 41, 43, 47, 53, 59, 61, 67, 71, 73))
    assert bad > 0


def test_deep_reduction_quotient_within_one(oracle):
    rng = np.random.default_rng(18)
    primes = [int(v) for v in oracle.primes_between(3, 1 << 18)]
    ymax = (1 << 32) - 1 + 3 * ((1 << 32) - 1) * ((1 << 18) - 1)      # largest limb fold
    assert ymax < 1 << 52
    for q in primes[::7] + primes[-50:] + primes[:50]:
        m = np.concatenate([rng.integers(0, ymax // q, 400, dtype=np.int64),
                            np.arange(ymax // q - 40, ymax // q + 1, dtype=np.int64),
                            np.arange(0, 40, dtype=np.int64)])
        y = np.concatenate([m * q, m * q + 1, m * q + q - 1, m * q + q // 2])
        y = y[(y >= 0) & (y <= ymax)]
        r = 1.0 / q                                                       # correctly rounded
        qt = np.trunc(y.astype(np.float64) * r).astype(np.int64)   # exact: y < 2^53
        rem = y - qt * q                                              # fma(-qt, q, y), exact
        assert rem.min() >= -q and rem.max() < 2 * q, q
        rem = np.where(rem < 0, rem + q, rem)
        rem = np.where(rem >= q, rem - q, rem)
        assert np.array_equal(rem, y % q)
