"""Arithmetic facts the kernels rely on, checked exhaustively on the CPU.

Stage B reduces x = j + s (x < 2^21) modulo a prime q < 4096 with one f32 multiply by a
reciprocal rounded up to the smallest float >= (1/q)(1 + 2^-23) and a truncation, with no
correction step.  numpy's float32 multiply rounds to nearest like the GPU's v_mul_f32.
"""
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
    for q in (int(v) for v in O.primes_between(3, 4096)):
        qt = (xf * rcp_up(q)).astype(np.uint32)
        assert np.array_equal(qt, x // q), q


def test_plain_reciprocal_is_not_exact():
    # the rounding-up matters: 1/q rounded to nearest under-estimates some quotients
    x = np.arange(0, 1 << 21, dtype=np.uint32)
    bad = sum(not np.array_equal((x.astype(np.float32) * np.float32(1.0 / q)).astype(np.uint32), x // q)
              for q in (37, 41, 43, 47, 53, 59, 61, 67, 71, 73))
    assert bad > 0
