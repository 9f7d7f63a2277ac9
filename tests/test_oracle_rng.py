"""The project random stream in the oracle: Philox4x32-10 known answers and
distribution moments (the reference's own RNG tests are moment tests,
src/tests/rand_tests.hpp:30-177)."""
import numpy as np
import pytest

import oracle


# Known-answer vectors of Philox4x32-10 from the Random123 distribution
# (kat_vectors: philox4x32 10 rounds).
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_kat(ctr, key, expect):
    assert tuple(int(x) for x in oracle.philox(ctr, key)) == expect


def test_u01_range_and_moments():
    u = oracle.rng_fill(1234, 0, 0, 0, 0, 0, 0, 200000)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 3e-3
    assert abs(u.var() - 1 / 12) < 2e-3


def test_streams_differ_by_key_and_counter():
    a = oracle.rng_fill(1, 0, 0, 0, 0, 0, 0, 16)
    b = oracle.rng_fill(1, 1, 0, 0, 0, 0, 0, 16)
    c = oracle.rng_fill(1, 0, 5, 0, 0, 0, 0, 16)
    assert not np.allclose(a, b) and not np.allclose(a, c)
    # the stream is deterministic
    assert np.array_equal(a, oracle.rng_fill(1, 0, 0, 0, 0, 0, 0, 16))


def test_normal_moments():
    z = oracle.rng_fill(99, 3, 0, 0, 1, 0, 0, 400000)
    assert abs(z.mean()) < 6e-3
    assert abs(z.var() - 1) < 1e-2
    assert abs(np.mean(z ** 3)) < 2e-2
    assert abs(np.mean(z ** 4) - 3) < 5e-2


@pytest.mark.parametrize("a", [0.4, 1.0, 1.5, 2.5, 502.5])
def test_gamma_moments(a):
    g = oracle.rng_fill(7, 1, 0, 0, 2, a, 0, 200000)
    assert g.min() >= 0
    assert abs(g.mean() - a) < 0.02 * a + 0.01
    assert abs(g.var() - a) < 0.05 * a + 0.01


def test_sinvchi2_moments():
    # scaled-Inv-chi2(nu, s2): mean nu s2/(nu-2), var 2 nu^2 s2^2/((nu-2)^2 (nu-4))
    nu, s2 = 10.0, 0.7
    x = oracle.rng_fill(11, 0, 0, 0, 3, nu, s2, 400000)
    assert abs(x.mean() - nu * s2 / (nu - 2)) < 5e-3
    var = 2 * nu ** 2 * s2 ** 2 / ((nu - 2) ** 2 * (nu - 4))
    assert abs(x.var() - var) / var < 0.05
