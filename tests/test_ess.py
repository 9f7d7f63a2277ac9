"""The benchmark's ESS estimator against the oracle's scalar implementation
and against the closed form for AR(1) and independent series."""
import numpy as np

import oracle
from bmagwa_amd import ess


def test_ess_matches_oracle_implementation():
    rng = np.random.default_rng(3)
    X = np.zeros((5, 3000))
    for k, rho in enumerate([0.0, 0.3, 0.6, 0.9, 0.97]):
        z = 0.0
        for t in range(3000):
            z = rho * z + rng.normal()
            X[k, t] = z
    got = ess.ess_geyer(X)
    ref = np.array([oracle.ess_geyer(x) for x in X])
    assert np.allclose(got, ref, rtol=1e-9)
    for k, rho in enumerate([0.0, 0.3, 0.6, 0.9]):
        assert got[k] / 3000 == __import__("pytest").approx((1 - rho) / (1 + rho), rel=0.3)


def test_constant_and_indicator_series():
    M = np.zeros((3, 100), dtype=bool)
    M[1] = True
    M[2, ::2] = True
    e = ess.ess_geyer(M)
    assert e[0] == 100 and e[1] == 100
    sizes = np.array([2, 1, 0, 2])
    loci = np.array([5, 7, 7, 5, 9])
    Mi = ess.indicator_matrix(sizes, loci, [5, 7, 9])
    assert Mi.tolist() == [[True, False, False, True], [True, True, False, False],
                           [False, False, False, True]]


def _ar1(rng, rho, T, x0=0.0):
    z, out = x0, np.zeros(T)
    for t in range(T):
        z = rho * z + rng.normal()
        out[t] = z
    return out


def test_multichain_ess_closed_forms():
    """Independent chains of an AR(1) series: ESS close to C T (1-rho)/(1+rho);
    one chain: close to the single-series estimator; chains stuck in
    different states: far below the sample count; nothing moves: 0."""
    import pytest
    rng = np.random.default_rng(7)
    C, T = 8, 4000
    for rho in (0.0, 0.5, 0.9):
        X = np.stack([_ar1(rng, rho, T)[None, :] for _ in range(C)])
        e = ess.ess_multichain(X)[0]
        assert e / (C * T) == pytest.approx((1 - rho) / (1 + rho), rel=0.25)
    x = _ar1(rng, 0.6, T)
    assert ess.ess_multichain(x[None, None, :])[0] == pytest.approx(ess.ess_geyer(x[None, :])[0], rel=0.05)
    # half the chains always include the SNP, half never: the pooled variance
    # is all between-chain, the chains carry almost no information
    stuck = np.zeros((C, 1, T))
    stuck[: C // 2] = 1.0
    stuck[:, :, :5] = 0.5           # a few early values so the within-chain variance is not 0
    assert ess.ess_multichain(stuck)[0] < 0.01 * C * T
    assert ess.ess_multichain(np.zeros((C, 1, T)))[0] == 0.0
    assert ess.ess_multichain(np.ones((1, 1, T)))[0] == 0.0
    # mixing indicator chains: ESS of the same order as C T
    M = (rng.random((C, 1, T)) < 0.3).astype(float)
    assert ess.ess_multichain(M)[0] == pytest.approx(C * T, rel=0.15)
