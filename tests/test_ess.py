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
