"""Cholesky kernels of the oracle against a fresh factorisation, following
src/tests/symmmatrix_tests.hpp:41-235 (1e-10), including the two Givens-sign
2x2 known answers."""
import numpy as np
import pytest

import oracle
from oracle import C, lib, ptr


def chol(a):
    a = np.asfortranarray(a.copy())
    m = a.shape[0]
    ok = lib().orc_chol(ptr(a), m, m)
    return ok, np.triu(a)


def test_cholesky_kats():
    ok, u = chol(np.eye(10))
    assert ok and np.array_equal(u, np.eye(10))
    ok, u = chol(np.array([[1, 0.9], [0.9, 1.0]]))
    assert ok and u[0, 0] == pytest.approx(1) and u[0, 1] == pytest.approx(0.9)
    assert u[1, 1] == pytest.approx(0.4359, abs=1e-3)
    assert not chol(np.array([[1, 1], [1, 1.0]]))[0]
    assert not chol(np.array([[1, 1.1], [1.1, 1.0]]))[0]


def test_cholesky_matches_numpy():
    rng = np.random.default_rng(0)
    X = rng.normal(size=(50, 12))
    A = X.T @ X
    ok, u = chol(A)
    assert ok
    assert np.allclose(u, np.linalg.cholesky(A).T, atol=1e-12, rtol=1e-12)


def random_xx(rng, n=100, m=10):
    X = 10 * rng.normal(size=(n, m)) + 1
    return X.T @ X


def test_cholesky_update():
    rng = np.random.default_rng(1234)
    for _ in range(100):
        xx = random_xx(rng)
        m = xx.shape[0]
        ok, full = chol(xx)
        l = np.zeros((m, m), order="F")
        l[:m - 1, :m - 1] = chol(xx[:m - 1, :m - 1])[1]
        col = np.ascontiguousarray(xx[:, m - 1])
        assert lib().orc_chol_update(ptr(l), m, m - 1, ptr(col), 0.0)
        assert np.allclose(np.triu(l), full, atol=1e-10, rtol=0)


def test_cholesky_update_with_prior_term_and_failure():
    rng = np.random.default_rng(2)
    xx = random_xx(rng)
    m = xx.shape[0]
    A = xx + np.diag([0.0] * (m - 1) + [3.5])
    l = np.zeros((m, m), order="F")
    l[:m - 1, :m - 1] = chol(xx[:m - 1, :m - 1])[1]
    col = np.ascontiguousarray(xx[:, m - 1])
    assert lib().orc_chol_update(ptr(l), m, m - 1, ptr(col), 3.5)
    assert np.allclose(np.triu(l), chol(A)[1], atol=1e-10)
    # a duplicated column is not positive definite -> failure (src/symmmatrix.cpp:161)
    X = rng.normal(size=(20, 3))
    X = np.column_stack([X, X[:, 1]])
    xx = X.T @ X
    l = np.zeros((4, 4), order="F")
    l[:3, :3] = chol(xx[:3, :3])[1]
    col = np.ascontiguousarray(xx[:, 3])
    assert not lib().orc_chol_update(ptr(l), 4, 3, ptr(col), 0.0)


def test_cholesky_downdate():
    rng = np.random.default_rng(1234)
    for rep in range(100):
        xx = random_xx(rng)
        m = xx.shape[0]
        col_rem = int(rng.integers(0, m))
        ok, l = chol(xx)
        l = np.asfortranarray(l)
        lib().orc_chol_downdate(ptr(l), m, m, col_rem)
        keep = [i for i in range(m) if i != col_rem]
        ref = chol(xx[np.ix_(keep, keep)])[1]
        assert np.allclose(np.triu(l[:m - 1, :m - 1]), ref, atol=1e-10, rtol=0)


def test_cholesky_swapadj():
    rng = np.random.default_rng(1234)
    for rep in range(100):
        xx = random_xx(rng)
        m = xx.shape[0]
        col = int(round((m - 2) * rng.random()))
        ok, l = chol(xx)
        l = np.asfortranarray(l)
        v = rng.normal(size=m)
        v_ref = v.copy()
        lib().orc_chol_swapadj(ptr(l), m, m, col, ptr(v))
        perm = list(range(m))
        perm[col], perm[col + 1] = perm[col + 1], perm[col]
        ref = chol(xx[np.ix_(perm, perm)])[1]
        assert np.allclose(np.triu(l), ref, atol=1e-10, rtol=0)
        # v = U^-T b is rotated consistently: with b = U^T v before, after the
        # swap U'^T v' must equal the permuted b
        u0 = chol(xx)[1]
        b = u0.T @ v_ref
        assert np.allclose(np.triu(l).T @ v, b[perm], atol=1e-8)


@pytest.mark.parametrize("xx", [[[100.0, -50.0], [-50.0, 26.0]], [[100.0, 50.0], [50.0, 80.0]]])
def test_swapadj_givens_sign_kats(xx):
    """src/tests/symmmatrix_tests.hpp:187-233: both signs of r in drotg."""
    xx = np.array(xx)
    ok, l = chol(xx)
    l = np.asfortranarray(l)
    lib().orc_chol_swapadj(ptr(l), 2, 2, 0, None)
    assert l[0, 0] > 0 and l[1, 1] > 0
    ref = chol(xx[::-1, ::-1])[1]
    assert np.allclose(np.triu(l), ref, atol=1e-10)


def test_drotg_definition():
    for a, b in [(3.0, 4.0), (-3.0, 4.0), (4.0, -3.0), (0.0, 0.0), (0.0, 2.0)]:
        aa, bb, c, s = (C.c_double(a), C.c_double(b), C.c_double(), C.c_double())
        lib().orc_drotg(C.byref(aa), C.byref(bb), C.byref(c), C.byref(s))
        r = aa.value
        assert c.value * a + s.value * b == pytest.approx(r, abs=1e-14)
        assert -s.value * a + c.value * b == pytest.approx(0.0, abs=1e-14)
        if a or b:
            roe = a if abs(a) > abs(b) else b
            assert np.sign(r) == np.sign(roe)
