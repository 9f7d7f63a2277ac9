"""The drop-in Model on the GPU (FP64 Cholesky add / dchex remove / swap /
full factorisation) against the reference's known answers and the oracle
(1e-10 relative)."""
import math

import numpy as np
import pytest

import oracle
from fixtures import plinktest, pack_genotypes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bmg():
    import bmagwa_amd
    bmagwa_amd.lib()
    return bmagwa_amd


def test_loglik_kats(bmg):
    """src/tests/model_tests.hpp:36-90 (types A, AH; tau2 = 2, alpha = 2)."""
    body, y = plinktest()
    ctx = bmg.Ctx(5, 10, body, y, types=("A", "AH"))
    m = bmg.Model(ctx, n_plus_nu=5 + 4, nus2_plus_yy=4 * 1 + 0.58, inv_tau2_e0=0.0)
    tau2 = 2.0
    inv = 1.0 / (2.0 * 2.0 * tau2)
    m.add_term(1, 0, [inv, inv])
    syy = 0.58 - 0.51670588235294123702
    ll = -0.5 * math.log(tau2) - 0.5 * math.log(4.0 * 10.0 + 5.0 / tau2) \
        - 0.5 * 9 * math.log(4 + syy)
    assert m.log_likelihood() == pytest.approx(ll, abs=1e-8)
    m.add_term(5, 1, [inv, inv])
    syy = 0.58 - 0.56537856592135671274
    ll = -1.5 * math.log(tau2) - 0.5 * math.log(810.62499999999954525265) \
        - 0.5 * 9 * math.log(4 + syy)
    assert m.log_likelihood() == pytest.approx(ll, abs=1e-8)


def oracle_model_pair(n, p, body, y, types, nu=4.0, s2=1.0):
    cfg = dict(n=n, m_g=p, types=",".join(types), e_qg=2, var_qg=4, nu_sigma2=nu, s2_sigma2=s2,
               nu_tau2=[4, 4, 4, 4], s2_tau2=[1, 1, 1, 1], mu_alpha=2.0, n_rao=10,
               n_rao_burnin=1, thin=1, n_sample_tau2_and_missing=10, do_n_iter=1, seed=1)
    d = oracle.Data(n, p, body, y)
    s = oracle.Sampler(cfg, d)
    return s, d


@pytest.mark.parametrize("types", [("A",), ("AH",), ("A", "AH")])
def test_random_add_remove_sequences_match_oracle(bmg, types):
    rng = np.random.default_rng(len(types) * 31 + 7)
    n, p = 257, 60
    G = rng.integers(0, 3, size=(p, n))
    body = pack_genotypes(G)
    y = rng.normal(size=n) + 0.3 * G[3] - 0.2 * G[17]
    s, d = oracle_model_pair(n, p, body, y, types)
    pr = s.prior()
    om = s.current()
    ctx = bmg.Ctx(n, p, body, y, types=types)
    gm = bmg.Model(ctx, n_plus_nu=pr["n_plus_nu"], nus2_plus_yy=pr["nus2_plus_yy"])
    assert gm.log_likelihood() == pytest.approx(om.log_likelihood(), rel=1e-12)
    inmodel = []
    nt = len(types)
    for step in range(120):
        if inmodel and (rng.random() < 0.45 or len(inmodel) > 25):
            mi = int(rng.integers(0, len(inmodel)))
            om.remove_term(mi); gm.remove_term(mi)
            inmodel.pop(mi)
        else:
            j = int(rng.choice([x for x in range(p) if x not in inmodel]))
            t = int(rng.integers(0, nt))
            dv = rng.uniform(0.05, 3.0, size=2)
            om.add_term(j, t, dv); gm.add_term(j, t, dv)
            inmodel.append(j)
        lo, lg = om.log_likelihood(), gm.log_likelihood()
        assert gm.size() == om.size() == len(inmodel)
        assert lg == pytest.approx(lo, rel=1e-10), step
        if step % 17 == 0:
            om.compute_log_likelihood(); gm.compute_log_likelihood()
            assert gm.log_likelihood() == pytest.approx(om.log_likelihood(), rel=1e-10)
            a, b = gm.get(), om.get()
            assert np.array_equal(a["xx"], b["xx"])       # integer X^T X: bit-exact
            assert np.allclose(a["l"], b["l"], rtol=1e-11, atol=1e-11)


def test_non_positive_definite_gives_minus_infinity(bmg):
    """A SNP identical to the constant column with d = 0 makes X'X + D
    singular: the log-likelihood is -inf (src/model.hpp:476-484), which the
    sampler turns into a rejection; the oracle agrees."""
    rng = np.random.default_rng(0)
    n, p = 50, 10
    G = rng.integers(0, 3, size=(p, n))
    G[2] = 1                                   # x = 1: duplicate of the constant
    body = pack_genotypes(G)
    y = rng.normal(size=n)
    ctx = bmg.Ctx(n, p, body, y)
    s, d = oracle_model_pair(n, p, body, y, ("A",))
    pr = s.prior()
    m = bmg.Model(ctx, n_plus_nu=pr["n_plus_nu"], nus2_plus_yy=pr["nus2_plus_yy"])
    om = s.current()
    m.add_term(1, 0, [0.5]); om.add_term(1, 0, [0.5])
    assert np.isfinite(m.log_likelihood())
    assert m.log_likelihood() == pytest.approx(om.log_likelihood(), rel=1e-12)
    m.add_term(2, 0, [0.0]); om.add_term(2, 0, [0.0])
    assert m.log_likelihood() == -np.inf and om.log_likelihood() == -np.inf


def test_swapadj_matches_fresh_factor(bmg):
    rng = np.random.default_rng(5)
    n, p = 120, 12
    G = rng.integers(0, 3, size=(p, n))
    body = pack_genotypes(G)
    y = rng.normal(size=n)
    ctx = bmg.Ctx(n, p, body, y)
    m = bmg.Model(ctx, n_plus_nu=n + 1, nus2_plus_yy=1 + y @ y, inv_tau2_e0=0.0)
    for j in range(6):
        m.add_term(j, 0, [0.5])
    before = m.get()
    col = 3
    m.swapadj(col)
    after = m.get()
    perm = list(range(7)); perm[col], perm[col + 1] = perm[col + 1], perm[col]
    A = before["xx"] + np.triu(before["xx"], 1).T + np.diag(before["d"])
    ref = np.linalg.cholesky(A[np.ix_(perm, perm)]).T
    assert np.allclose(np.triu(after["l"]), ref, atol=1e-10)
