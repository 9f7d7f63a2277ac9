"""Model, Prior and delayed-rejection helpers of the oracle against the
reference's known answers (src/tests/model_tests.hpp, prior_tests.hpp,
sampler_tests.hpp:612-700) and against dense numpy algebra."""
import itertools
import math

import numpy as np
import pytest

import oracle
from fixtures import plinktest

LOG_HALF = math.log(0.5)

# Prior of src/tests/model_tests.hpp:48-56: nu_sigma=4, s2=1, nu_tau2=4,
# s2_tau2=1, alpha=2, e_qg=2, var_qg=4, inv_tau2_e=0, types {A, AH}
MODEL_KAT_CFG = dict(n=5, m_g=10, types="A,AH", e_qg=2, var_qg=4, nu_sigma2=4,
                     s2_sigma2=1, nu_tau2=[4, 4, 4, 4], s2_tau2=[1, 1, 1, 1],
                     mu_alpha=2.0, inv_tau2_e_const_val=0.0, n_rao=10, n_rao_burnin=1,
                     thin=1, n_sample_tau2_and_missing=10, do_n_iter=1, seed=1)


def sampler_on(cfg, body, y, recode=False):
    d = oracle.Data(cfg["n"], cfg["m_g"], body, y, recode=recode)
    return oracle.Sampler(cfg, d), d


def test_loglik_kats():
    body, y = plinktest()
    s, d = sampler_on(MODEL_KAT_CFG, body, y)
    m = s.current()
    tau2 = 4 * 1 / (4 - 2)
    inv = 1.0 / (2.0 * 2.0 * tau2)
    m.add_term(1, 0, [inv, inv])
    syy = 0.58 - 0.51670588235294123702
    ll = -0.5 * math.log(tau2) - 0.5 * math.log(4.0 * 10.0 + 5.0 / tau2) \
        - 0.5 * (4 + 5) * math.log(4 * 1 + syy)
    assert m.log_likelihood() == pytest.approx(ll, abs=1e-8)
    m.add_term(5, 1, [inv, inv])
    syy = 0.58 - 0.56537856592135671274
    ll = -0.5 * 3.0 * math.log(tau2) - 0.5 * math.log(810.62499999999954525265) \
        - 0.5 * (4 + 5) * math.log(4 * 1 + syy)
    assert m.log_likelihood() == pytest.approx(ll, abs=1e-8)


def dense_loglik(X, y, d, nus2_plus_yy, n_plus_nu):
    """log marginal likelihood of src/model.hpp:199-237 in dense algebra:
    0.5 sum_{i>=1} log d_i - log det chol(X'X + D) - (n+nu)/2 log(S)."""
    A = X.T @ X + np.diag(d)
    Lc = np.linalg.cholesky(A)
    v = np.linalg.solve(Lc, X.T @ y)
    S = nus2_plus_yy - v @ v
    return 0.5 * np.sum(np.log(d[1:])) - np.sum(np.log(np.diag(Lc))) \
        - 0.5 * n_plus_nu * math.log(S)


def test_loglik_matches_dense_algebra_random_models():
    rng = np.random.default_rng(11)
    n, p = 40, 25
    G = rng.integers(0, 3, size=(p, n))
    from fixtures import pack_genotypes
    body = pack_genotypes(G)
    y = rng.normal(size=n)
    cfg = dict(MODEL_KAT_CFG, n=n, m_g=p, types="A")
    s, d = sampler_on(cfg, body, y)
    pr = s.prior()
    for trial in range(20):
        m = s.new()
        m.copy_from(oracle.Model(s, oracle.lib().orc_sampler_current_model(s.h)))
        k = int(rng.integers(1, 8))
        snps = rng.choice(p, size=k, replace=False)
        dv = rng.uniform(0.05, 2.0, size=k)
        for j, dj in zip(snps, dv):
            m.add_term(int(j), 0, [dj, dj])
        X = np.column_stack([np.ones(n)] + [G[j].astype(float) for j in snps])
        dd = np.concatenate([[0.0], dv])
        ref = dense_loglik(X, y, dd, pr["nus2_plus_yy"], pr["n_plus_nu"])
        assert m.log_likelihood() == pytest.approx(ref, rel=1e-11, abs=1e-11)
        # removing from the middle (dchex path) and re-adding
        mid = k // 2
        m.remove_term(mid)
        keep = [j for i, j in enumerate(snps) if i != mid]
        kd = [x for i, x in enumerate(dv) if i != mid]
        X = np.column_stack([np.ones(n)] + [G[j].astype(float) for j in keep])
        ref = dense_loglik(X, y, np.concatenate([[0.0], kd]), pr["nus2_plus_yy"], pr["n_plus_nu"])
        assert m.log_likelihood() == pytest.approx(ref, rel=1e-11, abs=1e-11)
        full = m.log_likelihood()
        m.compute_log_likelihood()
        assert m.log_likelihood() == pytest.approx(full, rel=1e-12, abs=1e-12)


def test_incremental_equals_full():
    """src/tests/model_tests.hpp:246-312."""
    body, y = plinktest()
    s, d = sampler_on(MODEL_KAT_CFG, body, y)
    m = s.current()
    inv = 1.0 / 8.0
    l00 = m.log_likelihood()
    m.add_term(1, 0, [inv, inv])
    l11 = m.log_likelihood(); m.compute_log_likelihood(); l12 = m.log_likelihood()
    m.remove_term(0)
    assert m.log_likelihood() == pytest.approx(l00, abs=1e-6)
    assert l11 == pytest.approx(l12, abs=1e-6) and abs(l00 - l11) > 1e-6
    m.add_term(1, 0, [inv, inv]); m.add_term(7, 1, [inv, inv])
    l11 = m.log_likelihood(); m.compute_log_likelihood(); l12 = m.log_likelihood()
    m.remove_term(0); m.remove_term(0)
    assert m.log_likelihood() == pytest.approx(l00, abs=1e-6)
    assert l11 == pytest.approx(l12, abs=1e-6)


def fresh_model_logpost(s, sub, addtypes, inv, base=()):
    m = s.new()
    m.copy_from(oracle.Model(s, oracle.lib().orc_sampler_current_model(s.h)))
    Ns = [0] * 5
    nl = 0
    for (j, t) in base:
        m.add_term(j, t, [inv, inv])
        Ns[[0, 4][t]] += 1
        nl += 1
    lp = 0.0
    for j in sub:
        t = addtypes[j]
        ty = [0, 4][t]
        m.add_term(j, t, [inv, inv])
        lp += s.prior_add(Ns, nl, ty)
        Ns[ty] += 1
        nl += 1
    return lp + m.log_likelihood()


def test_exhaustive_modelset():
    """src/tests/model_tests.hpp:92-174."""
    body, y = plinktest()
    s, d = sampler_on(MODEL_KAT_CFG, body, y)
    s2, d2 = sampler_on(MODEL_KAT_CFG, body, y)
    inv = 1.0 / 8.0
    addtypes = [0, 1, 0, 0, 1]
    m = s.current()
    for j in range(5):
        m.add_term(j, addtypes[j], [inv, inv])
    probs = np.zeros(32)
    mx = oracle.C.c_double()
    oracle.lib().orc_exhaustive_modelset(m.h, 0, 5, oracle.ptr(probs), oracle.C.byref(mx))
    for i in range(32):
        sub = [j for j in range(5) if (i >> j) & 1]
        assert probs[i] == pytest.approx(fresh_model_logpost(s2, sub, addtypes, inv), abs=1e-10)
    assert mx.value == pytest.approx(probs.max())


def test_exhaustive_modelset_with_const_loci():
    """src/tests/model_tests.hpp:176-244."""
    body, y = plinktest()
    s, d = sampler_on(MODEL_KAT_CFG, body, y)
    s2, d2 = sampler_on(MODEL_KAT_CFG, body, y)
    inv = 1.0 / 8.0
    addtypes = [1, 0, 0, 1, 0, 0, 0, 0]
    m = s.current()
    for j in range(8):
        m.add_term(j, addtypes[j], [inv, inv])
    probs = np.zeros(64)
    mx = oracle.C.c_double()
    oracle.lib().orc_exhaustive_modelset(m.h, 2, 6, oracle.ptr(probs), oracle.C.byref(mx))
    for i in range(64):
        sub = [j + 2 for j in range(6) if (i >> j) & 1]
        # the reference's expected value adds the prior of the 6 free loci only
        m2 = s2.new()
        m2.copy_from(oracle.Model(s2, oracle.lib().orc_sampler_current_model(s2.h)))
        m2.add_term(0, 1, [inv, inv]); m2.add_term(1, 0, [inv, inv])
        Ns = [1, 0, 0, 0, 1]; nl = 2; lp = 0.0
        for j in sub:
            ty = [0, 4][addtypes[j]]
            m2.add_term(j, addtypes[j], [inv, inv])
            lp += s2.prior_add(Ns, nl, ty); Ns[ty] += 1; nl += 1
        assert probs[i] == pytest.approx(lp + m2.log_likelihood(), abs=1e-10)


def test_prior_changes_match_log_model():
    """src/tests/prior_tests.hpp:31-124 (nu_tau2=3, s2=1, e_qg=3, var_qg=9)."""
    body, y = plinktest()
    cfg = dict(MODEL_KAT_CFG, e_qg=3, var_qg=9, nu_sigma2=1, s2_sigma2=1,
               nu_tau2=[3, 3, 3, 3], mu_alpha=1.0)
    s, d = sampler_on(cfg, body, y)
    A, AH = 0, 4
    for Ns, new, n_old, t in [([0, 0, 0, 0, 0], [1, 0, 0, 0, 0], 0, A),
                              ([2, 0, 0, 0, 3], [3, 0, 0, 0, 3], 5, A),
                              ([0, 0, 0, 0, 0], [0, 0, 0, 0, 1], 0, AH),
                              ([3, 0, 0, 0, 2], [3, 0, 0, 0, 3], 5, AH)]:
        dl = s.prior_log_model(new) - s.prior_log_model(Ns)
        assert s.prior_add(Ns, n_old, t) == pytest.approx(dl, abs=1e-6)
        assert s.prior_rem(new, n_old + 1, t) == pytest.approx(-dl, abs=1e-6)
    assert s.prior_swi([3, 0, 0, 0, 3], AH, AH) == pytest.approx(0.0)
    dl = s.prior_log_model([2, 0, 0, 0, 3]) - s.prior_log_model([3, 0, 0, 0, 2])
    assert s.prior_swi([3, 0, 0, 0, 2], AH, A) == pytest.approx(dl, abs=1e-6)
    assert s.prior_swi([2, 0, 0, 0, 3], A, AH) == pytest.approx(-dl, abs=1e-6)


def test_beta_binomial_params():
    """src/prior.hpp:273-290 reproduces E and Var of the model size."""
    body, y = plinktest()
    for e_q, var_q in [(2, 4), (3, 9), (2, 2)]:
        s, d = sampler_on(dict(MODEL_KAT_CFG, e_qg=e_q, var_qg=var_q), body, y)
        pr = s.prior()
        a, b, m = pr["g_a"], pr["g_b"], 10
        mean = m * a / (a + b)
        var = m * a * b * (a + b + m) / ((a + b) ** 2 * (a + b + 1))
        assert mean == pytest.approx(e_q) and var == pytest.approx(var_q)


def test_exh_proposal_probabilities_kat():
    """src/tests/sampler_tests.hpp:612-700, all 8 closed forms."""
    q_add = [0.3, 0.4, 0.5, 0.6]; q_rem = [0.7, 0.6, 0.5, 0.4]
    b2n = np.array([0, 2, 1], dtype=np.uint8)
    dq_add = np.array([0.3, 0.5, 0.6]); dq_rem = np.array([0.7, 0.5, 0.4])
    lp = np.zeros(8)
    oracle.lib().orc_proposal_probs_exh(0, 3, oracle.u8(b2n), oracle.ptr(dq_add),
                                        oracle.ptr(dq_rem), 0.3 + 0.5 + 0.6, 0.6, 1, 4,
                                        None, oracle.ptr(lp))
    L = math.log; H = LOG_HALF; qa = q_add; qr = q_rem
    exp = [
        L(qa[0] / (qa[0] + qa[2] + qa[3])) + H + L(qa[2] / (qa[2] + qa[3])) + H + L(qa[3] / qa[3]) + H,
        L(qr[0] / (qr[0] + qr[1])) + H + L(qa[2] / (qa[2] + qa[3])) + H + L(qa[3] / qa[3]) + H,
        L(qa[0] / (qa[0] + qa[2])) + H + L(qa[2] / qa[2]) + H + L(qr[3] / (qr[1] + qr[3])),
        L(qr[3] / (qr[0] + qr[1] + qr[3])) + H + L(qa[2] / qa[2]) + H + L(qr[0] / (qr[1] + qr[0])),
        L(qa[0] / (qa[0] + qa[3])) + H + L(qr[2] / (qr[1] + qr[2])) + H + L(qa[3] / qa[3]) + H,
        L(qr[2] / (qr[0] + qr[1] + qr[2])) + H + L(qr[0] / (qr[1] + qr[0])) + H + L(qa[3] / qa[3]) + H,
        L(qa[0] / qa[0]) + H + L(qr[3] / (qr[3] + qr[1] + qr[2])) + L(qr[2] / (qr[1] + qr[2])),
        L(qr[3] / (qr[0] + qr[3] + qr[1] + qr[2])) + L(qr[2] / (qr[0] + qr[1] + qr[2])) + L(qr[0] / (qr[0] + qr[1])),
    ]
    assert np.allclose(lp, exp, atol=1e-10, rtol=0)


def test_exh_proposals_equal_brute_force_sequential_draws():
    """Every pattern's probability equals the product of the sequential
    without-replacement draws, enumerated by brute force for tiny m_g."""
    rng = np.random.default_rng(9)
    m_g, cl, ms = 6, 2, 3
    q_add = rng.uniform(0.1, 0.9, m_g); q_rem = rng.uniform(0.1, 0.9, m_g)
    # const loci {0,1}; moved variables 2,3,4 in "normalised" order; SNP 5 free
    moved = [2, 3, 4]
    z_add = q_add[moved].sum() + q_add[5]
    z_rem = q_rem[[0, 1]].sum()
    b2n = np.array([0, 1, 2], dtype=np.uint8)
    lp = np.zeros(8)
    oracle.lib().orc_proposal_probs_exh(0, ms, oracle.u8(b2n), oracle.ptr(q_add[moved].copy()),
                                        oracle.ptr(q_rem[moved].copy()), z_add, z_rem, cl, m_g,
                                        None, oracle.ptr(lp))
    for pattern in range(8):
        # bit j set: moved[j] is in the starting model (the move removes it)
        inm = set([0, 1] + [moved[j] for j in range(ms) if (pattern >> j) & 1])
        isadd = [not ((pattern >> j) & 1) for j in range(ms)]
        # forward draws in normalised order: additions in order, removals in reverse
        za = sum(q_add[i] for i in range(m_g) if i not in inm)
        zr = sum(q_rem[i] for i in inm)
        madd, mrem = m_g - len(inm), len(inm)
        lpp = 0.0
        rem_pos = [j for j in range(ms) if not isadd[j]][::-1]
        ri = 0
        for j in range(ms):
            if madd > 0 and mrem > 0:
                lpp += LOG_HALF
            if isadd[j]:
                i = moved[j]; lpp += math.log(q_add[i] / za); za -= q_add[i]; madd -= 1
            else:
                i = moved[rem_pos[ri]]; ri += 1
                lpp += math.log(q_rem[i] / zr); zr -= q_rem[i]; mrem -= 1
        assert lp[pattern] == pytest.approx(lpp, abs=1e-12)


def test_residual_term_in_the_cancelling_regime():
    """Two identical SNP columns and a phenotype that they and two others fit
    exactly, with a tiny prior precision: S = nu s2 + y'y - v'v cancels to
    noise (the reference's formula, src/model.hpp:226,498), and the model
    forms S directly instead; it matches the regularised least-squares
    residual |y - X mu|^2 + mu'D mu of the augmented system [X; D^1/2],
    solved by SVD in numpy, and the log-likelihood stays finite."""
    rng = np.random.default_rng(17)
    n, p = 50, 20
    G = rng.integers(0, 3, size=(p, n))
    G[4] = G[3]                                     # perfect LD
    from fixtures import pack_genotypes
    body = pack_genotypes(G)
    snps = (3, 4, 7, 11)
    X = np.column_stack([np.ones(n)] + [G[j].astype(float) for j in snps])
    y = X @ np.array([3.0, 7.0, 7.0, -11.0, 5.0])   # no noise: an exact fit
    cfg = dict(MODEL_KAT_CFG, n=n, m_g=p, types="A", nu_sigma2=0, s2_sigma2=1)
    s, d = sampler_on(cfg, body, y)
    m = s.current()
    m.compute_log_likelihood()
    dv = 1e-14
    for j in snps:
        m.add_term(j, 0, [dv, dv])
    st = m.get()
    D = np.diag([0.0] + [dv] * len(snps))
    A = np.vstack([X, np.sqrt(D)])
    b = np.concatenate([y, np.zeros(len(snps) + 1)])
    mu, *_ = np.linalg.lstsq(A, b, rcond=None)
    s_true = float(np.sum((A @ mu - b) ** 2))
    assert s_true > 0
    naive = s.prior()["nus2_plus_yy"] - st["v"] @ st["v"]
    assert abs(naive - s_true) > 0.1 * s_true       # the cancelled value is noise
    assert st["syx"] == pytest.approx(s_true, rel=1e-3)
    assert np.isfinite(m.log_likelihood())
    # removing an inner column (the dchex path) and the full refactorisation
    # keep it too
    m.remove_term(1)
    X2 = np.delete(X, 2, axis=1)
    A2 = np.vstack([X2, np.sqrt(np.diag([0.0] + [dv] * 3))])
    b2 = np.concatenate([y, np.zeros(4)])
    mu2, *_ = np.linalg.lstsq(A2, b2, rcond=None)
    s2_true = float(np.sum((A2 @ mu2 - b2) ** 2))
    assert m.get()["syx"] == pytest.approx(s2_true, rel=1e-3)
    m.compute_log_likelihood()
    assert m.get()["syx"] == pytest.approx(s2_true, rel=1e-3)
    assert np.isfinite(m.log_likelihood())
