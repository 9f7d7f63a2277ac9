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


def _sampler_dataset(n, p, seed):
    from bmagwa_amd import synth
    thresh, rho, _ = synth.snp_layout(p, (0.05, 0.5), (0.8, 0.9, 10, 30), seed)
    body = oracle.synth_genotypes(seed, n, p, thresh, rho)
    dec = lambda snps: np.stack([oracle.decode(body, n, int(j), 0) for j in snps])
    y, causal, _ = synth.phenotype(dec, n, p, 6, 0.6, seed)
    cfg = synth.sampler_cfg(n, p, 6, seed=29, chain=3)
    return body, y, cfg


OPS = [(0, 17, 0), (0, 120, 0), (0, 121, 0), (0, 333, 0), (0, 480, 0), (1, 2, 0), (0, 250, 0),
       (2, 0, 0), (3, 0, 0), (2, 0, 0), (3, 0, 0)]


def _oracle_ops(cfg, body, y, n, p, d):
    s = oracle.Sampler(cfg, oracle.Data(n, p, body, y))
    m = s.current()
    m.compute_log_likelihood()
    h = oracle.lib().orc_sampler_current_model(s.h)
    out = []
    for op, arg, _ in OPS:
        if op == 0:
            m.add_term(arg, 0, [d, d])
            out.append(("add", m.size(), m.log_likelihood()))
        elif op == 1:
            m.remove_term(arg)
            out.append(("remove", m.size(), m.log_likelihood()))
        elif op == 2:
            oracle.lib().orc_model_sample_beta_sigma2(s.h, h)
            k = m.cols()
            b = np.ctypeslib.as_array(oracle.lib().orc_model_beta(h), shape=(k,)).copy()
            out.append(("beta", m.get()["sigma2"], b))
        else:
            pv = np.zeros(3)
            oracle.lib().orc_model_compute_pve(h, oracle.ptr(pv))
            yh = np.ctypeslib.as_array(oracle.lib().orc_model_y_hat(h), shape=(n,)).copy()
            out.append(("pve", pv, s.rb(yh)))
    return out


def _compare_ops(got, want):
    for g, w in zip(got, want):
        assert g[0] == w[0]
        if g[0] in ("add", "remove"):
            assert g[1] == w[1]
            assert g[2] == pytest.approx(w[2], rel=1e-10)
        elif g[0] == "beta":
            assert g[1] == pytest.approx(w[1], rel=1e-10)
            assert np.allclose(g[2], w[2], rtol=1e-9, atol=1e-12)
        else:
            assert np.allclose(g[1], w[1], rtol=1e-10, atol=1e-14)
            assert np.allclose(g[2], w[2], rtol=1e-8, atol=1e-12)


def test_sampler_model_draws_pve_and_rb_match_oracle(bmg):
    """A model created from a Sampler's options (bmg_model_create_sampler)
    carries its prior and random stream: add / remove, sample_beta_sigma2,
    compute_pve and the RB scan of a caller-held model (bmg_model_rb)
    reproduce the oracle's Model and RaoBlackwellizer."""
    n, p = 400, 600
    body, y, cfg = _sampler_dataset(n, p, 41)
    s0 = oracle.Sampler(cfg, oracle.Data(n, p, body, y))
    d = s0.prior()["inv_tau2_alpha2_A"]
    want = _oracle_ops(cfg, body, y, n, p, d)
    ctx = bmg.Ctx(n, p, body, y)
    m = bmg.Model(ctx, cfg=cfg)
    got = []
    for op, arg, _ in OPS:
        if op == 0:
            m.add_term(arg, 0, [d, d])
            got.append(("add", m.size(), m.log_likelihood()))
        elif op == 1:
            m.remove_term(arg)
            got.append(("remove", m.size(), m.log_likelihood()))
        elif op == 2:
            m.sample_beta_sigma2()
            b, s2, _ = m.beta()
            got.append(("beta", s2, b))
        else:
            pv = m.compute_pve()
            _, _, yh = m.beta()
            got.append(("pve", pv, m.rb(yh)))
    _compare_ops(got, want)
    # commit / rollback: a proposal copied back and forth
    prop = bmg.Model(ctx, cfg=cfg)
    prop.rollback_from(m)
    prop.add_term(5, 0, [d, d])
    assert prop.size() == m.size() + 1
    prop.rollback_from(m)
    assert prop.log_likelihood() == m.log_likelihood()
    prop.add_term(5, 0, [d, d])
    m.commit_from(prop)
    assert m.size() == prop.size() and m.log_likelihood() == prop.log_likelihood()


def test_cpp_caller_matches_oracle(bmg, tmp_path):
    """tests/cpp/model_caller: a C++ HipModel (the binding a maintainer adds
    to the reference) driven through include/bmagwa_hip.h reproduces the
    oracle's log-likelihoods, draws, PVE and RB probabilities."""
    import ctypes
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "model_caller")
    assert os.path.exists(exe), "build the caller first (__graft_entry__.build())"
    n, p = 400, 600
    body, y, cfg = _sampler_dataset(n, p, 43)
    s0 = oracle.Sampler(cfg, oracle.Data(n, p, body, y))
    d = s0.prior()["inv_tau2_alpha2_A"]
    want = _oracle_ops(cfg, body, y, n, p, d)
    o = bmg.opts_from_cfg(cfg)
    blob = (np.uint32(n).tobytes() + np.uint64(p).tobytes() + body.tobytes() + y.astype(np.float64).tobytes()
            + ctypes.string_at(ctypes.byref(o), ctypes.sizeof(o)) + np.uint32(len(OPS)).tobytes())
    for op, arg, dd in OPS:
        blob += np.int32(op).tobytes() + np.uint32(arg).tobytes() + np.float64(d).tobytes()
    inp = tmp_path / "in.bin"
    inp.write_bytes(blob)
    print("model_caller n=%d p=%d ops=%d" % (n, p, len(OPS)), flush=True)
    res = subprocess.run([exe, str(inp)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    lines = res.stdout.split("\n")
    got, i = [], 0
    while i < len(lines) and lines[i]:
        f = lines[i].split()
        if f[0] in ("add", "remove"):
            got.append((f[0], int(f[2]), float(f[3])))
        elif f[0] == "beta":
            got.append(("beta", float(f[1]), np.array([float(v) for v in f[2:]])))
        else:
            pv = np.array([float(v) for v in f[1:4]])
            i += 1
            pr = np.array([float(v) for v in lines[i].split()[1:]])
            got.append(("pve", pv, pr))
        i += 1
    assert len(got) == len(OPS)
    _compare_ops(got, want)


def test_residual_term_in_the_cancelling_regime(bmg):
    """The device model forms S directly where nu s2 + y'y - v'v has
    cancelled (two identical columns, an exact fit, prior precision 1e-14):
    it matches the regularised least-squares residual of the augmented
    system and the oracle's guarded value, and the log-likelihood is
    finite."""
    rng = np.random.default_rng(17)
    n, p = 50, 20
    G = rng.integers(0, 3, size=(p, n))
    G[4] = G[3]
    body = pack_genotypes(G)
    snps = (3, 4, 7, 11)
    X = np.column_stack([np.ones(n)] + [G[j].astype(float) for j in snps])
    y = X @ np.array([3.0, 7.0, 7.0, -11.0, 5.0])
    from bmagwa_amd import synth
    cfg = synth.sampler_cfg(n, p, 2, seed=3, nu_sigma2=0, s2_sigma2=1, R2mode_sigma2=0.0)
    dv = 1e-14
    ctx = bmg.Ctx(n, p, body, y)
    m = bmg.Model(ctx, cfg=cfg)
    s = oracle.Sampler(cfg, oracle.Data(n, p, body, y))
    om = s.current()
    om.compute_log_likelihood()
    for j in snps:
        m.add_term(j, 0, [dv, dv])
        om.add_term(j, 0, [dv, dv])
    A = np.vstack([X, np.sqrt(np.diag([0.0] + [dv] * len(snps)))])
    b = np.concatenate([y, np.zeros(len(snps) + 1)])
    mu, *_ = np.linalg.lstsq(A, b, rcond=None)
    s_true = float(np.sum((A @ mu - b) ** 2))
    g = m.get()
    assert g["syx"] == pytest.approx(s_true, rel=1e-3)
    assert g["syx"] == pytest.approx(om.get()["syx"], rel=1e-6)
    assert np.isfinite(m.log_likelihood())
