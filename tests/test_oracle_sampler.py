"""Sampler-level checks of the oracle: the weighted sampler, the RB scan
closed form (src/tests/raoblackwellizer_tests.hpp:31-93), and posterior
inclusion probabilities against exact enumeration of all 2^7 models
(pattern of src/tests/sampler_tests.hpp:230-270 with tau2 held fixed)."""
import itertools
import math

import numpy as np
import pytest

import oracle
from fixtures import SMALL_MODELSPACE_TAU2_INI, small_modelspace


def test_sumtree_totals_after_updates():
    """Totals after zero/un-zero sequences, like the reference's
    DiscreteDistributionTests.Updates (values 0.1..0.5, 1e-15)."""
    w = np.array([0.1, 0.2, 0.3, 0.4, 0.5])
    t = oracle.SumTree(w)
    assert t.total() == pytest.approx(1.5, abs=1e-15)
    t.zero(0); assert t.zeroed(0) and t.total() == pytest.approx(1.4, abs=1e-15)
    t.zero(4); assert t.total() == pytest.approx(0.9, abs=1e-15)
    t.zero(2); assert t.total() == pytest.approx(0.6, abs=1e-15)
    t.unzero(0); assert not t.zeroed(0) and t.total() == pytest.approx(0.7, abs=1e-15)
    tw = t.total()
    t.zero(3); t.unzero(2); t.zero(2); t.unzero(3)
    assert t.total() == tw            # recomputed, no drift
    for i in range(5):
        t.zero(i)
    assert t.total() == 0.0
    rng = np.random.default_rng(1)
    for i in range(5):
        t.unzero(i)
    for _ in range(20000):
        i = int(rng.integers(0, 5))
        t.zero(i); t.zero((i + 2) % 5); t.unzero((i + 2) % 5); t.unzero(i)
        assert t.total() >= 0
    assert t.total() == pytest.approx(1.5, abs=1e-15)


def test_sumtree_pick_is_inverse_cdf_bruteforce():
    rng = np.random.default_rng(2)
    for n in [1, 5, 64, 65, 300, 5000]:
        w = rng.uniform(0.01, 1.0, n)
        t = oracle.SumTree(w)
        zeroed = rng.random(n) < 0.3
        for i in np.flatnonzero(zeroed):
            t.zero(int(i))
        eff = np.where(zeroed, 0.0, w)
        if eff.sum() == 0:
            continue
        cdf = np.cumsum(eff) / eff.sum()
        for u in rng.random(200):
            i = t.pick(u)
            assert not zeroed[i]
            lo = cdf[i - 1] if i > 0 else 0.0
            # exact inverse-CDF up to rounding of the partial sums
            assert lo - 1e-12 <= u <= cdf[i] + 1e-12


def test_sumtree_sampling_frequencies():
    w = np.array([0.1, 0.2, 0.3, 0.4, 0.5, 0.0, 2.0])
    t = oracle.SumTree(w)
    t.zero(3)
    u = oracle.rng_fill(5, 0, 0, 0, 0, 0, 0, 200000)
    picks = np.array([t.pick(x) for x in u])
    eff = w.copy(); eff[3] = 0
    freq = np.bincount(picks, minlength=7) / len(picks)
    assert np.allclose(freq, eff / eff.sum(), atol=3e-3)


def small_sampler(cfg_over=None):
    body, y = small_modelspace()
    cfg = dict(SMALL_MODELSPACE_TAU2_INI)
    cfg.update(cfg_over or {})
    d = oracle.Data(30, 7, body, y, recode=True)
    s = oracle.Sampler(cfg, d)
    return s, d


def test_rb_empty_model_closed_form():
    """src/tests/raoblackwellizer_tests.hpp:31-93."""
    s, d = small_sampler()
    m = s.current()
    m.compute_log_likelihood()
    oracle.lib().orc_model_sample_beta_sigma2(s.h, m.h)
    pves = np.zeros(3)
    oracle.lib().orc_model_compute_pve(m.h, oracle.ptr(pves))
    m.compute_log_likelihood()
    p_r = s.rb(np.zeros(30))
    sigma2 = m.get()["sigma2"]
    inv = s.prior()["inv_tau2_alpha2_A"]
    dm = oracle.DataModel(d, ["A"])
    yv = d.y
    n = 30
    for j in range(7):
        x = dm.get(0, j)
        xx, yx, xs, rs = x @ x, yv @ x, x.sum(), yv.sum()
        det = n * (xx + inv) - xs * xs
        rxxr = (rs * rs * (xx + inv) - 2.0 * yx * rs * xs + yx * yx * n) / det
        pip = math.exp(0.5 * (math.log(n) - math.log(det)) + 0.5 * math.log(inv)
                       - 0.5 * (rs * rs / n - rxxr) / sigma2
                       + s.prior_add([0] * 5, 0, 0))
        assert p_r[j] == pytest.approx(pip / (1 + pip), abs=1e-10)


def exact_pips(s, d, inv):
    """Posterior inclusion over all 2^7 models with tau2, alpha fixed."""
    base = oracle.Model(s, oracle.lib().orc_sampler_new_model(s.h))
    empty = s.current()
    lps = []
    subs = []
    for mask in range(1 << 7):
        base.copy_from(empty)
        sub = [j for j in range(7) if (mask >> j) & 1]
        for j in sub:
            base.add_term(j, 0, [inv])
        Ns = [len(sub), 0, 0, 0, 0]
        lps.append(base.log_likelihood() + s.prior_log_model(Ns))
        subs.append(sub)
    lps = np.array(lps)
    w = np.exp(lps - lps.max()); w /= w.sum()
    pip = np.zeros(7)
    for wi, sub in zip(w, subs):
        pip[sub] += wi
    return pip


@pytest.mark.parametrize("dr", [0, 3])
def test_posterior_pips_match_exact_enumeration(dr):
    """MCMC-frequency and RB PIPs of a plain PMV run vs exact enumeration,
    with tau2/alpha sampled once at the start and then held fixed
    (n_sample_tau2_and_missing larger than the run, src/tests/testdata/
    onesnp_notau2.ini:27)."""
    n_iter = 200000
    s, d = small_sampler(dict(n_sample_tau2_and_missing=10 ** 9, n_rao=50, n_rao_burnin=100,
                              thin=1, delay_rejection=dr, max_move_size=4, seed=77))
    s.init_flat()
    s.setup()
    inv = s.current().get()["d"]
    s.iterate(n_iter)
    # after setup tau2/alpha are fixed: the shared value of the SNP terms
    inv_a = s.prior()["inv_tau2_alpha2_A"]
    exact = exact_pips(s, d, inv_a)
    th = s.thinned()
    sizes, loci = th["sizes"], th["loci"]
    burn = 20000
    counts = np.zeros(7)
    off = 0
    kept = 0
    for t, k in enumerate(sizes):
        if t >= burn:
            counts[loci[off:off + k]] += 1
            kept += 1
        off += k
    freq = counts / kept
    assert np.max(np.abs(freq - exact)) < 0.02, (freq, exact)
    p_rao = s.p_rao()
    assert np.max(np.abs(p_rao - exact)) < 0.02, (p_rao, exact)


def test_ess_geyer_sanity():
    rng = np.random.default_rng(0)
    x = rng.normal(size=20000)
    assert 0.8 * 20000 < oracle.ess_geyer(x) < 1.2 * 20000
    # AR(1) with rho = 0.9: ESS ~ n (1 - rho) / (1 + rho)
    z = np.zeros(20000)
    for i in range(1, 20000):
        z[i] = 0.9 * z[i - 1] + rng.normal()
    assert oracle.ess_geyer(z) == pytest.approx(20000 * 0.1 / 1.9, rel=0.25)


def test_geometric_cdf_and_naive_sampling():
    v = np.zeros(5)
    oracle.lib().orc_geometric_dist_cdf(5, 0.2, oracle.ptr(v))
    assert np.allclose(v, 1 - 0.8 ** np.arange(1, 6))
    assert oracle.lib().orc_sample_discrete_naive(oracle.ptr(v), 5, 0.0) == 0
    assert oracle.lib().orc_sample_discrete_naive(oracle.ptr(v), 5, 0.999) == 4
