"""The batched PMV sampler on the GPU against the oracle under the same
random stream: identical moves, jump distances, move sizes and model sizes
per iteration, log-likelihoods within 1e-10 relative; several chains of one
batch each reproduce their own oracle run."""
import numpy as np
import pytest

import oracle
from bmagwa_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bmg():
    import bmagwa_amd
    bmagwa_amd.lib()
    return bmagwa_amd


def dataset(n, p, ld, n_causal, seed, types="A"):
    thresh, rho, _ = synth.snp_layout(p, (0.05, 0.5), ld, seed)
    body = oracle.synth_genotypes(seed, n, p, thresh, rho)
    dec = lambda snps: np.stack([oracle.decode(body, n, int(j), 0) for j in snps])
    y, causal, beta = synth.phenotype(dec, n, p, n_causal, 0.5, seed, dominance=(types == "AH"))
    return body, y, causal


def oracle_run(cfg, body, y, n_iter):
    d = oracle.Data(cfg["n"], cfg["m_g"], body, y)
    s = oracle.Sampler(cfg, d)
    s.init_flat()
    assert s.setup() == 0
    s.iterate(n_iter)
    return s


def compare(tr_g, tr_o, n_iter):
    assert len(tr_g["move"]) == n_iter
    first_bad = None
    for key in ("move", "jd", "ms", "size"):
        neq = np.flatnonzero(tr_g[key] != tr_o[key])
        if len(neq):
            first_bad = neq[0] if first_bad is None else min(first_bad, neq[0])
    assert first_bad is None, "diverged at iteration %d" % first_bad
    ll_o, ll_g = tr_o["loglik"], tr_g["loglik"]
    assert np.allclose(ll_g, ll_o, rtol=1e-10, atol=0), np.max(np.abs(ll_g - ll_o) / np.abs(ll_o))


@pytest.mark.parametrize("types,ld,use_ind", [("A", (0.9, 0.9, 10, 30), 1), ("A", None, 0),
                                              ("AH", (0.8, 0.9, 10, 20), 1)])
def test_trajectory_parity_single_chain(bmg, types, ld, use_ind):
    n, p, n_iter = 300, 400, 3000
    body, y, causal = dataset(n, p, ld, 5, 1234, types)
    cfg = synth.sampler_cfg(n, p, 5, types=types, seed=77, n_rao=100, n_rao_burnin=5,
                            thin=10, n_sample_tau2_and_missing=10, max_move_size=10,
                            use_individual_tau2=use_ind, do_n_iter=n_iter)
    s = oracle_run(cfg, body, y, n_iter)
    ctx = bmg.Ctx(n, p, body, y, types=(types,))
    b = bmg.Batch([ctx], [0], [cfg], dict(trace_every_iter=1))
    b.setup()
    b.run(n_iter)
    compare(b.trace(0), s.trace(), n_iter)
    assert np.allclose(b.p_rao(0), s.p_rao(), rtol=1e-8, atol=1e-10)
    assert np.array_equal(np.sort(b.loci(0)), np.sort(s.loci()))


@pytest.mark.parametrize("types,ld,mms,dr", [("A", (0.9, 0.9, 10, 30), 8, 8), ("A", (0.8, 0.95, 5, 20), 10, 4),
                                            ("AH", (0.8, 0.9, 10, 20), 6, 6)])
def test_trajectory_parity_delayed_rejection(bmg, types, ld, mms, dr):
    """Delayed rejection on the device (second stage of add/remove and
    state-change moves over all 2^ms sub-models): the trajectory is the
    oracle's, including the moves the second stage accepts, with the
    move-size limit equal to and above the delayed-rejection limit."""
    n, p, n_iter = 300, 400, 3000
    body, y, causal = dataset(n, p, ld, 5, 4321, types)
    cfg = synth.sampler_cfg(n, p, 5, types=types, seed=91, n_rao=100, n_rao_burnin=5,
                            thin=10, n_sample_tau2_and_missing=10, max_move_size=mms,
                            delay_rejection=dr, do_n_iter=n_iter)
    s = oracle_run(cfg, body, y, n_iter)
    ctx = bmg.Ctx(n, p, body, y, types=(types,))
    b = bmg.Batch([ctx], [0], [cfg], dict(trace_every_iter=1))
    b.setup()
    b.run(n_iter)
    tr = b.trace(0)
    compare(tr, s.trace(), n_iter)
    # the second stage did run: accepted moves that kept fewer SNPs than drawn
    part = (tr["jd"] > 0) & (tr["jd"] < tr["ms"]) & (tr["move"] != 1)
    assert part.sum() > 0
    assert np.allclose(b.p_rao(0), s.p_rao(), rtol=1e-8, atol=1e-10)
    assert np.array_equal(np.sort(b.loci(0)), np.sort(s.loci()))


@pytest.mark.parametrize("types,ld,dr", [("A", (0.9, 0.9, 10, 30), 8), ("AH", (0.8, 0.9, 10, 20), 6)])
def test_factor_in_hbm_matches_oracle(bmg, types, ld, dr):
    """Capacity 256 keeps the working Cholesky factor in a per-chain global
    array instead of LDS; forced on small models, the trajectory (delayed
    rejection included) is still the oracle's."""
    n, p, n_iter = 300, 400, 3000
    body, y, causal = dataset(n, p, ld, 5, 4321, types)
    cfg = synth.sampler_cfg(n, p, 5, types=types, seed=91, n_rao=100, n_rao_burnin=5,
                            thin=10, n_sample_tau2_and_missing=10, max_move_size=8,
                            delay_rejection=dr, do_n_iter=n_iter)
    s = oracle_run(cfg, body, y, n_iter)
    ctx = bmg.Ctx(n, p, body, y, types=(types,))
    b = bmg.Batch([ctx], [0], [cfg], dict(trace_every_iter=1, capacity=256))
    b.setup()
    b.run(n_iter)
    compare(b.trace(0), s.trace(), n_iter)
    assert np.allclose(b.p_rao(0), s.p_rao(), rtol=1e-8, atol=1e-10)


def test_models_beyond_the_lds_capacity(bmg):
    """220 causal SNPs at h2 = 0.9 (n = 2,000, p = 3,000): the models grow past
    the 160 columns an LDS-resident factor holds, the batch picks the
    256-column capacity from the prior, and the trajectory over 3,000
    iterations is the oracle's."""
    n, p, nc, n_iter = 2000, 3000, 220, 3000
    thresh, rho, _ = synth.snp_layout(p, (0.05, 0.5), None, 77)
    body = oracle.synth_genotypes(77, n, p, thresh, rho)
    dec = lambda snps: np.stack([oracle.decode(body, n, int(j), 0) for j in snps])
    y, causal, _ = synth.phenotype(dec, n, p, nc, 0.9, 77)
    cfg = synth.sampler_cfg(n, p, nc, seed=3, n_rao=100, n_rao_burnin=5, thin=10, do_n_iter=n_iter,
                            delay_rejection=4)
    s = oracle_run(cfg, body, y, n_iter)
    ctx = bmg.Ctx(n, p, body, y)
    b = bmg.Batch([ctx], [0], [cfg], dict(trace_every_iter=1, trace_capacity=256 * (n_iter // 10 + 8)))
    b.setup()
    b.run(n_iter)
    tr = b.trace(0)
    compare(tr, s.trace(), n_iter)
    assert tr["size"].max() > 160
    assert np.allclose(b.p_rao(0), s.p_rao(), rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("capacity", [64, 160])
def test_chains_leave_a_small_capacity_with_the_same_trajectory(bmg, capacity):
    """The same growing models (n = 2,000, p = 3,000, 220 causal, h2 = 0.9)
    in a launch of 64 or 160 columns: a chain whose model comes within one
    move of the capacity continues in the 256-column kernel from that
    iteration, and the trajectory over 2,000 iterations is still the
    oracle's (nothing is rejected for want of room)."""
    n, p, nc, n_iter = 2000, 3000, 220, 2000
    thresh, rho, _ = synth.snp_layout(p, (0.05, 0.5), None, 77)
    body = oracle.synth_genotypes(77, n, p, thresh, rho)
    dec = lambda snps: np.stack([oracle.decode(body, n, int(j), 0) for j in snps])
    y, causal, _ = synth.phenotype(dec, n, p, nc, 0.9, 77)
    cfg = synth.sampler_cfg(n, p, nc, seed=3, n_rao=100, n_rao_burnin=5, thin=10, do_n_iter=n_iter,
                            delay_rejection=4)
    s = oracle_run(cfg, body, y, n_iter)
    ctx = bmg.Ctx(n, p, body, y)
    b = bmg.Batch([ctx], [0], [cfg], dict(trace_every_iter=1, capacity=capacity))
    b.setup()
    b.run(n_iter)
    tr = b.trace(0)
    compare(tr, s.trace(), n_iter)
    assert tr["size"].max() + 1 > capacity - 21
    assert np.allclose(b.p_rao(0), s.p_rao(), rtol=1e-8, atol=1e-10)


def test_models_past_the_device_limit_stop_with_an_error(bmg):
    """AH coding (two columns a locus) with 220 causal SNPs at h2 = 0.9:
    the models grow past the 256 columns the device holds, and the run stops
    with BMG_E_CAPACITY naming the chain instead of rejecting the moves that
    do not fit (the reference stops the same way at its MAXMODELSIZE)."""
    n, p, nc = 2000, 3000, 220
    thresh, rho, _ = synth.snp_layout(p, (0.05, 0.5), None, 78)
    body = oracle.synth_genotypes(78, n, p, thresh, rho)
    dec = lambda snps: np.stack([oracle.decode(body, n, int(j), 0) for j in snps])
    y, causal, _ = synth.phenotype(dec, n, p, nc, 0.9, 78, dominance=True)
    cfg = synth.sampler_cfg(n, p, nc, types="AH", seed=5, n_rao=100, n_rao_burnin=5, thin=10,
                            do_n_iter=20000)
    ctx = bmg.Ctx(n, p, body, y, types=("AH",))
    b = bmg.Batch([ctx], [0], [cfg], dict(trace_every_iter=1))
    b.setup()
    with pytest.raises(bmg.BmgError, match="device limit"):
        for _ in range(20):
            b.run(1000)
    st = b.state(0)
    assert 2 * st["size"] + 1 + 2 * 20 + 1 > 256


def test_batch_of_chains_each_match_their_oracle_run(bmg):
    n, p, n_iter = 256, 300, 1500
    body, y, _ = dataset(n, p, (0.9, 0.9, 10, 30), 4, 99)
    cfgs = [synth.sampler_cfg(n, p, 4, seed=1000 + c, chain=c, n_rao=100, n_rao_burnin=5,
                              thin=10, max_move_size=8, do_n_iter=n_iter) for c in range(4)]
    ctx = bmg.Ctx(n, p, body, y)
    b = bmg.Batch([ctx], [0] * 4, cfgs, dict(trace_every_iter=1))
    b.setup()
    b.run(n_iter // 3)
    b.run(n_iter - n_iter // 3)       # runs resume exactly
    for c in range(4):
        s = oracle_run(cfgs[c], body, y, n_iter)
        compare(b.trace(c), s.trace(), n_iter)


def test_rb_scan_matches_oracle(bmg):
    n, p = 300, 350
    body, y, _ = dataset(n, p, (0.9, 0.9, 10, 30), 5, 5)
    cfg = synth.sampler_cfg(n, p, 5, seed=3, n_rao=50, thin=10, do_n_iter=500)
    s = oracle_run(cfg, body, y, 200)
    ctx = bmg.Ctx(n, p, body, y)
    b = bmg.Batch([ctx], [0], [cfg], dict(trace_every_iter=1))
    b.setup()
    b.run(200)
    yh = np.random.default_rng(1).normal(size=n) * 0.1
    pg = b.rb(0, yh)
    po = s.rb(yh)
    assert np.allclose(pg, po, rtol=1e-9, atol=1e-12)


def test_uneven_datasets_and_steady_state_graph(bmg):
    """Chains on three datasets with different sample sizes (the same SNP
    panel, different individuals) in one batch: every chain reproduces its
    own oracle run, and the steady-state hipGraph replay gives the same
    trajectories, samples and RB PIPs as the eager run, bit for bit."""
    p, n_iter, per_ds = 400, 1600, 2
    sets = [dataset(n, p, (0.9, 0.9, 10, 30), 4, 40 + n) for n in (180, 333, 520)]
    cfgs, chain_ctx = [], []
    for d, (body, y, _) in enumerate(sets):
        n = len(y)
        for k in range(per_ds):
            c = len(cfgs)
            cfgs.append(synth.sampler_cfg(n, p, 4, seed=700 + c, chain=c, n_rao=100, n_rao_burnin=3,
                                          thin=10, max_move_size=8, do_n_iter=n_iter))
            chain_ctx.append(d)
    out = {}
    for g in (0, 1):
        ctxs = [bmg.Ctx(len(y), p, body, y) for body, y, _ in sets]
        b = bmg.Batch(ctxs, chain_ctx, cfgs, dict(trace_every_iter=1, use_graph=g))
        b.setup()
        b.run(700)                    # burn-in ends inside this call
        b.run(n_iter - 700)           # whole segments: replayed when g = 1
        st = b.graph_stats()
        if g:
            assert st["builds"] == 1 and st["replays"] >= 8, st
        else:
            assert st["replays"] == 0
        out[g] = [(b.trace(c), b.thinned(c), b.p_rao(c)) for c in range(len(cfgs))]
        b.close()
    for c in range(len(cfgs)):
        (t0, h0, r0), (t1, h1, r1) = out[0][c], out[1][c]
        for key in t0:
            assert np.array_equal(t0[key], t1[key]), (c, key)
        for key in h0:
            assert np.array_equal(h0[key], h1[key]), (c, key)
        assert np.array_equal(r0, r1)
    for c in (0, 3, 5):
        body, y, _ = sets[chain_ctx[c]]
        s = oracle_run(cfgs[c], body, y, n_iter)
        compare(out[1][c][0], s.trace(), n_iter)
        assert np.allclose(out[1][c][2], s.p_rao(), rtol=1e-8, atol=1e-10)


def test_output_files_in_the_reference_format(bmg, tmp_path):
    """bmg_batch_write_outputs writes Sampler::sample's binary files
    (src/sampler.cpp:560-584,687-728,846-858; the README's value types): the
    thinned files hold the oracle run's samples (same stream: identical loci
    and sizes, log-likelihoods, sigma2, alpha and PVE within 1e-10), the
    per-iteration files its moves, jump distances and move sizes, the prior
    file the model-size prior of src/prior.hpp:169-183 at each sample, and
    _rao.dat the RB PIPs."""
    import math
    n, p, n_iter = 256, 300, 1200
    body, y, _ = dataset(n, p, (0.9, 0.9, 10, 30), 4, 99)
    cfg = synth.sampler_cfg(n, p, 4, seed=12, n_rao=100, n_rao_burnin=3, thin=10, max_move_size=8,
                            do_n_iter=n_iter)
    ctx = bmg.Ctx(n, p, body, y)
    b = bmg.Batch([ctx], [0], [cfg], dict(trace_every_iter=1))
    b.setup()
    b.run(n_iter)
    base = str(tmp_path / "run1")
    b.write_outputs(0, base)
    s = oracle_run(cfg, body, y, n_iter)
    th, tr = s.thinned(), s.trace()
    rd = lambda name, dt: np.fromfile(base + name, dtype=dt)
    assert np.array_equal(rd("_modelsize.dat", np.uint32), th["sizes"])
    assert np.array_equal(rd("_loci.dat", np.uint32), th["loci"])
    for name, key in (("_log_likelihood.dat", "loglik"), ("_sigma2.dat", "sigma2"), ("_alpha.dat", "alpha")):
        assert np.allclose(rd(name, np.float64), th[key], rtol=1e-10, atol=0), name
    assert np.allclose(rd("_pve.dat", np.float64).reshape(-1, 3), th["pve"], rtol=1e-9, atol=1e-14)
    assert np.array_equal(rd("_move_type.dat", np.uint8), tr["move"])
    assert np.array_equal(rd("_jumpdistance.dat", np.uint8), tr["jd"])
    assert np.array_equal(rd("_move_size.dat", np.uint8), tr["ms"])
    assert len(rd("_move_type.dat", np.uint8)) == n_iter and len(th["sizes"]) == n_iter // 10
    # model-size prior of one effect type: beta-binomial with the solved (a, b)
    pr = s.prior()
    lp = rd("_log_prior.dat", np.float64)
    ref = [math.lgamma(1.0 + k) + (math.lgamma(pr["g_a"] + k) + math.lgamma(pr["g_b"] + p - k)
                                    - math.lgamma(1.0 + k)) for k in th["sizes"].astype(float)]
    assert np.allclose(lp, ref, rtol=1e-12, atol=1e-9)
    assert np.allclose(rd("_rao.dat", np.float64), s.p_rao(), rtol=1e-8, atol=1e-10)
    assert not (tmp_path / "run1_types.dat").exists()


@pytest.mark.parametrize("dr", [0, 10])
def test_reference_cpu_configuration_trajectory(bmg, dr):
    """The reference's CPU-runnable workload shape: n = 1,000, p = 10,000
    independent SNPs (MAF ~ U(0.05, 0.5), Hardy-Weinberg), 10 causal SNPs,
    h2 = 0.3, additive coding, one chain, without delayed rejection and with
    the reference example's delay_rejection = 10 (max_move_size = 20): the
    device reproduces the oracle's trajectory (moves, jump distances, move
    sizes, model sizes identical; log-likelihoods within 1e-10) over 3,000
    iterations, and the RB PIPs."""
    n, p, n_iter = 1000, 10000, 3000
    thresh, rho, _ = synth.snp_layout(p, (0.05, 0.5), None, 2024)
    body = oracle.synth_genotypes(2024, n, p, thresh, rho)
    dec = lambda snps: np.stack([oracle.decode(body, n, int(j), 0) for j in snps])
    y, causal, _ = synth.phenotype(dec, n, p, 10, 0.3, 2024)
    cfg = synth.sampler_cfg(n, p, 10, seed=5, n_rao=500, n_rao_burnin=2, thin=10, do_n_iter=n_iter,
                            delay_rejection=dr)
    ctx = bmg.Ctx(n, p, body, y)
    b = bmg.Batch([ctx], [0], [cfg], dict(trace_every_iter=1))
    b.setup()
    b.run(n_iter)
    s = oracle_run(cfg, body, y, n_iter)
    compare(b.trace(0), s.trace(), n_iter)
    assert np.allclose(b.p_rao(0), s.p_rao(), rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("dr", [0, 3])
def test_model_size_limit_matches_exact_enumeration(bmg, dr):
    """max_model_size = 2 on the 7-SNP small-modelspace fixture
    (src/tests/testdata/small_modelspace_tau2.ini) with tau2 and alpha held
    at their set-up draw: no sample exceeds the limit, and the device run's
    RB PIPs match exact enumeration over the 29 models of at most 2 SNPs
    (the beta-binomial prior truncated at 2) to 0.01, with and without
    delayed rejection.  Sample frequencies are held to 0.03: on this fixture
    they sit up to 0.018 from the exact PIPs without any limit, in the
    oracle's runs as in the device's, while the RB PIPs agree to 0.001."""
    from fixtures import small_modelspace, SMALL_MODELSPACE_TAU2_INI
    body, y = small_modelspace()
    n_iter, cap = 400000, 2
    cfg = dict(SMALL_MODELSPACE_TAU2_INI)
    cfg.update(n_sample_tau2_and_missing=10 ** 9, n_rao=50, n_rao_burnin=100, thin=1,
               delay_rejection=dr, max_move_size=4, seed=77, do_n_iter=n_iter)
    d = oracle.Data(30, 7, body, y, recode=True)
    s = oracle.Sampler(cfg, d)
    s.init_flat()
    assert s.setup() == 0
    inv_a = s.prior()["inv_tau2_alpha2_A"]
    base = oracle.Model(s, oracle.lib().orc_sampler_new_model(s.h))
    empty = s.current()
    lps, subs = [], []
    for mask in range(1 << 7):
        sub = [j for j in range(7) if (mask >> j) & 1]
        base.copy_from(empty)
        for j in sub:
            base.add_term(j, 0, [inv_a])
        lps.append(base.log_likelihood() + s.prior_log_model([len(sub), 0, 0, 0, 0]))
        subs.append(sub)
    lps = np.array(lps)
    size = np.array([len(sub) for sub in subs])

    def pips(keep):
        w = np.where(keep, np.exp(lps - lps[keep].max()), 0.0)
        w /= w.sum()
        out = np.zeros(7)
        for wi, sub in zip(w, subs):
            out[sub] += wi
        return out

    exact = pips(size <= cap)
    # the limit matters here: about a third of the untruncated posterior
    # lies on larger models
    assert np.max(np.abs(pips(size >= 0) - exact)) > 0.1

    ctx = bmg.Ctx(30, 7, body, y, recode=True)
    b = bmg.Batch([ctx], [0], [cfg], dict(max_model_size=cap, trace_capacity=cap * (n_iter + 16)))
    b.setup()
    b.run(n_iter)
    th = b.thinned(0)
    sizes, loci = th["sizes"], th["loci"]
    assert sizes.max() <= cap
    burn = 20000
    off = int(sizes[:burn].sum())
    counts = np.bincount(loci[off:].astype(np.int64), minlength=7)
    freq = counts / (len(sizes) - burn)
    assert np.max(np.abs(freq - exact)) < 0.03, (freq, exact)
    p_rao = b.p_rao(0)
    assert np.max(np.abs(p_rao - exact)) < 0.01, (p_rao, exact)


@pytest.mark.parametrize("types,dr", [("A", 0), ("A", 10), ("AH", 4)])
def test_covariates_trajectory_matches_oracle(bmg, types, dr, tmp_path):
    """Two covariates beside the constant, the shape of the reference's own
    example (testdata/testdata.ini: m_e = 2, delay_rejection = 10): the
    chain's trajectory equals the oracle's, so do its RB PIPs, and the
    thinned PVE carries the covariates' share (src/model.hpp:345-392)."""
    n, p, n_iter = 600, 1500, 2000
    body, y0, _ = dataset(n, p, (0.9, 0.9, 10, 30), 6, 61, types)
    rng = np.random.default_rng(9)
    E = np.column_stack([rng.normal(size=n), rng.integers(0, 2, size=n).astype(float)])
    y = y0 + E @ np.array([0.8, -0.5])
    cfg = synth.sampler_cfg(n, p, 6, types=types, seed=4, n_rao=100, n_rao_burnin=3, thin=10,
                            do_n_iter=n_iter, delay_rejection=dr, m_e=2, inv_tau2_e_val=0.001)
    d = oracle.Data(n, p, body, y, e_cov=E)
    s = oracle.Sampler(cfg, d)
    s.init_flat()
    assert s.setup() == 0
    s.iterate(n_iter)
    ctx = bmg.Ctx(n, p, body, y, types=(types,), e_cov=E)
    b = bmg.Batch([ctx], [0], [cfg], dict(trace_every_iter=1))
    b.setup()
    b.run(n_iter)
    compare(b.trace(0), s.trace(), n_iter)
    assert np.allclose(b.p_rao(0), s.p_rao(), rtol=1e-8, atol=1e-10)
    base = str(tmp_path / "cov")
    b.write_outputs(0, base)
    pve = np.fromfile(base + "_pve.dat", dtype=np.float64).reshape(-1, 3)
    po = s.thinned()["pve"]
    assert pve.shape == po.shape
    assert np.allclose(pve, po, rtol=1e-8, atol=1e-12)
    assert (pve[:, 2] > 0).all()


def test_covariates_need_matching_options(bmg):
    """m_e of the options must be the dataset's number of covariates."""
    n, p = 100, 200
    body, y, _ = dataset(n, p, None, 2, 3)
    ctx = bmg.Ctx(n, p, body, y, e_cov=np.random.default_rng(1).normal(size=(n, 2)))
    cfg = synth.sampler_cfg(n, p, 2, seed=1, m_e=1)
    with pytest.raises(bmg.BmgError, match="m_e"):
        bmg.Batch([ctx], [0], [cfg])


def test_save_beta_matches_oracle(bmg, tmp_path):
    """save_beta = 1 (src/sampler.cpp:584,710-713): the beta of every
    thinned sample, in _beta.dat and through bmg_batch_thinned_beta, equals
    the oracle's (the constant and covariates first, then the loci)."""
    n, p, n_iter = 400, 900, 1500
    body, y0, _ = dataset(n, p, (0.9, 0.9, 10, 30), 5, 71, "AH")
    E = np.random.default_rng(3).normal(size=(n, 1))
    y = y0 + 0.4 * E[:, 0]
    cfg = synth.sampler_cfg(n, p, 5, types="AH", seed=8, n_rao=100, n_rao_burnin=3, thin=10,
                            do_n_iter=n_iter, save_beta=1, m_e=1, inv_tau2_e_val=0.01)
    s = oracle.Sampler(cfg, oracle.Data(n, p, body, y, e_cov=E))
    s.init_flat()
    assert s.setup() == 0
    s.iterate(n_iter)
    ctx = bmg.Ctx(n, p, body, y, types=("AH",), e_cov=E)
    b = bmg.Batch([ctx], [0], [cfg], dict(trace_every_iter=1))
    b.setup()
    b.run(n_iter)
    compare(b.trace(0), s.trace(), n_iter)
    want = s.thinned()["beta"]
    got = b.thinned_beta(0)
    assert got.shape == want.shape and len(got) > 2 * (n_iter // 10)
    assert np.allclose(got, want, rtol=1e-9, atol=1e-12)
    base = str(tmp_path / "sb")
    b.write_outputs(0, base)
    assert np.array_equal(np.fromfile(base + "_beta.dat", dtype=np.float64), got)
