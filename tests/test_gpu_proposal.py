"""A proposal table set from outside (bmg_batch_set_p_proposal, the entry
the learned per-SNP proposal uses) gives the device chain the oracle's
trajectory when the oracle is given the same table at the same iteration."""
import numpy as np
import pytest

import oracle
from bmagwa_amd import synth
from test_gpu_sampler import compare, dataset

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bmg():
    import bmagwa_amd
    bmagwa_amd.lib()
    return bmagwa_amd


@pytest.mark.parametrize("dr", [0, 6])
def test_set_proposal_table_matches_oracle(bmg, dr):
    n, p, first, rest = 300, 400, 400, 3000
    body, y, causal = dataset(n, p, (0.9, 0.9, 10, 30), 5, 4321)
    cfg = synth.sampler_cfg(n, p, 5, seed=91, n_rao=100, n_rao_burnin=2, thin=10,
                            n_sample_tau2_and_missing=10, max_move_size=10, delay_rejection=dr,
                            do_n_iter=first + rest)
    rng = np.random.default_rng(5)
    table = rng.uniform(0, 1, size=p) ** 4
    table[causal] = 0.9
    d = oracle.Data(n, p, body, y)
    s = oracle.Sampler(cfg, d)
    s.init_flat()
    assert s.setup() == 0
    s.iterate(first)
    assert s.set_p_proposal(table) == 0
    s.iterate(rest)

    ctx = bmg.Ctx(n, p, body, y)
    b = bmg.Batch([ctx], [0], [cfg], dict(trace_every_iter=1))
    b.setup()
    b.run(first)
    b.set_p_proposal(0, table)
    assert np.array_equal(b.p_proposal(0), table)
    b.run(rest)
    compare(b.trace(0), s.trace(), first + rest)
    assert np.array_equal(b.p_proposal(0), table)      # adaptation is over
    assert np.allclose(b.p_rao(0), s.p_rao(), rtol=1e-8, atol=1e-10)


def test_set_proposal_table_rejects_bad_values(bmg):
    n, p = 100, 50
    body, y, _ = dataset(n, p, None, 2, 7)
    cfg = synth.sampler_cfg(n, p, 2, seed=3, do_n_iter=100)
    ctx = bmg.Ctx(n, p, body, y)
    b = bmg.Batch([ctx], [0], [cfg], {})
    with pytest.raises(RuntimeError):
        b.set_p_proposal(0, np.full(p, 0.5))       # before setup
    b.setup()
    bad = np.full(p, 0.5)
    bad[7] = 1.2
    with pytest.raises(RuntimeError):
        b.set_p_proposal(0, bad)
    with pytest.raises(ValueError):
        b.set_p_proposal(0, np.full(p + 1, 0.5))


def test_learned_snp_proposal_keeps_the_trajectory(bmg):
    """proposal = 2: the per-SNP network trains on the burn-in RB scans
    without touching the chains, and at the end of burn-in its table
    becomes each chain's p_proposal.  The oracle, given that table at the
    same iteration, follows the device chain exactly."""
    n, p = 300, 400
    n_rao, burn, rest = 100, 3, 2500
    body, y, causal = dataset(n, p, (0.9, 0.9, 10, 30), 5, 4321)
    cfgs = [synth.sampler_cfg(n, p, 5, seed=91 + c, chain=c, n_rao=n_rao, n_rao_burnin=burn, thin=10,
                              n_sample_tau2_and_missing=10, max_move_size=10,
                              do_n_iter=n_rao * burn + rest) for c in range(2)]
    ctx = bmg.Ctx(n, p, body, y)
    b = bmg.Batch([ctx], [0, 0], cfgs, dict(trace_every_iter=1, proposal=2, learn_window=10,
                                            learn_steps=30, learn_batch=512, learn_every=n_rao))
    b.setup()
    b.run(n_rao * burn + rest)
    st = b.learn_state(0)
    assert st["frozen"] and st["rounds"] >= 1
    for c in range(2):
        table = b.p_proposal(c)
        assert np.all((table > 0) & (table < 1))
        # the network ranks the causal SNPs above the typical SNP
        assert np.median(table[causal]) > np.median(table)
        d = oracle.Data(n, p, body, y)
        s = oracle.Sampler(cfgs[c], d)
        s.init_flat()
        assert s.setup() == 0
        s.iterate(n_rao * burn)
        assert not np.array_equal(s.p_proposal(), table)    # the plain running mean differs
        assert s.set_p_proposal(table) == 0
        s.iterate(rest)
        compare(b.trace(c), s.trace(), n_rao * burn + rest)


def test_both_learned_proposals_match_exact_enumeration(bmg):
    """The whole learned pipeline on the device (proposal = 3: both networks
    trained during burn-in, the swap table and the per-SNP table frozen at
    its end) on the 7-SNP small-modelspace fixture, tau2 and alpha held at
    their set-up draw: after burn-in the RB PIPs of four chains match exact
    enumeration over all 128 models to 0.01 and the sample frequencies to
    0.03 (the tolerances the plain sampler meets on this fixture)."""
    from fixtures import small_modelspace, SMALL_MODELSPACE_TAU2_INI
    body, y = small_modelspace()
    n_iter, nc = 250000, 4
    cfgs = []
    for c in range(nc):
        cfg = dict(SMALL_MODELSPACE_TAU2_INI)
        cfg.update(n_sample_tau2_and_missing=10 ** 9, n_rao=50, n_rao_burnin=40, thin=1,
                   delay_rejection=0, max_move_size=4, seed=77 + c, chain=c, do_n_iter=n_iter)
        cfgs.append(cfg)
    ctx = bmg.Ctx(30, 7, body, y, recode=True)
    b = bmg.Batch([ctx], [0] * nc, cfgs,
                  dict(proposal=3, learn_window=3, learn_steps=20, learn_batch=256, learn_every=50,
                       trace_every_iter=1, trace_capacity=7 * (n_iter + 16)))
    b.setup()
    b.run(n_iter)
    st = b.learn_state(0)
    assert st["frozen"] and st["rounds"] >= 2
    burn = 40 * 50 + 1000          # past the freeze at iteration 2,000
    # exact posterior under each chain's own tau2 / alpha (drawn at set-up
    # from its seed, then held)
    for c in range(nc):
        d = oracle.Data(30, 7, body, y, recode=True)
        s = oracle.Sampler(cfgs[c], d)
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
        w = np.exp(lps - lps.max())
        w /= w.sum()
        exact = np.zeros(7)
        for wi, sub in zip(w, subs):
            exact[sub] += wi
        tr = b.trace(c)
        assert (tr["move"][burn:] == 3).sum() > 0.05 * (n_iter - burn)      # learned swaps drawn
        p_rao = b.p_rao(c)
        assert np.max(np.abs(p_rao - exact)) < 0.01, (c, p_rao, exact)
        th = b.thinned(c)
        sizes, loci = th["sizes"], th["loci"]
        off = int(sizes[:burn].sum())
        freq = np.bincount(loci[off:].astype(np.int64), minlength=7) / (len(sizes) - burn)
        assert np.max(np.abs(freq - exact)) < 0.03, (c, freq, exact)


def test_learned_snp_features_match_oracle(bmg):
    """The per-SNP proposal network's features, computed on the device in a
    chain's state during burn-in, against the oracle's restatement with the
    same inputs (the device's marginal scores, band, chains' mean table and
    the chain's model and running RB mean)."""
    n, p, H = 300, 500, 12
    body, y, causal = dataset(n, p, (0.9, 0.95, 10, 40), 6, 4321)
    cfg = synth.sampler_cfg(n, p, 6, seed=91, n_rao=100, n_rao_burnin=5, thin=10, max_move_size=10,
                            do_n_iter=400)
    d = oracle.Data(n, p, body, y)
    ctx = bmg.Ctx(n, p, body, y)
    b = bmg.Batch([ctx], [0], [cfg], dict(proposal=2, learn_window=H, learn_steps=10, learn_batch=256,
                                          learn_every=100))
    b.setup()
    b.run(300)                                  # three burn-in blocks: a model, a trained round
    z, pbar, band = b.learn_static(0)
    pre = oracle.DataModel(d, ["A"]).self_moments()
    loci = b.loci(0)
    assert len(loci) > 0
    mi = np.full(p, -1, np.int32)
    mi[loci] = np.arange(len(loci))
    pp = b.p_proposal(0)
    # the model SNPs, their neighbours and a spread of others
    js = sorted(set([int(j) for l in loci for j in range(max(0, l - H), min(p, l + H + 1))]
                    + list(range(0, p, 7))))
    fd = b.learn_snp_features(0, 0, js)
    for i, j in enumerate(js):
        fo = oracle.learn_snp_features(pre, 0, n, band, H, z, pbar, mi, pp, j)
        assert np.allclose(fd[i], fo, rtol=1e-5, atol=1e-5), (j, fd[i], fo)
