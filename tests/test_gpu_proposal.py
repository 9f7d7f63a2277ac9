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
