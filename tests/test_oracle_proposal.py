"""A proposal table from outside the sampler (the learned per-SNP proposal
sets one at the end of burn-in) leaves the posterior unchanged: q_add and
q_rem enter the add/remove move's Hastings ratio and the delayed-rejection
stage's proposal probabilities (src/sampler.hpp:1126-1202,
src/sampler.cpp:982-1049).  Checked against exact enumeration of all 2^7
models with tau2/alpha held fixed, with a deliberately skewed table; the
same run with the removal weights dropped from the reverse proposal term
misses the exact PIPs by 0.2 (checked once by editing the oracle), ten
times the tolerance."""
import numpy as np
import pytest

from test_oracle_learned import freq_after_burn
from test_oracle_sampler import exact_pips, small_sampler


def skewed_table(p, seed):
    rng = np.random.default_rng(seed)
    t = rng.uniform(0.0, 1.0, size=p) ** 3     # most SNPs rarely proposed
    t[rng.integers(p)] = 1.0                   # one SNP removed only at the floor rate q_rem_min
    return t


@pytest.mark.parametrize("dr,seed", [(0, 21), (3, 22)])
def test_fixed_proposal_table_keeps_the_posterior(dr, seed):
    # no adaptation at all (n_rao_burnin = 0, adaptation off): the table set
    # after setup is the proposal for the whole run
    s, d = small_sampler(dict(n_sample_tau2_and_missing=10 ** 9, n_rao=50, n_rao_burnin=0,
                              thin=1, delay_rejection=dr, max_move_size=4, seed=seed))
    s.init_flat()
    s.setup()
    assert s.set_move_mix([1.0, 0.0, 0.0]) == 0
    table = skewed_table(7, seed)
    assert s.set_p_proposal(table) == 0
    assert np.array_equal(s.p_proposal(), table)
    s.iterate(300000)
    assert np.array_equal(s.p_proposal(), table)     # not adapted afterwards
    exact = exact_pips(s, d, s.prior()["inv_tau2_alpha2_A"])
    freq = freq_after_burn(s, 20000)
    assert np.max(np.abs(freq - exact)) < 0.02, (freq, exact)


def test_set_p_proposal_rejects_values_outside_the_unit_interval():
    s, d = small_sampler()
    s.init_flat()
    s.setup()
    before = s.p_proposal()
    bad = np.full(7, 0.5)
    bad[3] = 1.5
    assert s.set_p_proposal(bad) != 0
    bad[3] = np.nan
    assert s.set_p_proposal(bad) != 0
    assert np.array_equal(s.p_proposal(), before)
