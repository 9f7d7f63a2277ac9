"""The learned swap move of the oracle (weights frozen) leaves the posterior
unchanged: its Metropolis-Hastings ratio carries the forward and reverse
partner-draw probabilities.  Checked against exact enumeration of all 2^7
models with tau2/alpha held fixed, with deliberately skewed weights.

The runs use add/remove moves and swaps half and half, every swap a learned
swap, and no state changes (move mix 0.5/0.5/0): the plain swap's walk over
free neighbours is not exactly reversible, which biases even the plain
sampler by ~0.015 here, and would hide the learned move's own error.  With the Hastings term of the learned move
dropped, the same runs miss the exact PIPs by 0.08-0.11 (checked once by
editing the oracle), far outside the 0.02 tolerance."""
import numpy as np
import pytest

import oracle
from test_oracle_sampler import exact_pips, small_sampler


def skewed_weights(p, H, seed):
    rng = np.random.default_rng(seed)
    w = np.exp(rng.normal(scale=2.0, size=(p, 2 * H))).astype(np.float32)
    w[rng.random(size=w.shape) < 0.2] = 0.0     # some partners never proposed
    return w


def freq_after_burn(s, burn):
    th = s.thinned()
    counts = np.zeros(7)
    off = kept = 0
    for t, k in enumerate(th["sizes"]):
        if t >= burn:
            counts[th["loci"][off:off + k]] += 1
            kept += 1
        off += k
    return counts / kept


@pytest.mark.parametrize("H,seed", [(3, 11), (6, 12)])
def test_learned_swap_pips_match_exact_enumeration(H, seed):
    # every drawn swap is a learned swap (the plain swap is left out: its
    # own exact-enumeration error would hide the learned move's)
    p_learned = 1.0
    s, d = small_sampler(dict(n_sample_tau2_and_missing=10 ** 9, n_rao=50, n_rao_burnin=100,
                              thin=1, delay_rejection=0, max_move_size=4, seed=seed))
    s.init_flat()
    s.setup()
    assert s.set_move_mix([0.5, 0.5, 0.0]) == 0
    assert s.set_learned(skewed_weights(7, H, seed), H, p_learned) == 0
    s.iterate(400000)
    exact = exact_pips(s, d, s.prior()["inv_tau2_alpha2_A"])
    freq = freq_after_burn(s, 20000)
    assert np.max(np.abs(freq - exact)) < 0.02, (freq, exact)
    tr = s.trace()
    n_learned = int((tr["move"] == 3).sum())
    assert n_learned > 0.8 * 0.5 * p_learned * len(tr["move"])
    # learned swaps were accepted (the move is live, not a no-op)
    assert int(((tr["move"] == 3) & (tr["jd"] > 0)).sum()) > 1000


def test_learned_swap_off_is_the_plain_sampler():
    """set_learned(None) leaves the random stream and trajectory untouched."""
    runs = []
    for switch in [False, True]:
        s, d = small_sampler(dict(seed=5, delay_rejection=0))
        s.init_flat()
        s.setup()
        if switch:
            s.set_learned(None, 3, 0.5)
        s.iterate(3000)
        runs.append(s.trace())
    for k in ["move", "jd", "size", "loglik"]:
        assert np.array_equal(runs[0][k], runs[1][k])


def test_set_learned_rejects_bad_arguments():
    s, d = small_sampler()
    assert s.set_learned(np.ones((7, 2 * 40), np.float32), 40, 0.5) != 0
    assert s.set_learned(np.ones((7, 6), np.float32), 3, 1.5) != 0
