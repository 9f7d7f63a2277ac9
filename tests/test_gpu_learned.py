"""The learned swap on the GPU against the oracle: with the same frozen
weight table and random stream, every iteration (learned swaps included)
takes the same move, jump and model size, log-likelihoods within 1e-10."""
import numpy as np
import pytest

from test_gpu_sampler import compare, dataset, oracle_run  # noqa: F401  (fixtures)
from bmagwa_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bmg():
    import bmagwa_amd
    bmagwa_amd.lib()
    return bmagwa_amd


def skewed_table(p, H, seed):
    rng = np.random.default_rng(seed)
    w = np.exp(rng.normal(scale=1.5, size=(p, 2 * H))).astype(np.float32)
    w[rng.random(size=w.shape) < 0.1] = 0.0
    return w


@pytest.mark.parametrize("H,p_learned,use_ind", [(8, 0.4, 1), (32, 0.9, 0), (1, 0.5, 1)])
def test_learned_swap_trajectory_parity(bmg, H, p_learned, use_ind):
    n, p, n_iter = 300, 400, 3000
    body, y, _ = dataset(n, p, (0.9, 0.9, 10, 30), 5, 4321)
    cfg = synth.sampler_cfg(n, p, 5, seed=91, n_rao=100, n_rao_burnin=5, thin=10,
                            max_move_size=10, use_individual_tau2=use_ind, do_n_iter=n_iter)
    w = skewed_table(p, H, 7 + H)
    import oracle
    d = oracle.Data(n, p, body, y)
    s = oracle.Sampler(cfg, d)
    s.init_flat()
    assert s.setup() == 0
    assert s.set_learned(w, H, p_learned) == 0
    s.iterate(n_iter)
    ctx = bmg.Ctx(n, p, body, y)
    b = bmg.Batch([ctx], [0], [cfg], dict(trace_every_iter=1))
    b.setup()
    b.set_learned(0, w, H, p_learned)
    b.run(n_iter)
    tr = b.trace(0)
    compare(tr, s.trace(), n_iter)
    assert (tr["move"] == 3).sum() > 0.5 * 0.15 * p_learned * n_iter   # a share of the swaps
    assert ((tr["move"] == 3) & (tr["jd"] > 0)).sum() > 0


def test_learned_swap_switched_off_is_plain(bmg):
    n, p, n_iter = 256, 300, 1000
    body, y, _ = dataset(n, p, (0.9, 0.9, 10, 30), 4, 99)
    cfg = synth.sampler_cfg(n, p, 4, seed=5, n_rao=100, n_rao_burnin=5, thin=10,
                            max_move_size=8, do_n_iter=n_iter)
    ctx = bmg.Ctx(n, p, body, y)
    b = bmg.Batch([ctx], [0], [cfg], dict(trace_every_iter=1))
    b.setup()
    b.set_learned(0, skewed_table(p, 4, 1), 4, 0.5)
    b.set_learned(0, None, 4, 0.5)
    b.run(n_iter)
    s = oracle_run(cfg, body, y, n_iter)
    compare(b.trace(0), s.trace(), n_iter)
