"""Online learned proposal on the GPU against the oracle: marginal scores,
band X^T X, pair features and the swap labels of the first burn-in RB block
(the device and oracle trajectories coincide until the first table exists);
then a full burn-in: training rounds, freezing, a finite table, learned swaps
accepted, PIPs close to a plain run."""
import numpy as np
import pytest

import oracle
from bmagwa_amd import synth
from test_gpu_sampler import dataset

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bmg():
    import bmagwa_amd
    bmagwa_amd.lib()
    return bmagwa_amd


H = 12
LCFG = dict(proposal=1, learn_window=H, learn_steps=20, learn_batch=1024, p_learned=0.8,
            learn_every=100, trace_every_iter=1)


def setup_pair(bmg, n, p, n_iter, seed=4321, burnin=3):
    body, y, causal = dataset(n, p, (0.9, 0.95, 10, 40), 6, seed)
    cfg = synth.sampler_cfg(n, p, 6, seed=91, n_rao=100, n_rao_burnin=burnin, thin=10,
                            max_move_size=10, use_individual_tau2=0, do_n_iter=n_iter)
    d = oracle.Data(n, p, body, y)
    s = oracle.Sampler(cfg, d)
    s.init_flat()
    assert s.setup() == 0
    ctx = bmg.Ctx(n, p, body, y)
    b = bmg.Batch([ctx], [0], [cfg], LCFG)
    b.setup()
    return s, d, b, body, y, causal, cfg


def test_static_inputs_and_features(bmg):
    n, p = 300, 500
    s, d, b, body, y, _, _ = setup_pair(bmg, n, p, 200)
    z, pbar, band = b.learn_static(0)
    zo = s.marginal_z()
    assert np.allclose(z, zo, rtol=1e-5, atol=1e-4)
    bo = np.zeros(p * (H + 1), np.int64)
    oracle.lib().orc_band_xtx(oracle.u8(body), (n + 3) // 4, n, p, H + 1, 0, 0,
                              oracle.ptr(bo, oracle.C.c_int64))
    bo = bo.reshape(p, H + 1)
    for j in range(p):
        w = min(H + 1, p - j)
        assert np.array_equal(band[j, :w], bo[j, :w]), j
    pre = oracle.DataModel(d, ["A"]).self_moments()
    rng = np.random.default_rng(3)
    js, os_ = [], []
    while len(js) < 400:
        j = int(rng.integers(0, p)); o = int(rng.integers(-H, H + 1))
        if o != 0 and 0 <= j + o < p:
            js.append(j); os_.append(o)
    fd = b.learn_features(0, js, os_)
    pb = rng.uniform(0, 1, p).astype(np.float32)  # pbar on the device is flat here
    pb[:] = pbar
    for i, (j, o) in enumerate(zip(js, os_)):
        fo = oracle.learn_pair_features(pre, 0, n, band, H, z, pb, j, o)
        assert np.allclose(fd[i], fo, rtol=1e-5, atol=1e-5), (j, o, fd[i], fo)


def test_first_block_labels_match_oracle(bmg):
    n, p = 300, 500
    s, d, b, *_ = setup_pair(bmg, n, p, 100)
    b.run(100)                      # one segment, then the first (burn-in) RB block
    s.iterate(100)
    jg, og, lg, cnt = b.learn_samples(0)
    jo, oo, lo = s.swap_labels(H)
    assert cnt == len(jo) > 0
    kg = np.lexsort((og, jg)); ko = np.lexsort((oo, jo))
    assert np.array_equal(jg[kg], jo[ko]) and np.array_equal(og[kg], oo[ko])
    assert np.allclose(lg[kg], lo[ko], rtol=1e-5, atol=1e-4), np.max(np.abs(lg[kg] - lo[ko]))
    st = b.learn_state(0)
    assert st["rounds"] == 1 and not st["frozen"]


def test_burn_in_trains_freezes_and_keeps_posterior(bmg):
    n, p, n_iter = 400, 1500, 12000
    body, y, causal = dataset(n, p, (0.9, 0.95, 10, 40), 6, 77)
    runs = {}
    for prop in [0, 1]:
        cfgs = [synth.sampler_cfg(n, p, 6, seed=300 + c, chain=c, n_rao=100, n_rao_burnin=10,
                                  thin=10, max_move_size=10, do_n_iter=n_iter) for c in range(4)]
        ctx = bmg.Ctx(n, p, body, y)
        bc = dict(LCFG, proposal=prop, trace_every_iter=1)
        b = bmg.Batch([ctx], [0] * 4, cfgs, bc)
        b.setup()
        b.run(n_iter)
        pips = np.mean([b.p_rao(c) for c in range(4)], axis=0)
        runs[prop] = (b, pips)
        if prop == 1:
            st = b.learn_state(0)
            assert st["frozen"] and st["rounds"] >= 2 and st["labels"] > 0
            w = b.learn_table(0)
            assert np.all(np.isfinite(w)) and np.all(w >= 0)
            assert np.all(w[:H, :H][np.tril_indices(H)] >= 0)
            tr = b.trace(0)
            learned = tr["move"] == 3
            assert learned.sum() > 0.05 * n_iter      # 0.15 x 0.8 of the moves
            assert (learned & (tr["jd"] > 0)).sum() > 0
    diff = np.abs(runs[0][1] - runs[1][1])
    assert diff.max() < 0.1, diff.max()


def test_learned_run_is_reproducible(bmg):
    """Two learned batches with the same seeds train the same network and run
    the same chains: the label ring is filled in a fixed order and the
    network's gradients are summed in a fixed order, so the ring, the
    parameters, the frozen table and every trajectory are identical."""
    n, p, n_iter = 300, 900, 3000
    body, y, _ = dataset(n, p, (0.9, 0.95, 10, 40), 6, 31)
    outs = []
    for rep in range(2):
        cfgs = [synth.sampler_cfg(n, p, 6, seed=40 + c, chain=c, n_rao=100, n_rao_burnin=6, thin=10,
                                  max_move_size=10, do_n_iter=n_iter) for c in range(3)]
        ctx = bmg.Ctx(n, p, body, y)
        b = bmg.Batch([ctx], [0] * 3, cfgs, dict(LCFG, use_graph=1))
        b.setup()
        b.run(n_iter)
        st = b.learn_state(0)
        assert st["frozen"] and st["rounds"] >= 2
        outs.append(dict(ring=b.learn_samples(0), params=b.learn_params(0), table=b.learn_table(0),
                         traces=[b.trace(c) for c in range(3)], p_rao=[b.p_rao(c) for c in range(3)]))
        b.close()
    a, c = outs
    for x, z in zip(a["ring"], c["ring"]):
        assert np.array_equal(np.asarray(x), np.asarray(z))
    assert np.array_equal(a["params"], c["params"])
    assert np.array_equal(a["table"], c["table"])
    for ta, tc in zip(a["traces"], c["traces"]):
        for key in ta:
            assert np.array_equal(ta[key], tc[key]), key
    for ra, rc in zip(a["p_rao"], c["p_rao"]):
        assert np.array_equal(ra, rc)


def test_trained_network_gains_over_the_move_without_it(bmg):
    """Ablation on an LD dataset: the same learned-swap move with the trained
    network's table against the same move with a flat table (every partner
    within H equally likely: the move without the network) and against the
    plain swap.  Same chains, seeds and iterations; every drawn swap is the
    variant under test (about 12,000 swaps each, so an acceptance rate has a
    standard error near 0.4 %); the trained table's swaps are accepted more
    often than either."""
    n, p, burn, n_iter, nch = 600, 3000, 3000, 10000, 8
    body, y, _ = dataset(n, p, (0.9, 0.95, 10, 40), 10, 2024)
    cfgs = [synth.sampler_cfg(n, p, 10, seed=60 + c, chain=c, n_rao=100, n_rao_burnin=burn // 100, thin=10,
                              max_move_size=10, do_n_iter=burn + n_iter + 10) for c in range(nch)]
    ctx = bmg.Ctx(n, p, body, y)
    b = bmg.Batch([ctx], [0] * nch, cfgs, dict(LCFG, p_learned=1.0))
    b.setup()
    b.run(burn)
    assert b.learn_state(0)["frozen"]
    trained = b.learn_table(0)
    b.close()

    def accept(table):
        bb = bmg.Batch([bmg.Ctx(n, p, body, y)], [0] * nch, cfgs, dict(trace_every_iter=1))
        bb.setup()
        if table is not None:
            bb.set_learned(0, table, H, 1.0)
        bb.run(burn + n_iter)
        mv = 3 if table is not None else 1
        acc = tot = 0
        for c in range(nch):
            tr = bb.trace(c)
            m, jd = tr["move"][burn:], tr["jd"][burn:]
            acc += int(((m == mv) & (jd > 0)).sum())
            tot += int((m == mv).sum())
        bb.close()
        assert tot > 8000
        return acc / tot

    flat = np.ones_like(trained)
    flat[trained == 0] = 0.0        # same support: partners inside the genome
    r_trained, r_flat, r_plain = accept(trained), accept(flat), accept(None)
    assert r_trained > r_flat + 0.015, (r_trained, r_flat, r_plain)
    assert r_trained > 1.3 * r_plain, (r_trained, r_flat, r_plain)
