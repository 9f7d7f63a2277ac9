"""The learned-proposal network's bf16 MFMA kernels against the CPU oracle
(same rounding points, fp32 vs double accumulation): forward outputs,
gradients of every parameter block, and Adam steps."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mlp():
    import bmagwa_amd
    m = bmagwa_amd.Mlp()
    yield m
    m.close()


BLOCKS = [("W1", 0, 4096), ("b1", 4096, 4352), ("W2", 4352, 4352 + 65536),
          ("b2", 69888, 70144), ("w3", 70144, 70400), ("b3", 70400, 70401)]


def data(seed, rows):
    rng = np.random.default_rng(seed)
    X = (rng.normal(size=(rows, 16)) * rng.uniform(0.1, 3.0, size=16)).astype(np.float32)
    lab = (rng.normal(size=rows) * 3.0).astype(np.float32)
    return X, lab


@pytest.mark.parametrize("rows", [1, 64, 1000, 4096])
def test_forward_matches_oracle(mlp, rows):
    p = oracle.mlp_init(rows, scale3=0.3)
    X, _ = data(rows, rows)
    mlp.set_params(p)
    got = mlp.forward(X)
    want = oracle.mlp_forward(p, X)
    assert np.allclose(got, want, rtol=2e-3, atol=2e-3 * np.abs(want).max()), \
        np.max(np.abs(got - want))


@pytest.mark.parametrize("rows", [64, 777, 4096])
def test_gradients_match_oracle(mlp, rows):
    p = oracle.mlp_init(10 + rows, scale3=0.3)
    X, lab = data(20 + rows, rows)
    mlp.set_params(p)
    got = mlp.gradients(X, lab)
    _, want = oracle.mlp_gradients(p, X, lab)
    for name, a, b in BLOCKS:
        err = np.linalg.norm(got[a:b] - want[a:b]) / max(np.linalg.norm(want[a:b]), 1e-12)
        assert err < 5e-3, (name, err)


def test_adam_steps_match_oracle(mlp):
    p = oracle.mlp_init(7, scale3=0.3)
    X, lab = data(8, 2048)
    mlp.set_params(p)
    q = p.copy()
    m = np.zeros_like(q)
    v = np.zeros_like(q)
    for t in range(1, 4):
        _, g = oracle.mlp_gradients(q, X, lab)
        oracle.mlp_adam(q, m, v, g.astype(np.float32), 1e-3, t)
        mlp.adam_step(X, lab, 1e-3)
    got = mlp.get_params()
    # every parameter moved by about lr per step; the two agree far closer
    assert np.max(np.abs(got - q)) < 3e-4
    assert np.median(np.abs(got - p)) > 1e-3


def test_device_training_reduces_the_loss(mlp):
    rng = np.random.default_rng(5)
    p = oracle.mlp_init(5)
    X = rng.normal(size=(4096, 16)).astype(np.float32)
    lab = (np.tanh(X[:, 0]) * 3 + X[:, 1] * X[:, 2]).astype(np.float32)
    mlp.set_params(p)
    l0 = oracle.mlp_gradients(p, X, lab)[0]
    for _ in range(100):
        mlp.adam_step(X, lab, 3e-3)
    l1 = oracle.mlp_gradients(mlp.get_params(), X, lab)[0]
    assert l1 < 0.3 * l0, (l0, l1)
