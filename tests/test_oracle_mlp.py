"""The oracle of the learned-proposal network: bf16 rounding against its
definition, the exact gradients against central finite differences, the
bf16-rounded gradients against the exact ones, and Adam against its
published update rule.  (The network has no reference counterpart.)"""
import numpy as np
import pytest

import oracle


def bf16_numpy(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32)


def test_bf16_round_to_nearest_even():
    f = oracle.lib().orc_bf16_round
    assert f(1.0) == 1.0
    assert f(1.0 + 2 ** -8) == 1.0                      # tie -> even (1.0)
    assert f(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6        # tie -> even (1 + 2^-6)
    assert f(1.0 + 2 ** -8 + 2 ** -12) == 1.0 + 2 ** -7  # above the tie
    assert f(-2.5) == -2.5
    assert np.isnan(f(float("nan")))
    rng = np.random.default_rng(0)
    x = (rng.normal(size=2000) * 10.0 ** rng.integers(-8, 8, 2000)).astype(np.float32)
    got = np.array([f(float(v)) for v in x], np.float32)
    assert np.array_equal(got, bf16_numpy(x))


def problem(seed, B=96):
    rng = np.random.default_rng(seed)
    p = oracle.mlp_init(seed, scale3=0.3)
    X = rng.normal(size=(B, oracle.MLP_IN)).astype(np.float32)
    lab = (rng.normal(size=B) * 2.0).astype(np.float32)
    return p, X, lab


def test_exact_gradients_match_finite_differences():
    p, X, lab = problem(1)
    loss, g = oracle.mlp_gradients(p, X, lab, exact=True)
    rng = np.random.default_rng(2)
    idx = []
    for name, size in [("W1", 16 * 256), ("b1", 256), ("W2", 256 * 256), ("b2", 256),
                       ("w3", 256), ("b3", 1)]:
        idx += list(oracle.MLP_OFF[name] + rng.integers(0, size, 6))
    for i in idx:
        h = 2e-4 * max(1.0, abs(float(p[i])))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        hp, hm = float(pp[i]) - float(p[i]), float(p[i]) - float(pm[i])
        fd = (oracle.mlp_loss_exact(pp, X, lab) - oracle.mlp_loss_exact(pm, X, lab)) / (hp + hm)
        assert fd == pytest.approx(g[i], rel=2e-2, abs=2e-6), (i, fd, g[i])


def test_rounded_gradients_close_to_exact():
    p, X, lab = problem(3, B=256)
    _, ge = oracle.mlp_gradients(p, X, lab, exact=True)
    _, gr = oracle.mlp_gradients(p, X, lab, exact=False)
    for name, (a, b) in [("W1", (0, 4096)), ("W2", (4352, 4352 + 65536)), ("w3", (70144, 70400))]:
        err = np.linalg.norm(gr[a:b] - ge[a:b]) / np.linalg.norm(ge[a:b])
        assert err < 0.03, (name, err)


def test_adam_update_rule():
    rng = np.random.default_rng(4)
    p = rng.normal(size=oracle.MLP_NPAR).astype(np.float32)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    p0 = p.copy()
    g = rng.normal(size=oracle.MLP_NPAR).astype(np.float32)
    oracle.mlp_adam(p, m, v, g, 1e-3, 1)
    # first step: m_hat = g, v_hat = g^2 -> p -= lr * g / (|g| + eps)
    assert np.allclose(p, p0 - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-6, atol=1e-7)
    assert np.allclose(m, 0.1 * g, rtol=1e-6)


def test_training_reduces_the_loss():
    rng = np.random.default_rng(5)
    p = oracle.mlp_init(5)
    X = rng.normal(size=(256, 16)).astype(np.float32)
    lab = (np.tanh(X[:, 0]) * 3 + X[:, 1] * X[:, 2]).astype(np.float32)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    l0 = oracle.mlp_gradients(p, X, lab)[0]
    for t in range(1, 61):
        _, g = oracle.mlp_gradients(p, X, lab)
        oracle.mlp_adam(p, m, v, g.astype(np.float32), 3e-3, t)
    assert oracle.mlp_gradients(p, X, lab)[0] < 0.3 * l0
