"""The network oracle (oracle/mlp.py), pinned without the GPU: its analytic
loss gradient against central finite differences, and its BF16 rounding
against hand-computed values."""
import numpy as np

from oracle import mlp


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    n_in, n = 12, 40
    X = rng.standard_normal((n, n_in))
    t = rng.standard_normal(n)
    w = rng.standard_normal(mlp.n_weights(n_in)) * 0.3
    _, g = mlp.loss_grad(X, t, w)
    eps = 1e-6
    for k in rng.choice(w.size, 60, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[k] += eps
        wm[k] -= eps
        fd = (mlp.loss(X, t, wp) - mlp.loss(X, t, wm)) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_gradient_layout_and_bias_terms():
    # a network whose output is b3 alone: only db3 is nonzero, equal to 2*mean(y - t)
    n_in = 5
    w = np.zeros(mlp.n_weights(n_in))
    w[-1] = 0.5
    X = np.ones((4, n_in))
    t = np.array([0.0, 1.0, 2.0, 3.0])
    l, g = mlp.loss_grad(X, t, w)
    assert np.allclose(g[:-1], 0)
    assert np.isclose(g[-1], 2 * np.mean(0.5 - t))
    assert np.isclose(l, np.mean((0.5 - t) ** 2))


def test_bf16_rounding():
    a = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 1 + 2 ** -7], np.float32)
    np.testing.assert_array_equal(mlp.bf16_round(a),
                                  np.array([1.0, 1.0, 1 + 2 ** -6, -2.5, 1 + 2 ** -7], np.float32))


def _node_case(seed=3, m=12, dim=2):
    rng = np.random.default_rng(seed)
    N = 40
    x, y, z = rng.random(N), rng.random(N), np.zeros(N)
    cloud = np.stack([np.concatenate([[i], rng.choice(np.delete(np.arange(N), i), m - 1, replace=False)])
                      for i in range(N)])
    Lp = rng.standard_normal((N, m)) + np.eye(N, m) * 4
    res = rng.standard_normal(N)
    w = rng.standard_normal(mlp.n_weights(m * dim + m)) * 0.2
    return x, y, z, dim, cloud, 0.1, Lp, res, w


def test_correction_bf16_activation_path(monkeypatch):
    """The BF16-activation evaluation is the network forward pass with two
    rounding points: with the BF16 rounding made the identity it equals
    forward() (pinned by finite differences above) up to the FP32 casts the
    kernel also applies before rounding; with BF16 rounding it moves the
    correction only by BF16-sized amounts."""
    args = _node_case()
    exact = mlp.correction(*args, 5, bf16_weights=False, bf16_activations=False)
    rounded = mlp.correction(*args, 5, bf16_weights=False, bf16_activations=True)
    assert exact != rounded
    assert abs(rounded - exact) <= 2e-2 * abs(exact)
    monkeypatch.setattr(mlp, "bf16_round", lambda a: np.asarray(a, dtype=np.float64))
    same = mlp.correction(*args, 5, bf16_weights=False, bf16_activations=True)
    assert abs(same - exact) <= 1e-6 * abs(exact)
