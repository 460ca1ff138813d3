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


def test_shuffles_known_answer_and_uniform():
    # hand-worked: draws 7, 3, 10, 1 on 5 rows swap (4, 2), (3, 3), (2, 1), (1, 1)
    assert mlp.shuffles(5, 1, [7, 3, 10, 1]) == [[0, 4, 1, 3, 2]]
    # successive epochs continue from the previous order
    two = mlp.shuffles(5, 2, [7, 3, 10, 1, 7, 3, 10, 1])
    assert two[0] == [0, 4, 1, 3, 2] and two[1] == [0, 2, 4, 3, 1]
    # Fisher-Yates reaches every permutation of 4 rows about equally often
    # (a swap partner drawn below i instead of i + 1 would give cycles only)
    rng = np.random.default_rng(5)
    counts = {}
    for _ in range(4800):
        p = tuple(mlp.shuffles(4, 1, rng.integers(0, 2 ** 62, 3))[0])
        counts[p] = counts.get(p, 0) + 1
    assert len(counts) == 24
    assert min(counts.values()) > 120 and max(counts.values()) < 280


def test_adam_constant_gradient_closed_form(monkeypatch):
    """With a constant gradient g the moments are m_k = (1 - b1^k) g and
    v_k = (1 - b2^k) g^2, so step k moves w by lr g / (|g| + eps / sqrt(1 - b2^k))."""
    g = np.array([1e-3, -2.0, 5e-7, 0.0, 3e-9])
    monkeypatch.setattr(mlp, "loss_grad", lambda X, t, w: (0.0, g.copy()))
    w0 = np.array([0.5, -0.25, 1.0, 2.0, 0.0])
    lr = 1e-2
    w, grads = mlp.adam_fit(np.zeros((8, 1)), np.zeros(8), w0, epochs=3, batch=4, lr=lr, draws=np.arange(21))
    assert len(grads) == 6
    want = w0.copy()
    for k in range(1, 7):
        want -= lr * g / (np.abs(g) + mlp.ADAM_EPS / np.sqrt(1 - mlp.ADAM_B2 ** k))
    np.testing.assert_allclose(w, want, rtol=1e-12, atol=1e-15)


def test_adam_batches_follow_the_shuffle(monkeypatch):
    seen = []

    def fake(X, t, w):
        seen.append(X[:, 0].astype(int).tolist())
        return 0.0, np.zeros_like(w)

    monkeypatch.setattr(mlp, "loss_grad", fake)
    X = np.arange(10, dtype=float)[:, None]
    draws = np.random.default_rng(1).integers(0, 2 ** 62, 18)
    mlp.adam_fit(X, np.zeros(10), np.zeros(3), epochs=2, batch=4, lr=1.0, draws=draws)
    o = mlp.shuffles(10, 2, draws)
    # two full batches per epoch; the last two rows of each shuffle are left out
    assert seen == [o[0][0:4], o[0][4:8], o[1][0:4], o[1][4:8]]


def test_node_features_scaling():
    x, y, z, dim, cloud, h, Lp, res, w = _node_case(seed=4)
    i = 7
    f, sigma = mlp.node_features(x, y, z, dim, cloud, h, res, i)
    m = cloud.shape[1]
    off = f[:m * dim].reshape(m, dim)
    np.testing.assert_array_equal(off[0], 0.0)                      # the node itself leads its stencil
    np.testing.assert_allclose(off[3], [(x[cloud[i, 3]] - x[i]) / h, (y[cloud[i, 3]] - y[i]) / h], rtol=1e-15)
    u = f[m * dim:]
    assert np.isclose(np.sqrt(np.mean(u * u)), 1.0, rtol=1e-14)     # residuals over their RMS
    assert np.isclose(sigma, np.sqrt(np.mean(res[cloud[i]] ** 2)), rtol=1e-15)
    res0 = res.copy()
    res0[cloud[i]] = 0.0
    f0, s0 = mlp.node_features(x, y, z, dim, cloud, h, res0, i)
    assert s0 == 0 and not f0[m * dim:].any()
