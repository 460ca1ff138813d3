"""CPU oracle of the learned smoother's network.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the GPU network
kernels; the product package never imports it.

There is no reference counterpart: the reference smooths with SOR only.  This
float64 numpy restatement defines the network the GPU kernels implement
(multigrid_meshless_amd/csrc/mlp.hip) and is itself pinned by central finite
differences of the loss (tests/test_oracle_mlp.py).

Training: mean squared loss, its gradient, and Adam over shuffled batches
(adam_fit), pinned by finite differences and closed forms of the first
steps (tests/test_oracle_mlp.py).

Weight vector layout: W1[H][in], b1[H], W2[H][H], b2[H], w3[H], b3 (H = 64).
Node features on a level with stencil width m (dimension dim):
  offsets (x_cj - x_i) / h for j < m and each axis, then r_cj / sigma_i for
  j < m, sigma_i the RMS of the stencil residuals; the node's correction is
  f(features) * sigma_i / a_ii.
"""
from __future__ import annotations

import numpy as np

H = 64


def n_weights(n_in: int) -> int:
    return H * n_in + H + H * H + H + H + 1


def unpack(w: np.ndarray, n_in: int):
    w = np.asarray(w, dtype=np.float64)
    o = 0
    W1 = w[o:o + H * n_in].reshape(H, n_in); o += H * n_in
    b1 = w[o:o + H]; o += H
    W2 = w[o:o + H * H].reshape(H, H); o += H * H
    b2 = w[o:o + H]; o += H
    w3 = w[o:o + H]; o += H
    b3 = w[o]
    return W1, b1, W2, b2, w3, b3


def forward(X: np.ndarray, w: np.ndarray):
    """y = w3 . relu(W2 relu(W1 x + b1) + b2) + b3 for each row x of X."""
    X = np.asarray(X, dtype=np.float64)
    W1, b1, W2, b2, w3, b3 = unpack(w, X.shape[1])
    Z1 = X @ W1.T + b1
    A1 = np.maximum(Z1, 0)
    Z2 = A1 @ W2.T + b2
    A2 = np.maximum(Z2, 0)
    return A2 @ w3 + b3, (Z1, A1, Z2, A2)


def loss(X, t, w) -> float:
    y, _ = forward(X, w)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def loss_grad(X: np.ndarray, t: np.ndarray, w: np.ndarray):
    """Mean squared error and its gradient with respect to the weight vector."""
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[0]
    W1, b1, W2, b2, w3, b3 = unpack(w, X.shape[1])
    y, (Z1, A1, Z2, A2) = forward(X, w)
    dy = 2.0 * (y - np.asarray(t, dtype=np.float64)) / n
    dw3 = A2.T @ dy
    db3 = dy.sum()
    dZ2 = np.outer(dy, w3) * (Z2 > 0)
    dW2 = dZ2.T @ A1
    db2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * (Z1 > 0)
    dW1 = dZ1.T @ X
    db1 = dZ1.sum(0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dw3, [db3]])
    return float(np.mean((y - t) ** 2)), g


# Adam constants as the GPU holds them (FP32 literals)
ADAM_B1 = float(np.float32(0.9))
ADAM_B2 = float(np.float32(0.999))
ADAM_EPS = float(np.float32(1e-8))


def shuffles(n: int, epochs: int, draws) -> list:
    """Row order of each epoch: Fisher-Yates from the last row down, swap
    partner draw % (i + 1), with `draws` the consecutive splitmix64 outputs
    of the seed (epochs * (n - 1) of them)."""
    order = list(range(n))
    out, q = [], 0
    for _ in range(epochs):
        for i in range(n - 1, 0, -1):
            j = int(draws[q]) % (i + 1)
            q += 1
            order[i], order[j] = order[j], order[i]
        out.append(list(order))
    return out


def adam_fit(X, t, w, epochs: int, batch: int, lr: float, draws):
    """Adam on the MSE loss over rows (X, t): each epoch takes floor(n / batch)
    steps over consecutive batches of its shuffle; step k moves
    w -= lr sqrt(1 - b2^k) / (1 - b1^k) m / (sqrt(v) + eps) with
    m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2 from zero.  Returns the
    final weights and every step's gradient."""
    X = np.asarray(X, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64).copy()
    n = X.shape[0]
    B = min(batch, n)
    M = np.zeros_like(w)
    V = np.zeros_like(w)
    grads, k = [], 0
    for order in shuffles(n, epochs, draws):
        for b0 in range(0, n - B + 1, B):
            k += 1
            idx = order[b0:b0 + B]
            _, g = loss_grad(X[idx], t[idx], w)
            M = ADAM_B1 * M + (1 - ADAM_B1) * g
            V = ADAM_B2 * V + (1 - ADAM_B2) * g * g
            lr_t = lr * np.sqrt(1 - ADAM_B2 ** k) / (1 - ADAM_B1 ** k)
            w -= lr_t * M / (np.sqrt(V) + ADAM_EPS)
            grads.append(g)
    return w, grads


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def node_features(x, y, z, dim, cloud, h, res, i):
    """Feature vector of node i (reference order arrays) and its sigma."""
    c = cloud[i]
    crd = [x, y, z][:dim]
    off = np.stack([(a[c] - a[i]) / h for a in crd], 1).ravel()
    r = res[c]
    sigma = np.sqrt(np.mean(r * r))
    # scaled by the reciprocal, as the kernels do (rounding points of BF16 inputs agree)
    return np.concatenate([off, r * (1.0 / sigma) if sigma > 0 else 0 * r]), sigma


def correction(x, y, z, dim, cloud, h, Lp, res, w, i, bf16_weights=True, bf16_activations=True):
    """The learned correction of interior node i: f(features) * sigma / a_ii,
    as the inference kernel evaluates it: BF16-rounded W2, w3 and residual
    columns of W1 (bf16_weights) and BF16-rounded matrix-core inputs -- the
    scaled residuals and the first hidden layer -- (bf16_activations), with
    sums in float64 here (FP32 on the GPU)."""
    feats, sigma = node_features(x, y, z, dim, cloud, h, res, i)
    if sigma == 0:
        return 0.0
    w = np.asarray(w, dtype=np.float32).copy()
    n_in = feats.size
    m = cloud.shape[1]
    if bf16_weights:
        W1 = w[:H * n_in].reshape(H, n_in)
        W1[:, m * dim:] = bf16_round(W1[:, m * dim:])
        o = H * n_in + H
        w[o:o + H * H] = bf16_round(w[o:o + H * H])
        o += H * H + H
        w[o:o + H] = bf16_round(w[o:o + H])
    if not bf16_activations:
        out, _ = forward(feats[None, :], w.astype(np.float64))
        return float(out[0]) * sigma / Lp[i, 0]
    W1, b1, W2, b2, w3, b3 = unpack(w.astype(np.float64), n_in)
    u = bf16_round(feats[m * dim:].astype(np.float32)).astype(np.float64)
    z1 = W1[:, :m * dim] @ feats[:m * dim] + b1 + W1[:, m * dim:] @ u
    a1 = bf16_round(np.maximum(z1, 0).astype(np.float32)).astype(np.float64)
    a2 = np.maximum(W2 @ a1 + b2, 0)
    return float(a2 @ w3 + b3) * sigma / Lp[i, 0]
