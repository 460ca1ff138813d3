"""Learned smoother on the GPU against the network oracle (oracle/mlp.py).

No reference counterpart (the reference smooths with SOR only).  Bars:
  * forward outputs, loss and loss gradient of the FP32 training kernels equal
    the float64 oracle to 1e-5 / 1e-4 relative, at 60 and 280 inputs;
  * the training rows (offsets, scaled residuals, targets) equal the oracle's
    node features to FP32 rounding, and Adam over shuffled batches moves the
    weights as the oracle's Adam does;
  * one learned sweep changes p on every interior node by the oracle's
    correction (BF16 matrix cores: BF16 weights and activations, FP32
    accumulation) to 1e-3 of the largest correction, and leaves boundary
    nodes alone;
  * training on (error, residual) pairs of the plain V-cycle lowers the loss;
    the learned solve reaches the tolerance, and what it gains over plain SOR
    (cycles, seconds, or a fallback) is printed.
"""
import os
import sys

import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def problem(mgm, dim, sizes, mult=None):
    pts = [mgm.Points.generate(dim, n, 11 + k, 0.2) for k, n in enumerate(sizes)]
    cfg = mgm.config(dimension=dim, poly_degree=3, coarse_poly_degree=3, phs_degree=3,
                     cloud_size_multiplier=mult or (3.5 if dim == 3 else 2.0), omega=1.0, num_relax=3,
                     transfer_rows=mgm.ROWS_BY_TAG, sweep_order=mgm.ORDER_COLOUR)
    P = mgm.Problem(cfg, pts).setup()
    P.dirichlet_rows()
    return P


def he_weights(n_in, seed):
    rng = np.random.default_rng(seed)
    w = np.zeros(mlp.n_weights(n_in), np.float32)
    W1, b1, W2, b2, w3, b3 = mlp.unpack(np.zeros(mlp.n_weights(n_in)), n_in)
    o = 0
    w[o:o + W1.size] = rng.standard_normal(W1.size) * np.sqrt(2 / n_in); o += W1.size
    w[o:o + 64] = 0.1 * rng.standard_normal(64); o += 64
    w[o:o + W2.size] = rng.standard_normal(W2.size) * np.sqrt(2 / 64); o += W2.size
    w[o:o + 64] = 0.1 * rng.standard_normal(64); o += 64
    w[o:o + 64] = rng.standard_normal(64) / 8; o += 64
    w[o] = 0.05
    return w


@pytest.mark.parametrize("dim,sizes,width", [(2, (24, 12), 20 * 2 + 20), (3, (16, 8), 70 * 3 + 70)])
def test_forward_and_gradient_match_oracle(gpu, dim, sizes, width):
    """2D m = 20 (60 inputs) and the 3D cube's m = 70 (280 inputs)."""
    P = problem(gpu, dim, sizes)
    n_in = P.mlp_inputs()
    assert n_in == width
    w = he_weights(n_in, 1)
    P.set_mlp_weights(w)
    np.testing.assert_array_equal(P.mlp_weights(), w)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((300, n_in)).astype(np.float32)
    t = rng.standard_normal(300).astype(np.float32)
    y = P.mlp_forward(X)
    y_ref, _ = mlp.forward(X, w)
    assert np.abs(y - y_ref).max() <= 1e-5 * max(1.0, np.abs(y_ref).max())
    loss, g = P.mlp_gradient(X, t)
    loss_ref, g_ref = mlp.loss_grad(X, t.astype(np.float64), w.astype(np.float64))
    assert abs(loss - loss_ref) <= 1e-5 * loss_ref
    assert np.abs(g - g_ref).max() <= 1e-4 * np.abs(g_ref).max()


def test_training_features_match_oracle(gpu):
    """mgm_mlp_features (the builder the trainer uses) on the 3D m = 70 cloud:
    offsets / h, residuals / sigma and the target e a_ii / sigma of every
    sampled node equal oracle.mlp.node_features up to FP32 rounding, and a
    node whose stencil residual vanishes gets zero residual inputs and target."""
    P = problem(gpu, 3, (16, 8))
    geo = P.geometry(0)
    N = geo["x"].size
    cl, Lp = P.cloud(0), P.matrix(0, "lap_Poison")
    m = cl.shape[1]
    rng = np.random.default_rng(8)
    res = rng.standard_normal(N)
    err = rng.standard_normal(N)
    interior = np.flatnonzero(geo["btag"] == 0)
    nodes = np.concatenate([rng.choice(interior, 400, replace=False), np.flatnonzero(geo["btag"] != 0)[:8]])
    quiet = nodes[0]
    res[cl[quiet]] = 0.0
    X, t = P.mlp_features(nodes, res, err)
    h = P.info(0)["d_avg"]
    Xw = np.zeros(X.shape)
    tw = np.zeros(t.shape)
    for b, i in enumerate(nodes):
        f, sigma = mlp.node_features(geo["x"], geo["y"], geo["z"], 3, cl, h, res, i)
        Xw[b] = f
        tw[b] = err[i] * Lp[i, 0] / sigma if sigma > 0 else 0.0
    assert not X[0, m * 3:].any() and t[0] == 0
    np.testing.assert_allclose(X[:, :m * 3], Xw[:, :m * 3], rtol=0, atol=2e-7 * np.abs(Xw[:, :m * 3]).max())
    np.testing.assert_allclose(X[:, m * 3:], Xw[:, m * 3:], rtol=3e-7, atol=1e-30)
    np.testing.assert_allclose(t, tw, rtol=3e-7, atol=1e-30)
    print(f"  {nodes.size} rows of {X.shape[1]} inputs, max |t| {np.abs(t).max():.3g}", flush=True)


@pytest.mark.parametrize("epochs,batch", [(2, 512), (1, 128)])
def test_adam_fit_matches_oracle(gpu, oracle, epochs, batch):
    """mgm_mlp_fit (FP32 Adam over shuffled batches, as the trainer runs it)
    against oracle.mlp.adam_fit (float64) at the 280-input width.  Each step
    moves a weight by about lr times the sign of its averaged gradient, so
    weights whose gradient is near zero at some step (below 1e-4 of the
    largest, about 100x the FP32 gradient error) may move differently: every
    other weight -- over 95% of them -- must move as the oracle's to 0.1% of
    lr, and those few by at most 2 lr per step."""
    P = problem(gpu, 3, (16, 8))
    n_in = P.mlp_inputs()
    w0 = he_weights(n_in, 5)
    P.set_mlp_weights(w0)
    rng = np.random.default_rng(9)
    n = 512
    X = rng.standard_normal((n, n_in)).astype(np.float32)
    t = rng.standard_normal(n).astype(np.float32)
    lr, seed = 1e-4, 17
    l0, l1 = P.mlp_fit(X, t, epochs=epochs, batch=batch, lr=lr, seed=seed)
    w = P.mlp_weights().astype(np.float64)
    draws = oracle.splitmix64(seed, epochs * (n - 1))
    w_ref, grads = mlp.adam_fit(X, t.astype(np.float64), w0.astype(np.float64), epochs, batch, lr, draws)
    steps = len(grads)
    assert steps == epochs * (n // batch)
    robust = np.ones(w0.size, bool)
    for g in grads:
        robust &= np.abs(g) > 1e-4 * np.abs(g).max()
    d, d_ref = w - w0, w_ref - w0
    print(f"  {steps} steps, {robust.mean():.4f} of weights with a clear gradient sign, "
          f"loss {l0:.5f} -> {l1:.5f}", flush=True)
    assert robust.mean() > 0.95
    assert np.abs(d - d_ref)[robust].max() <= 1e-3 * lr
    assert np.abs(d[~robust]).max(initial=0) <= 2 * lr * steps
    assert abs(l0 - mlp.loss(X, t, w0.astype(np.float64))) <= 1e-5 * l0
    assert abs(l1 - mlp.loss(X, t, w)) <= 1e-5 * l1
    assert l1 < l0


def windowed(monkeypatch, block_rows):
    """Lay levels above 1,000 rows out in Morton blocks of block_rows rows, so
    their full passes and learned corrections take the LDS-window form."""
    monkeypatch.setenv("MGM_BLOCKED_ROWS", "1000")
    monkeypatch.setenv("MGM_BLOCK_ROWS", str(block_rows))


@pytest.mark.parametrize("dim,sizes,blocks", [(2, (40, 20, 10), 0), (3, (16, 8), 0), (3, (16, 8), 512)])
def test_learned_sweep_matches_oracle(gpu, monkeypatch, dim, sizes, blocks):
    print(f"[gpu learned sweep] dim {dim} sizes {sizes} blocks of {blocks or 'all'} rows", flush=True)
    if blocks:
        windowed(monkeypatch, blocks)
    P = problem(gpu, dim, sizes)
    assert (P.info(0)["window_blocks"] > 0) == bool(blocks)
    n_in = P.mlp_inputs()
    w = he_weights(n_in, 3)
    P.set_mlp_weights(w)
    P.use_learned_smoother(True)
    P.set_smoother(1.0, 1)                       # relax = the learned sweep alone
    geo = P.geometry(0)
    N = geo["x"].size
    rng = np.random.default_rng(4)
    interior = geo["btag"] == 0
    p0 = rng.standard_normal(N)
    s = np.where(interior, rng.standard_normal(N), 0.0)
    P.set_field(0, "p", p0)
    P.set_field(0, "source", s)
    P.relax(0)
    dp = P.field(0, "p") - p0
    cl, Lp = P.cloud(0), P.matrix(0, "lap_Poison")
    res = np.where(interior, s - (Lp * p0[cl]).sum(1), 0.0)
    h = P.info(0)["d_avg"]
    want = np.zeros(N)
    for i in np.flatnonzero(interior):
        want[i] = mlp.correction(geo["x"], geo["y"], geo["z"], dim, cl, h, Lp, res, w, i)
    assert np.all(dp[~interior] == 0)
    err = np.abs(dp - want).max() / np.abs(want).max()
    print(f"  max |dp - oracle| / max |oracle| = {err:.2e}", flush=True)
    assert err < 1e-3


@pytest.mark.parametrize("block_rows", [512, 1024])
def test_learned_window_form_bitwise(gpu, monkeypatch, block_rows):
    """The LDS-window form of the correction (levels in Morton blocks) equals
    the gathering form bit for bit: one sweep from a random p, and whole
    learned solves in FP64 and with the FP32 inner cycle."""
    print(f"[gpu learned window] 3D 24/12, blocks of {block_rows} rows", flush=True)
    windowed(monkeypatch, block_rows)
    P = problem(gpu, 3, (24, 12))
    assert P.info(0)["window_blocks"] > 1
    P.set_mlp_weights(he_weights(P.mlp_inputs(), 3))
    P.use_learned_smoother(True)
    geo = P.geometry(0)
    N = geo["x"].size
    rng = np.random.default_rng(12)
    p0 = rng.standard_normal(N)
    s = np.where(geo["btag"] == 0, rng.standard_normal(N), 0.0)
    out = {}
    for form in ("1", "0"):
        monkeypatch.setenv("MGM_LEARNED_WINDOW", form)
        P.set_smoother(1.0, 1)
        P.set_field(0, "p", p0)
        P.set_field(0, "source", s)
        P.relax(0)
        sweep = P.field(0, "p")
        P.set_smoother(1.0, 3)
        solves = []
        for prec in (0, 1):
            P.zero_fields()
            P.set_field(0, "source", s)
            r = P.solve(tol=1e-8, max_cycles=100, precision=prec)
            solves.append((r["history"], P.field(0, "p")))
        out[form] = (sweep, solves)
    assert not np.array_equal(out["1"][0], p0)
    np.testing.assert_array_equal(out["1"][0], out["0"][0])
    for (h1, p1), (h0, p0_) in zip(out["1"][1], out["0"][1]):
        np.testing.assert_array_equal(h1, h0)
        np.testing.assert_array_equal(p1, p0_)
    print(f"  cycles FP64 {len(out['1'][1][0][0]) - 1}, FP32 inner {len(out['1'][1][1][0]) - 1}", flush=True)


def test_learned_level_mask(gpu):
    """set_learned_levels: a level left out relaxes exactly as plain SOR, a
    level kept relaxes exactly as with no mask; an empty mask restores all."""
    print("[gpu learned level mask] dim 2 sizes (40, 20, 10)", flush=True)
    P = problem(gpu, 2, (40, 20, 10))
    P.set_mlp_weights(he_weights(P.mlp_inputs(), 3))
    P.set_smoother(1.0, 2)

    def relax(l, learned, mask):
        P.use_learned_smoother(learned)
        P.set_learned_levels(mask)
        N = P.geometry(l)["x"].size
        r = np.random.default_rng(7 + l)
        P.set_field(l, "p", r.standard_normal(N))
        P.set_field(l, "source", r.standard_normal(N))
        P.relax(l)
        return P.field(l, "p")

    for l in range(P.L):
        plain = relax(l, False, None)
        full = relax(l, True, None)
        assert l > 0 or not np.array_equal(plain, full)   # level 0's stencils fit the network
        off = [1] * P.L
        off[l] = 0
        on = [0] * P.L
        on[l] = 1
        assert np.array_equal(relax(l, True, off), plain)
        assert np.array_equal(relax(l, True, on), full)
        assert np.array_equal(relax(l, True, None), full)
    with pytest.raises(Exception):
        P.set_learned_levels([1] * (P.L + 1))


def test_training_and_learned_solve(gpu):
    from multigrid_meshless_amd import cases
    P = problem(gpu, 2, (64, 32, 16))
    l0, l1 = P.train_smoother(snapshots=6, samples_per_snapshot=4096, epochs=20, batch=2048, lr=2e-3, seed=5)
    print(f"[gpu learned training] loss {l0:.4f} -> {l1:.4f}", flush=True)
    assert l1 < 0.5 * l0
    g = P.geometry(0)
    f = cases.fourier_rhs(g["x"], g["y"], 0 * g["x"])
    f[g["btag"] != 0] = 0.0
    out = {}
    for learned in (False, True):
        P.use_learned_smoother(learned)
        P.zero_fields()
        P.set_field(0, "source", f)
        r = P.solve(tol=1e-8, max_cycles=300)
        out[learned] = r
        assert r["history"][-1] <= 1e-8
    fb = P.learned_fell_back()
    print(f"  plain: {out[False]['cycles']} cycles {out[False]['seconds']:.4f}s; learned: "
          f"{out[True]['cycles']} cycles {out[True]['seconds']:.4f}s; fell back: {fb}", flush=True)
    if not fb:
        # every learned cycle reduced the residual at least as much as the plain one
        h = out[True]["history"]
        rho_plain = h[1] / h[0]
        assert np.all(h[2:] / h[1:-1] <= rho_plain * (1 + 1e-12))
    # inside the reduced-precision V-cycles as well: the FP64 residual reaches the tolerance
    P.use_learned_smoother(True)
    for prec in (1, 2):
        P.zero_fields()
        P.set_field(0, "source", f)
        r = P.solve(tol=1e-10, max_cycles=300, precision=prec)
        print(f"  learned, precision {prec}: {r['cycles']} cycles {r['seconds']:.4f}s "
              f"fell back {P.learned_fell_back()}", flush=True)
        assert r["history"][-1] <= 1e-10


def annulus_problem(mgm, n, levels, seed0):
    """The 2D annulus r in [0.5, 1] (cases.annulus), `levels` clouds halving in
    spacing; PHS r^3 + cubic polynomials, multiplier 2 (m = 20); Dirichlet rows
    on both circles."""
    from multigrid_meshless_amd import cases
    sizes = [((n - 1) >> k) + 1 for k in range(levels)]
    pts = [cases.annulus(s, seed0 + k) for k, s in enumerate(sizes)]
    cfg = mgm.config(dimension=2, poly_degree=3, coarse_poly_degree=3, phs_degree=3, cloud_size_multiplier=2.0,
                     omega=1.0, num_relax=3, transfer_rows=mgm.ROWS_BY_TAG, sweep_order=mgm.ORDER_COLOUR)
    P = mgm.Problem(cfg, pts).setup()
    P.dirichlet_rows()
    return P


def annulus_rhs(g, seed=2, modes=16):
    rng = np.random.default_rng(seed)
    f = np.zeros_like(g["x"])
    for _ in range(modes):
        k = rng.integers(1, 5, 2)
        f += rng.standard_normal() * np.sin(np.pi * (k[0] * g["x"] + k[1] * g["y"]))
    f[g["btag"] != 0] = 0.0
    return f


def test_learned_smoother_annulus(gpu):
    """Trained on a coarser annulus hierarchy, used on a finer one: the learned
    solve reaches the tolerance, its field agrees with the plain solve's, and
    it needs fewer V-cycles (or reports that it fell back to SOR)."""
    print("[gpu learned] annulus 181/91/46/23, network trained on 91/46/23", flush=True)
    P = annulus_problem(gpu, 181, 4, 41)
    f = annulus_rhs(P.geometry(0))
    P.zero_fields()
    P.set_field(0, "source", f)
    plain = P.solve(tol=1e-8, max_cycles=300)
    p_plain = P.field(0, "p")
    assert plain["history"][-1] <= 1e-8
    T = annulus_problem(gpu, 91, 3, 51)
    assert T.mlp_inputs() == P.mlp_inputs() == 20 * 2 + 20
    l0, l1 = T.train_smoother(snapshots=8, samples_per_snapshot=4096, epochs=30, batch=2048, lr=1e-3, seed=7)
    assert l1 < l0
    P.set_mlp_weights(T.mlp_weights())
    P.use_learned_smoother(True)
    P.zero_fields()
    P.set_field(0, "source", f)
    r = P.solve(tol=1e-8, max_cycles=300)
    fb = P.learned_fell_back()
    diff = np.abs(P.field(0, "p") - p_plain).max() / np.abs(p_plain).max()
    print(f"  plain {plain['cycles']} cycles {plain['seconds']:.4f}s; learned {r['cycles']} cycles "
          f"{r['seconds']:.4f}s, fell back {fb}, field diff {diff:.2e}, loss {l0:.3f} -> {l1:.3f}", flush=True)
    assert r["history"][-1] <= 1e-8
    assert diff <= 1e-6
    assert fb or r["cycles"] < plain["cycles"]
