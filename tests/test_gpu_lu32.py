"""The matrix-core weight kernel (FP32 LU on MFMA + FP64 iterative refinement)
against the oracle.

MGM_LU_KERNEL=mfma forces it for every build of the test's problems.  Bars as
for the FP64 Gauss-Jordan kernel: weights (Dx, Dy, Dz, Laplacian),
lap_Poison and transfer rows within 1e-11 row-relative of the oracle's pivoted
LU (hf/rbf.h:401-421, hf/rbf.h:382-399, hf/mesh_functions.h:533-670), on
  * every golden fixture case (2D and 3D, PHS r^3 and r^5, polynomial degrees
    2-4: systems of 30 to 60 unknowns, 2 to 4 tiles of 16);
  * the 1M cube's stencils (m = 70, 90 unknowns, 6 tiles) on a 24^3 hierarchy;
  * the lid-driven cavity's stencils (PHS r^5 + degree 4, m = 30, 45
    unknowns), the ones that need the most refinement steps.
"""
import numpy as np
import pytest

from conftest import golden_cases, load_golden
from pipeline import case_meta, num_levels, oracle_setup
from test_gpu_parity import gpu_problem

pytestmark = pytest.mark.gpu


@pytest.fixture()
def mfma(monkeypatch):
    monkeypatch.setenv("MGM_LU_KERNEL", "mfma")
    monkeypatch.delenv("MGM_LU_VARIANT", raising=False)


def _row_rel(G, W):
    sc = np.abs(W).max(axis=1)
    zero = sc == 0
    assert not np.abs(G[zero]).any(), "rows the oracle leaves empty must stay empty"
    return np.abs(G - W).max(axis=1)[~zero] / sc[~zero]


@pytest.mark.parametrize("name", golden_cases())
def test_fixture_cases(name, gpu, oracle, mfma):
    g = load_golden(name)
    dim, poly, phs, mult, _ = case_meta(g)
    print(f"[gpu lu32] fixture {name}", flush=True)
    P = gpu_problem(gpu, g)
    mpn = P.info(0)["m"] + P.info(0)["n"]
    assert P.weight_kernel_choice(0) == (-(-mpn // 16), 0)
    ref = oracle_setup(oracle, g, weights=True, row_mode=0, method=1)
    worst = 0.0
    L = num_levels(g)
    for k in range(L):
        rows = {key: ref[k][key] for key in ("Dx", "Dy", "lap", "lap_Poison")}
        if dim == 3:
            rows["Dz"] = ref[k]["Dz"]
        if k > 0:
            rows["restr"] = ref[k]["restr"]
        if k < L - 1:
            rows["prol"] = ref[k]["prol"]
        for key, W in rows.items():
            e = _row_rel(P.matrix(k, key), W).max()
            worst = max(worst, e)
            assert e <= 1e-11, (name, k, key, e)
    print(f"  worst row-relative difference from the oracle's LU: {worst:.2e}", flush=True)


@pytest.mark.parametrize("kind", ["cube_m70", "cavity_m30"])
def test_benchmark_stencils(kind, gpu, oracle, mfma):
    if kind == "cube_m70":
        dim, sizes, poly, phs, mult = 3, (24, 12, 7), 3, 3, 3.5
    else:
        dim, sizes, poly, phs, mult = 2, (65, 33, 17), 4, 5, 2.0
    print(f"[gpu lu32] {kind}: {dim}D {sizes}, PHS r^{phs}, degree {poly}, multiplier {mult}", flush=True)
    cfg = gpu.config(dimension=dim, poly_degree=poly, coarse_poly_degree=poly, phs_degree=phs,
                     cloud_size_multiplier=mult, omega=1.0, num_relax=3, transfer_rows=gpu.ROWS_BY_TAG,
                     sweep_order=gpu.ORDER_COLOUR)
    P = gpu.Problem(cfg, [gpu.Points.generate(dim, n, 1 + k, 0.2) for k, n in enumerate(sizes)]).setup()
    assert P.weight_kernel_choice(0)[1] == 0
    ref = _oracle_rows(oracle, P, dim, phs)
    worst = 0.0
    for k in range(P.L):
        for key, W in ref[k].items():
            e = _row_rel(P.matrix(k, key), W).max()
            worst = max(worst, e)
            assert e <= 1e-11, (kind, k, key, e)
    print(f"  worst row-relative difference from the oracle's LU: {worst:.2e}", flush=True)


def _oracle_rows(oracle, P, dim, phs):
    L = P.L
    lv = []
    for k in range(L):
        info = P.info(k)
        lv.append(dict(P.geometry(k), cloud=P.cloud(k), nb=info["nb"], poly=info["poly_degree"],
                       rp=P.points(k, "restriction"), pp=P.points(k, "prolongation"),
                       ctag=np.zeros(info["N"], np.uint8)))
    out = []
    for k, d in enumerate(lv):
        Dx, Dy, Dz, lap = oracle.weights(d, dim, phs, d["poly"], method=1)
        rows = dict(Dx=Dx, Dy=Dy, lap=lap, lap_Poison=oracle.lap_poisson(d, dim, Dx, Dy, Dz, lap))
        if dim == 3:
            rows["Dz"] = Dz
        if k > 0:
            rows["restr"] = oracle.restriction(lv[k - 1], d, dim, phs, lv[k - 1]["poly"], row_mode=1, method=1)
        if k < L - 1:
            rows["prol"] = oracle.prolongation(d, lv[k + 1], dim, phs, lv[k + 1]["poly"], row_mode=1, method=1)
        out.append(rows)
    return out


@pytest.mark.parametrize("kernel", ["gj", "mfma"])
def test_one_launch_over_levels_of_uneven_size(kernel, gpu, oracle, monkeypatch):
    """Setup solves the systems of all levels in one launch per tile class
    (16 t >= unknowns + 4 for the Gauss-Jordan grid of t tiles); here the
    levels' systems differ in size (PHS r^5 with degree-4 polynomials on level
    0: 45 unknowns, class 4; degree 3 below: 30 unknowns, levels 1 and 2
    together in class 3).  Each level's rows equal those of a build of that
    level alone, bit for bit, and the oracle's LU to 1e-11."""
    monkeypatch.setenv("MGM_LU_KERNEL", kernel)
    monkeypatch.delenv("MGM_LU_VARIANT", raising=False)
    dim, sizes, phs = 2, (65, 33, 17), 5
    cfg = gpu.config(dimension=dim, poly_degree=4, coarse_poly_degree=3, phs_degree=phs,
                     cloud_size_multiplier=2.0, omega=1.0, num_relax=3, transfer_rows=gpu.ROWS_BY_TAG,
                     sweep_order=gpu.ORDER_COLOUR)
    P = gpu.Problem(cfg, [gpu.Points.generate(dim, n, 1 + k, 0.2) for k, n in enumerate(sizes)]).setup()
    assert [P.info(k)["m"] + P.info(k)["n"] for k in range(P.L)] == [45, 30, 30]
    keys = ["Dx", "Dy", "lap", "lap_Poison", "restr", "prol"]
    batched = [{key: P.matrix(k, key).copy() for key in keys
                if not (key == "restr" and k == 0) and not (key == "prol" and k == P.L - 1)} for k in range(P.L)]
    ref = _oracle_rows(oracle, P, dim, phs)
    for k in range(P.L):
        for key, W in ref[k].items():
            e = _row_rel(batched[k][key], W).max()
            assert e <= 1e-11, (kernel, k, key, e)
    # level by level: transfer rows of level k come from the systems of level
    # k - 1 (restriction) and k + 1 (prolongation), so rebuild all first
    for k in range(P.L):
        P.rebuild_weights(k)
    for k in range(P.L):
        for key, G in batched[k].items():
            np.testing.assert_array_equal(P.matrix(k, key), G, err_msg=f"{kernel} L{k} {key}")


@pytest.mark.parametrize("kernel", ["gj", "mfma"])
def test_interior_and_boundary_systems_in_one_launch(kernel, gpu, oracle, monkeypatch):
    """Interior systems (m = 70: 90 unknowns) and wide boundary systems
    (m_b = 72: 92 unknowns) of both levels share a tile class (6 tiles of 16),
    so setup solves all of them in one launch, each at its own size.  Every
    row within 1e-11 of the oracle's LU on the same stencils, and bit for bit
    the rows of per-level rebuilds."""
    monkeypatch.setenv("MGM_LU_KERNEL", kernel)
    monkeypatch.delenv("MGM_LU_VARIANT", raising=False)
    dim, sizes, phs = 3, (16, 8), 3
    print(f"[gpu lu32] interior 90 + boundary 92 unknowns in one launch, kernel {kernel}", flush=True)
    cfg = gpu.config(dimension=dim, poly_degree=3, coarse_poly_degree=3, phs_degree=phs,
                     cloud_size_multiplier=3.5, omega=1.0, num_relax=3, transfer_rows=gpu.ROWS_BY_TAG,
                     sweep_order=gpu.ORDER_COLOUR, boundary_cloud_size_multiplier=3.6)
    P = gpu.Problem(cfg, [gpu.Points.generate(dim, n, 1 + k, 0.2) for k, n in enumerate(sizes)]).setup()
    ops = ["Dx", "Dy", "Dz", "lap"]
    for k in range(P.L):
        assert P.info(k)["m"] + P.info(k)["n"] == 90 and P.boundary_info(k)[1] == 72
    keys = ops + ["lap_Poison", "restr", "prol"]
    batched = [{key: P.matrix(k, key).copy() for key in keys
                if not (key == "restr" and k == 0) and not (key == "prol" and k == P.L - 1)} for k in range(P.L)]
    wide = [{key: P.boundary_rows(k, key) for key in ops} for k in range(P.L)]
    ref = _oracle_rows(oracle, P, dim, phs)
    worst = 0.0
    for k in range(P.L):
        for key, W in ref[k].items():
            e = _row_rel(batched[k][key], W).max()
            worst = max(worst, e)
            assert e <= 1e-11, (kernel, k, key, e)
        geo = P.geometry(k)
        lv = dict(x=geo["x"], y=geo["y"], z=geo["z"], cloud=wide[k]["Dx"]["cloud"])
        Wo = dict(zip(ops, oracle.weights(lv, dim, phs, P.info(k)["poly_degree"], method=1)))
        for key in ops:
            e = _row_rel(wide[k][key]["W"], Wo[key]).max()
            worst = max(worst, e)
            assert e <= 1e-11, (kernel, k, "wide " + key, e)
    for k in range(P.L):
        P.rebuild_weights(k)
    for k in range(P.L):
        for key, G in batched[k].items():
            np.testing.assert_array_equal(P.matrix(k, key), G, err_msg=f"{kernel} L{k} {key}")
        for key in ops:
            np.testing.assert_array_equal(P.boundary_rows(k, key)["W"], wide[k][key]["W"],
                                          err_msg=f"{kernel} L{k} wide {key}")
    print(f"  worst row-relative difference from the oracle's LU: {worst:.2e}", flush=True)
