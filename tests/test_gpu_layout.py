"""Gather-sorted slabs give the plain slabs' results bit for bit.

On levels above MGM_GATHER_SORT_ROWS rows (262,144 by default) the streaming
passes -- SpMV, the colour steps of SOR, residuals, prolongation -- read their
slabs with each slice's entries stored in column order, every product written
to the LDS slot it has in the plain layout.  The row sums, and so every
result, must not change.  Here the threshold is set to 0, so every level of a
small hierarchy is stored both ways (MGM_GATHER_SORT=0 builds the plain one),
and the single-workgroup small-level relax is switched off so that the colour
steps run on every level.
"""
import numpy as np
import pytest

from multigrid_meshless_amd import cases

pytestmark = pytest.mark.gpu

CASES = [(2, (65, 33, 17)), (3, (21, 11, 7))]
FIELDS = ("p", "pprime", "res", "source")


def build(mgm, monkeypatch, sort, dim, sizes, seed=3):
    monkeypatch.setenv("MGM_GATHER_SORT", "1" if sort else "0")
    monkeypatch.setenv("MGM_GATHER_SORT_ROWS", "0")
    cfg = mgm.config(dimension=dim, poly_degree=3, coarse_poly_degree=3, phs_degree=3,
                     cloud_size_multiplier=3.5 if dim == 3 else 2.0, omega=1.0, num_relax=2,
                     transfer_rows=mgm.ROWS_BY_TAG, sweep_order=mgm.ORDER_COLOUR)
    pts = [mgm.Points.generate(dim, n, seed + k, 0.2) for k, n in enumerate(sizes)]
    return mgm.Problem(cfg, pts).setup()


def make_pair(gpu, monkeypatch, dim, sizes):
    monkeypatch.setenv("MGM_SMALL_ROWS", "0")
    print(f"[gpu layout] {dim}D levels {sizes}: plain vs gather-sorted", flush=True)
    A = build(gpu, monkeypatch, False, dim, sizes)
    B = build(gpu, monkeypatch, True, dim, sizes)
    return gpu, A, B


@pytest.fixture(params=CASES, ids=["2d", "3d"])
def pair(request, gpu, monkeypatch):
    return make_pair(gpu, monkeypatch, *request.param)


@pytest.fixture
def pair_2d(gpu, monkeypatch):
    """The cavity time steps are two-dimensional: the 2D hierarchy only."""
    return make_pair(gpu, monkeypatch, *CASES[0])


def same_fields(A, B, names=FIELDS):
    for lev in range(A.L):
        for k in names:
            np.testing.assert_array_equal(A.field(lev, k), B.field(lev, k), err_msg=f"level {lev} {k}")


def test_spmv_bitwise(pair):
    mgm, A, B = pair
    rng = np.random.default_rng(5)
    for lev in range(A.L):
        x = rng.standard_normal(A.info(lev)["N"])
        for name in ("lap_Poison", "lap", "Dx"):
            for P in (A, B):
                P.set_field(lev, "p", x)
                P.spmv(lev, name, "p", "dpdn")
            np.testing.assert_array_equal(A.field(lev, "dpdn"), B.field(lev, "dpdn"), err_msg=f"{lev} {name}")
            assert np.abs(A.field(lev, "dpdn")).max() > 0


@pytest.mark.parametrize("variant", ["H", "T", "F"])
def test_vcycles_bitwise(pair, variant):
    mgm, A, B = pair
    var = getattr(mgm, "VAR_" + variant)
    f = np.random.default_rng(7).standard_normal(A.info(0)["N"])
    for P in (A, B):
        P.dirichlet_rows()
        P.zero_fields()
        P.set_field(0, "source", f)
        P.vcycle(var, 3)
    same_fields(A, B)
    assert np.abs(A.field(0, "pprime" if variant == "T" else "p")).max() > 0


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_solve_bitwise(pair, precision):
    mgm, A, B = pair
    f = np.random.default_rng(9).standard_normal(A.info(0)["N"])
    runs = []
    for P in (A, B):
        P.dirichlet_rows()
        P.zero_fields()
        P.set_field(0, "source", f)
        runs.append(P.solve(1e-8, 60, precision=precision))
    assert runs[0]["cycles"] == runs[1]["cycles"]
    np.testing.assert_array_equal(runs[0]["history"], runs[1]["history"])
    np.testing.assert_array_equal(A.field(0, "p"), B.field(0, "p"))


def test_time_steps_bitwise(pair_2d):
    mgm, A, B = pair_2d
    for scheme in ("fs", "timple"):
        errs = []
        for P in (A, B):
            P.zero_fields()
            cases.lid_driven_cavity(P, timple=scheme == "timple")
            P.ns_set_params(**cases.ns_params(100.0, P.info(0)["d_avg"], 0.5, num_vcycles=3))
            errs.append([P.fs_step() if scheme == "fs" else P.timple_step() for _ in range(2)])
        np.testing.assert_array_equal(errs[0], errs[1])
        same_fields(A, B, ("p", "pprime", "u", "v", "u_new", "v_new", "p_old"))
    # the FP32 pressure cycle of a captured march
    for var in (mgm.VAR_F, mgm.VAR_T):
        runs = []
        for P in (A, B):
            P.zero_fields()
            cases.lid_driven_cavity(P, timple=var == mgm.VAR_T)
            P.ns_set_params(**cases.ns_params(100.0, P.info(0)["d_avg"], 0.5, num_vcycles=3))
            runs.append(P.ns_march(var, steps=2, precision=1))
        np.testing.assert_array_equal(runs[0]["errors"], runs[1]["errors"])
        same_fields(A, B, ("p", "pprime", "u", "v", "u_new", "v_new", "p_old"))


def test_default_thresholds_bitwise(gpu, monkeypatch):
    """The layout the 1M benchmark runs: level 0 above 262,144 rows gather-sorted
    at slice height 8, the level below it at slice height 4, plain slabs below;
    the same FP64 solve as with every slab plain."""
    print("[gpu layout] 3D 66^3/33^3/17^3/9^3 at the default thresholds", flush=True)
    monkeypatch.delenv("MGM_GATHER_SORT_ROWS", raising=False)
    sizes = (66, 33, 17, 9)
    runs = []
    for sort in (False, True):
        monkeypatch.setenv("MGM_GATHER_SORT", "1" if sort else "0")
        cfg = gpu.config(dimension=3, poly_degree=3, coarse_poly_degree=3, phs_degree=3,
                         cloud_size_multiplier=3.5, omega=1.0, num_relax=3,
                         transfer_rows=gpu.ROWS_BY_TAG, sweep_order=gpu.ORDER_COLOUR)
        P = gpu.Problem(cfg, [gpu.Points.generate(3, n, 1 + k, 0.2) for k, n in enumerate(sizes)]).setup()
        assert P.info(0)["N"] > 262144
        P.dirichlet_rows()
        f = np.random.default_rng(4).standard_normal(P.info(0)["N"])
        f[P.geometry(0)["btag"] != 0] = 0.0
        P.set_field(0, "source", f)
        r = P.solve(1e-8, 12)
        runs.append((r, P.field(0, "p")))
    np.testing.assert_array_equal(runs[0][0]["history"], runs[1][0]["history"])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
