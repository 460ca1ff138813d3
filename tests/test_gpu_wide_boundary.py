"""Wide boundary stencils (boundary_cloud_size_multiplier, no reference
counterpart) against the oracle.

With the option every boundary node gets a second stencil of m_b points
(itself, then its m_b - 1 nearest interior nodes: the reference's boundary
rule, hf/kdtree_functions.h:139-149, at a larger width) and its own
PHS + polynomial system (m_b + n unknowns) for the four derivative operators,
solved in the same batched weight build as the m-point systems.  Checked:
  * the wide stencils equal the oracle's kd-tree search at width m_b, bit for
    bit, mapped through the same BFS reordering; their first m points are the
    node's m-point stencil;
  * the wide rows are within 1e-11 of the oracle's pivoted-LU weights on
    those stencils (hf/rbf.h:401-421), with the Gauss-Jordan and the
    matrix-core kernels, up to 125 unknowns;
  * every m-point operator is bit for bit the one built without the option;
  * mgm_spmv_wide equals the oracle's ordered row sums (hf/mat_lib.h:70-89)
    over the m-point rows inside and the wide rows on the boundary.
"""
import numpy as np
import pytest

from conftest import load_golden
from pipeline import case_meta, level_inputs, num_levels, oracle_setup
from test_gpu_parity import gpu_problem

pytestmark = pytest.mark.gpu

# case: boundary multiplier (m_b = floor(mult * n)): gen3d m 40 -> 60 (80
# unknowns); cube3d_ref m 40 -> 105 (125 unknowns); gen2d m 20 -> 30
CASES = {"gen3d": 3.0, "cube3d_ref": 5.25, "gen2d": 3.0}


def _problem(gpu, g, bmult):
    dim, poly, phs, mult, _ = case_meta(g)
    cfg = gpu.config(dimension=dim, poly_degree=poly, coarse_poly_degree=3, phs_degree=phs,
                     cloud_size_multiplier=mult, omega=float(g["case_omega"]), num_relax=3,
                     transfer_rows=gpu.ROWS_REFERENCE, sweep_order=gpu.ORDER_NATURAL,
                     boundary_cloud_size_multiplier=bmult)
    pts = []
    for k in range(num_levels(g)):
        d = level_inputs(g, k)
        pts.append(gpu.Points.from_arrays(d["N"], d["nb"], d["num_elem"], d["d_avg"], d["x"], d["y"],
                                          d["z"], d["nx"], d["ny"], d["nz"], d["btag"], d["ctag"]))
    return gpu.Problem(cfg, pts).setup()


@pytest.mark.parametrize("kernel", ["auto", "mfma"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_wide_boundary_stencils(name, kernel, gpu, oracle, monkeypatch):
    if kernel == "auto":
        monkeypatch.delenv("MGM_LU_KERNEL", raising=False)
    else:
        monkeypatch.setenv("MGM_LU_KERNEL", kernel)
    monkeypatch.delenv("MGM_LU_VARIANT", raising=False)
    g = load_golden(name)
    dim, poly, phs, mult, _ = case_meta(g)
    bmult = CASES[name]
    print(f"[gpu wide boundary] {name} kernel {kernel} boundary multiplier {bmult}", flush=True)
    P = _problem(gpu, g, bmult)
    plain = gpu_problem(gpu, g)
    ref = oracle_setup(oracle, g, weights=False)
    ops = ["Dx", "Dy", "lap"] + (["Dz"] if dim == 3 else [])
    worst = 0.0
    for k in range(num_levels(g)):
        info = P.info(k)
        m, n, N = info["m"], info["n"], info["N"]
        nbw, mb = P.boundary_info(k)
        assert mb == int(np.floor(bmult * n + 1e-9)) and mb > m
        # m-point operators untouched
        np.testing.assert_array_equal(P.cloud(k), plain.cloud(k))
        for key in ops + ["lap_Poison"] + (["restr"] if k > 0 else []) + (["prol"] if k < P.L - 1 else []):
            np.testing.assert_array_equal(P.matrix(k, key), plain.matrix(k, key), err_msg=f"{name} L{k} {key}")
        # wide stencils: the oracle's search at width mb, through the BFS order
        d = level_inputs(g, k)
        wide_in = oracle.find_cloud_index(d["x"], d["y"], d["z"], d["btag"], d["ctag"], d["nb"], d["d_avg"], mb)
        q1 = ref[k]["q1"]
        q2 = np.empty_like(q1)
        q2[q1] = np.arange(len(q1), dtype=q1.dtype)
        rows = P.boundary_rows(k, "Dx")
        nodes, cloud = rows["nodes"], rows["cloud"]
        geo = P.geometry(k)
        ctag_ref = d["ctag"]   # the reference tests the corner tag by the new index (file-order array)
        expect_nodes = np.array([r for r in range(N) if geo["btag"][r] and not ctag_ref[r]], np.int32)
        np.testing.assert_array_equal(nodes, expect_nodes)
        np.testing.assert_array_equal(cloud, q2[wide_in[q1[nodes]]], err_msg=f"{name} L{k} wide stencils")
        np.testing.assert_array_equal(cloud[:, :m], P.cloud(k)[nodes])
        # wide weights against the oracle's LU on the same stencils
        lv = dict(x=geo["x"], y=geo["y"], z=geo["z"], cloud=cloud)
        Wo = dict(zip(["Dx", "Dy", "Dz", "lap"], oracle.weights(lv, dim, phs, info["poly_degree"], method=1)))
        for key in ops:
            G = P.boundary_rows(k, key)["W"]
            sc = np.abs(Wo[key]).max(axis=1)
            e = (np.abs(G - Wo[key]).max(axis=1) / sc).max()
            worst = max(worst, e)
            assert e <= 1e-11, (name, kernel, k, key, e)
        # mgm_spmv_wide: m-point rows inside, wide rows on the boundary
        f = np.random.default_rng(5 + k).standard_normal(N)
        P.set_field(k, "p", f)
        for key in ops:
            P.spmv_wide(k, key, "p", "dpdx")
            expect = oracle.spmv(P.matrix(k, key), P.cloud(k), f)
            expect[nodes] = oracle.spmv(P.boundary_rows(k, key)["W"], cloud, f)
            np.testing.assert_array_equal(P.field(k, "dpdx"), expect, err_msg=f"{name} L{k} spmv_wide {key}")
    print(f"  m_b {mb}: worst wide-row difference from the oracle's LU {worst:.2e}", flush=True)
