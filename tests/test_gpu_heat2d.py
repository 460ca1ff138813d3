"""The reference's CPU-runnable case on the GPU: steady 2D heat conduction /
Poisson on the unit square with the manufactured solution
u = sin(pi x) sin(pi y) (lap u = -2 pi^2 u, u = 0 on the boundary), FP64.

Three jittered Cartesian clouds 128^2 / 64^2 / 32^2 (jitter 0.2 h, seeds
0, 1, 2) are written as Gmsh v2 files and read back, the way the reference's
drivers get their levels (mg_heat_conduction.c:26-36); PHS r^3, cubic
polynomials, multiplier 2 (m = 20), Dirichlet identity rows, 3 SOR sweeps,
V-cycles to a relative residual of 1e-10.

With three levels the 32^2 coarsest level gets its 3 sweeps and no exact
solve, so a cycle reduces the residual by only ~1.5% and the solve takes
about 1,600 cycles; the oracle's own V-cycle on these clouds does the same.

Bars:
  * over the first 60 cycles every relative residual equals the oracle's
    V-cycle on the same operators in the same colour order (1e-12 relative:
    the norms sum in different orders), and the field after them is bitwise
    the oracle's;
  * the solve reaches 1e-10;
  * the solution is within 2e-4 of u: the discretisation error, which a
    direct sparse solve of the same level-0 system puts at 8.4e-5 for these
    clouds.
"""
import os
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = (128, 64, 32)


def test_manufactured_poisson_2d(gpu, oracle):
    mgm = gpu
    print("[gpu heat2d] 128^2/64^2/32^2 through Gmsh files, solve to 1e-10", flush=True)
    with tempfile.TemporaryDirectory() as d:
        pts = []
        for k, n in enumerate(SIZES):
            path = os.path.join(d, f"L{k}.msh")
            gen = mgm.Points.generate(2, n, k, 0.2, with_volume_elements=True, gmsh_path=path)
            back = mgm.Points.read_gmsh(path, 2)
            assert back.info()["N"] == gen.info()["N"]
            pts.append(back)
        cfg = mgm.config(dimension=2, poly_degree=3, coarse_poly_degree=3, phs_degree=3,
                         cloud_size_multiplier=2.0, omega=1.0, num_relax=3,
                         transfer_rows=mgm.ROWS_BY_TAG, sweep_order=mgm.ORDER_COLOUR)
        P = mgm.Problem(cfg, pts).setup()
    assert P.info(0)["m"] == 20
    P.dirichlet_rows()
    g = P.geometry(0)
    u = np.sin(np.pi * g["x"]) * np.sin(np.pi * g["y"])
    f = np.where(g["btag"] != 0, 0.0, -2 * np.pi ** 2 * u)
    lv = []
    for k in range(P.L):
        col = P.colours(k)
        dct = dict(cloud=P.cloud(k), Lp=P.matrix(k, "lap_Poison"), btag=P.geometry(k)["btag"],
                   rp=P.points(k, "restriction"), pp=P.points(k, "prolongation"),
                   order=np.lexsort((np.arange(len(col)), col)).astype(np.int32))
        if k > 0:
            dct["R"] = P.matrix(k, "restr")
        if k < P.L - 1:
            dct["P"] = P.matrix(k, "prol")
        lv.append(dct)
    H = oracle.Hierarchy(lv)
    H.lv[0]["s"][:] = f
    fnorm = oracle.l2_norm_gen(lv[0]["btag"], f)
    ohist = [1.0]
    for _ in range(60):
        H.vcycle(oracle.VAR_H, 1.0, 3)
        H.residual(oracle.VAR_H, 0)
        ohist.append(oracle.l2_norm_gen(lv[0]["btag"], H.lv[0]["res"]) / fnorm)
    P.zero_fields()
    P.set_field(0, "source", f)
    r60 = P.solve(tol=0.0, max_cycles=60)
    assert r60["cycles"] == 60
    np.testing.assert_allclose(r60["history"], ohist, rtol=1e-12, atol=0)
    np.testing.assert_array_equal(P.field(0, "p"), H.lv[0]["p"])

    P.zero_fields()
    P.set_field(0, "source", f)
    r = P.solve(tol=1e-10, max_cycles=4000)
    err = np.abs(P.field(0, "p") - u).max()
    print(f"  {r['cycles']} cycles in {r['seconds']:.3f} s, final {r['history'][-1]:.3e}, "
          f"max |p - sin(pi x) sin(pi y)| = {err:.3e}", flush=True)
    assert r["history"][-1] <= 1e-10
    assert err < 2e-4
