"""Reduced-precision solves (mgm_solve precision 1: FP32 inner V-cycle; 2: plus BF16
weights on levels >= 1) inside FP64 defect correction.

There is no reference counterpart (the reference solves in FP64 only).  The bar:
the FP64 residual of the defect-corrected iterate reaches the tolerance, and the
converged field equals the FP64 multigrid solution -- itself bit-identical to the
oracle's V-cycle -- to 1e-8 relative (max norm) when both are driven to 1e-12.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def problem(mgm, dim, sizes):
    from multigrid_meshless_amd import cases
    pts = [mgm.Points.generate(dim, n, 1 + k, 0.2) for k, n in enumerate(sizes)]
    cfg = mgm.config(dimension=dim, poly_degree=3, coarse_poly_degree=3, phs_degree=3,
                     cloud_size_multiplier=3.5 if dim == 3 else 2.0, omega=1.0, num_relax=3,
                     transfer_rows=mgm.ROWS_BY_TAG, sweep_order=mgm.ORDER_COLOUR)
    P = mgm.Problem(cfg, pts).setup()
    P.dirichlet_rows()
    g = P.geometry(0)
    f = cases.fourier_rhs(g["x"], g["y"], g["z"] if dim == 3 else 0 * g["x"])
    f[g["btag"] != 0] = 0.0
    return P, f


@pytest.mark.parametrize("dim,sizes", [(2, (64, 32, 16)), (3, (28, 14, 8))])
def test_mixed_precision_matches_fp64(gpu, dim, sizes):
    print(f"[gpu mixed] dim {dim} sizes {sizes}", flush=True)
    P, f = problem(gpu, dim, sizes)
    out = {}
    for prec in (0, 1, 2):
        P.zero_fields()
        P.set_field(0, "source", f)
        r = P.solve(tol=1e-12, max_cycles=400, precision=prec)
        out[prec] = (r, P.field(0, "p"))
        print(f"  precision {prec}: {r['cycles']} cycles, final {r['history'][-1]:.2e}", flush=True)
        assert r["history"][-1] <= 1e-12, (prec, r["history"][-1])
    p64 = out[0][1]
    scale = np.abs(p64).max()
    for prec in (1, 2):
        err = np.abs(out[prec][1] - p64).max() / scale
        print(f"  precision {prec}: max |p - p64| / max |p64| = {err:.2e}", flush=True)
        assert err < 1e-8, (prec, err)
    # the FP32 correction contracts about as fast as the FP64 cycle
    assert out[1][0]["cycles"] <= out[0][0]["cycles"] + 5


def test_mixed_precision_rejects_unsupported(gpu):
    P, f = problem(gpu, 2, (32, 16))
    P.set_field(0, "source", f)
    with pytest.raises(gpu.MgmError):
        P.solve(tol=1e-8, max_cycles=5, variant=gpu.VAR_T, precision=1)
    with pytest.raises(gpu.MgmError):
        P.solve(tol=1e-8, max_cycles=5, precision=3)
