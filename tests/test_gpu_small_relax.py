"""The single-workgroup relax of small levels gives the per-colour launches'
results bit for bit.

Levels of at most MGM_SMALL_ROWS device rows (12,000 by default) relax in one
launch: every sweep and colour class, and for the T / F variants the pin and
residual after each sweep, with the level's vectors in LDS.  Its sums run to
a fixed width with +0.0 in the padding slots, which leaves them unchanged.
MGM_SMALL_ROWS=0 switches it off, so each colour class is its own launch.
Here both forms run the same solves and time steps, in FP64 and in the FP32
and BF16-weight inner cycles, and every result must be identical.
"""
import numpy as np
import pytest

from multigrid_meshless_amd import cases

pytestmark = pytest.mark.gpu

CASES = [(2, (65, 33, 17)), (3, (21, 11, 7))]
NS_FIELDS = ("p", "pprime", "u", "v", "u_new", "v_new", "p_old")


def build(mgm, dim, sizes, seed=5):
    cfg = mgm.config(dimension=dim, poly_degree=3, coarse_poly_degree=3, phs_degree=3,
                     cloud_size_multiplier=3.5 if dim == 3 else 2.0, omega=1.0, num_relax=3,
                     transfer_rows=mgm.ROWS_BY_TAG, sweep_order=mgm.ORDER_COLOUR)
    pts = [mgm.Points.generate(dim, n, seed + k, 0.2) for k, n in enumerate(sizes)]
    return mgm.Problem(cfg, pts).setup()


def run_both(mgm, monkeypatch, dim, sizes, f):
    """f(problem) under each form.  Both problems are set up at the default
    threshold, which also picks the slice heights and so the padded device
    order that norms sum in; the relax form is chosen at launch."""
    monkeypatch.delenv("MGM_SMALL_ROWS", raising=False)
    probs = [build(mgm, dim, sizes) for _ in range(2)]
    out = []
    for P, small in zip(probs, (False, True)):
        if small:
            monkeypatch.delenv("MGM_SMALL_ROWS", raising=False)
        else:
            monkeypatch.setenv("MGM_SMALL_ROWS", "0")
        out.append((P, f(P)))
    monkeypatch.delenv("MGM_SMALL_ROWS", raising=False)
    return out


@pytest.mark.parametrize("case", CASES, ids=["2d", "3d"])
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_solve_bitwise(gpu, monkeypatch, case, precision):
    dim, sizes = case
    print(f"[gpu small relax] {dim}D {sizes} solve, precision {precision}", flush=True)

    def solve(P):
        P.dirichlet_rows()
        f = np.random.default_rng(11).standard_normal(P.info(0)["N"])
        f[P.geometry(0)["btag"] != 0] = 0.0
        P.zero_fields()
        P.set_field(0, "source", f)
        return P.solve(1e-9, 40, precision=precision)

    (A, ra), (B, rb) = run_both(gpu, monkeypatch, dim, sizes, solve)
    assert ra["cycles"] == rb["cycles"] > 3
    np.testing.assert_array_equal(ra["history"], rb["history"])
    np.testing.assert_array_equal(A.field(0, "p"), B.field(0, "p"))


@pytest.mark.parametrize("variant", ["T", "F"])
@pytest.mark.parametrize("precision", [0, 1])
def test_time_steps_bitwise(gpu, monkeypatch, variant, precision):
    mgm = gpu
    var = getattr(mgm, "VAR_" + variant)
    print(f"[gpu small relax] 2D cavity, variant {variant}, precision {precision}", flush=True)

    def march(P):
        P.zero_fields()
        cases.lid_driven_cavity(P, timple=var == mgm.VAR_T)
        P.ns_set_params(**cases.ns_params(100.0, P.info(0)["d_avg"], 0.5, num_vcycles=3))
        return P.ns_march(var, steps=3, precision=precision)

    (A, ra), (B, rb) = run_both(mgm, monkeypatch, 2, (33, 17, 9), march)
    assert np.all(np.isfinite(ra["errors"]))
    np.testing.assert_array_equal(ra["errors"], rb["errors"])
    for name in NS_FIELDS:
        np.testing.assert_array_equal(A.field(0, name), B.field(0, name), err_msg=name)
