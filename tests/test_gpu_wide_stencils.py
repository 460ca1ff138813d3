"""The 3D stencils and device layout of the 1M-point cube against the oracle.

The cube benchmark runs PHS r^3 + cubic polynomials with multiplier 3.5:
m = 70-point stencils, 90 x 90 local systems.  On its 100^3 hierarchy
  * level 0 (above 262,144 rows) is ordered in Morton blocks, stored at slice
    height 8 with gather-sorted copies, and each colour step takes a slice list;
  * the mid levels use slice height 4 and one launch per colour class;
  * the small levels relax in one workgroup (sums unrolled to 72 slots);
  * the weight kernel's autotuner picks 6x16 or 7x16 register tiles.
The thresholds behind these choices are environment variables, so a 24^3 /
12^3 / 7^3 hierarchy reproduces the same layout level by level
(MGM_BLOCKED_ROWS = MGM_GATHER_SORT_ROWS = 8000, MGM_SMALL_ROWS = 1000), and
the oracle can check it:
  * weights, lap_Poison and transfer rows under every weight-kernel variant
    the autotuner may pick: within 1e-11 row-relative of the oracle's pivoted
    LU (hf/rbf.h:401-421, hf/mesh_functions.h:533-670);
  * colour-ordered H / T / F V-cycles and the graph-captured solve: bit for bit
    the oracle's sequential sweep in colour order (hf/heat_conduction.h:25-112,
    hf/timple_solver.h:235-328, hf/fractional_step_solver.h:168-246).
Two narrower stencils cover the other unrolled widths of the single-workgroup
relax: degree-2 polynomials with multiplier 3.5 (m = 35, slice height 16) and
4 (m = 40, slice height 8).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = (24, 12, 7)
BENCH_LAYOUT = {"MGM_BLOCKED_ROWS": "8000", "MGM_GATHER_SORT_ROWS": "8000", "MGM_SMALL_ROWS": "1000"}
PER_COLOUR = dict(BENCH_LAYOUT, MGM_SMALL_ROWS="0")
# level 1 in small Morton blocks as well (the 1M cube's 125k-row level 1: 512-row blocks)
MID_BLOCKED = dict(BENCH_LAYOUT, MGM_MID_BLOCKED_ROWS="1000", MGM_MID_BLOCK_ROWS="256")
# id: (polynomial degree, multiplier, layout)
CASES = {
    "m70_bench_layout": (3, 3.5, BENCH_LAYOUT),
    "m70_per_colour": (3, 3.5, PER_COLOUR),
    "m70_mid_blocked": (3, 3.5, MID_BLOCKED),
    "m35_slice16": (2, 3.5, BENCH_LAYOUT),
    "m40_slice8": (2, 4.0, BENCH_LAYOUT),
}
_ORACLE_CACHE = {}


def oracle_rows(oracle, P, key):
    """The oracle's LU weights, lap_Poison and transfer rows on the GPU's own
    stencils and geometry (cached per stencil kind: identical across layouts)."""
    if key in _ORACLE_CACHE:
        return _ORACLE_CACHE[key]
    L = P.L
    lv = []
    for k in range(L):
        g = P.geometry(k)
        info = P.info(k)
        d = dict(g, cloud=P.cloud(k), nb=info["nb"], poly=info["poly_degree"],
                 rp=P.points(k, "restriction"), pp=P.points(k, "prolongation"),
                 ctag=np.zeros(info["N"], np.uint8))
        lv.append(d)
    out = []
    for k, d in enumerate(lv):
        Dx, Dy, Dz, lap = oracle.weights(d, 3, 3, d["poly"], method=1)
        rows = dict(Dx=Dx, Dy=Dy, Dz=Dz, lap=lap, lap_Poison=oracle.lap_poisson(d, 3, Dx, Dy, Dz, lap))
        if k > 0:
            rows["restr"] = oracle.restriction(lv[k - 1], d, 3, 3, lv[k - 1]["poly"], row_mode=1, method=1)
        if k < L - 1:
            rows["prol"] = oracle.prolongation(d, lv[k + 1], 3, 3, lv[k + 1]["poly"], row_mode=1, method=1)
        out.append(rows)
    _ORACLE_CACHE[key] = out
    return out


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request, gpu):
    poly, mult, env = CASES[request.param]
    mp = pytest.MonkeyPatch()
    for k, v in env.items():
        mp.setenv(k, v)
    mp.delenv("MGM_LU_VARIANT", raising=False)
    print(f"[gpu wide stencils] {request.param}: 3D {SIZES}, poly {poly}, multiplier {mult}, {env}", flush=True)
    cfg = gpu.config(dimension=3, poly_degree=poly, coarse_poly_degree=poly, phs_degree=3,
                     cloud_size_multiplier=mult, omega=1.0, num_relax=3, transfer_rows=gpu.ROWS_BY_TAG,
                     sweep_order=gpu.ORDER_COLOUR)
    P = gpu.Problem(cfg, [gpu.Points.generate(3, n, 1 + k, 0.2) for k, n in enumerate(SIZES)]).setup()
    yield request.param, P
    mp.undo()


def test_layout_is_the_benchmarks(case):
    name, P = case
    i0, i1, i2 = (P.info(k) for k in range(3))
    assert i0["blocked"] and i0["gather_sorted"] and not i0["small_relax"]
    mid = "MGM_MID_BLOCKED_ROWS" in CASES[name][2]
    assert i1["blocked"] == mid and (i1["window_blocks"] > 1) == mid
    assert not i1["gather_sorted"] and not i1["small_relax"]
    m = i0["m"]
    if m > 36:
        assert (i0["slice_height"], i1["slice_height"]) == (8, 8 if mid else 4)
    else:
        assert (i0["slice_height"], i1["slice_height"]) == (16, 16)
    assert i2["small_relax"] == (CASES[name][2]["MGM_SMALL_ROWS"] != "0")
    assert i0["device_rows"] > i0["N"]          # per-(block, colour) padding
    print(f"  m={m} mpn={m + i0['n']} colours {[P.info(k)['num_colours'] for k in range(3)]} "
          f"slice heights {[P.info(k)['slice_height'] for k in range(3)]}", flush=True)


def _lu_variants(mpn):
    t0 = next(t for t in range(2, 9) if 16 * t >= mpn + 4)
    return [(t, w) for t in (t0, t0 + 1) for w in (4, 1) if t <= 8 and not (w == 4 and t > 6)]


def _row_rel(G, W):
    sc = np.abs(W).max(axis=1)
    zero = sc == 0
    assert not np.abs(G[zero]).any(), "rows the oracle leaves empty must stay empty"
    return (np.abs(G - W).max(axis=1)[~zero] / sc[~zero]).max()


def test_weights_every_lu_variant(case, oracle, monkeypatch):
    name, P = case
    ref = oracle_rows(oracle, P, name.split("_")[0])
    mpn = P.info(0)["m"] + P.info(0)["n"]
    print(f"  autotuned weight kernel at mpn {mpn}: {P.weight_kernel_choice(0)}", flush=True)
    for tile, wps in _lu_variants(mpn):
        monkeypatch.setenv("MGM_LU_VARIANT", f"{tile},{wps}")
        for k in range(P.L):
            P.rebuild_weights(k)
        worst = 0.0
        for k in range(P.L):
            for key, W in ref[k].items():
                e = _row_rel(P.matrix(k, key), W)
                worst = max(worst, e)
                assert e <= 1e-11, (name, tile, wps, k, key, e)
        print(f"  tile {tile} wps {wps}: worst row-relative difference from the oracle {worst:.2e}", flush=True)
    monkeypatch.delenv("MGM_LU_VARIANT")
    for k in range(P.L):
        P.rebuild_weights(k)


def _oracle_hierarchy(oracle, P):
    lv = []
    for k in range(P.L):
        col = P.colours(k)
        d = dict(cloud=P.cloud(k), Lp=P.matrix(k, "lap_Poison"), lap=P.matrix(k, "lap"),
                 btag=P.geometry(k)["btag"], rp=P.points(k, "restriction"), pp=P.points(k, "prolongation"),
                 order=np.lexsort((np.arange(len(col)), col)).astype(np.int32))
        if k > 0:
            d["R"] = P.matrix(k, "restr")
        if k < P.L - 1:
            d["P"] = P.matrix(k, "prol")
        lv.append(d)
    return oracle.Hierarchy(lv)


def _rows(P, dirichlet):
    """Operator rows of the run: the Neumann rows the weights give, or identity
    rows on every boundary node (the cube benchmark's Dirichlet problem)."""
    for k in range(P.L):
        P.rebuild_weights(k)
    if dirichlet:
        P.dirichlet_rows()


# SOR factor 1; 0.85, under-relaxation, the single factor the benchmark's
# autotuner picks; and a factor per level
OMEGAS = [1.0, 0.85, (0.7, 0.6, 1.2)]


def _set_omega(P, omega):
    if isinstance(omega, tuple):
        P.set_smoother_levels(list(omega))
        assert P.smoother_levels() == [float(np.float32(w)) for w in omega]
    else:
        P.set_smoother(omega, 3)


@pytest.mark.parametrize("omega", OMEGAS)
@pytest.mark.parametrize("variant,dirichlet", [("H", True), ("H", False), ("T", False), ("F", False)])
def test_colour_vcycles_bitwise(case, oracle, variant, dirichlet, omega):
    name, P = case
    import multigrid_meshless_amd as mgm
    var = getattr(mgm, "VAR_" + variant)
    _set_omega(P, omega)
    _rows(P, dirichlet)
    H = _oracle_hierarchy(oracle, P)
    f = np.random.default_rng(11).standard_normal(P.info(0)["N"])
    if dirichlet:
        f[P.geometry(0)["btag"] != 0] = 0.0
    H.lv[0]["s"][:] = f
    P.zero_fields()
    P.set_field(0, "source", f)
    for _ in range(3):
        H.vcycle(var, list(omega) if isinstance(omega, tuple) else omega, 3)
    P.vcycle(var, 3)
    pf = "pprime" if variant == "T" else "p"
    for k in range(P.L):
        np.testing.assert_array_equal(P.field(k, pf), H.lv[k]["p"], err_msg=f"{name} {variant} L{k}")
        np.testing.assert_array_equal(P.field(k, "res"), H.lv[k]["res"], err_msg=f"{name} {variant} L{k} res")
        np.testing.assert_array_equal(P.field(k, "source"), H.lv[k]["s"], err_msg=f"{name} {variant} L{k} s")
    assert np.abs(P.field(0, pf)).max() > 0


@pytest.mark.parametrize("omega", OMEGAS)
def test_graph_solve_matches_oracle(case, oracle, omega):
    """mgm_solve replays one captured V-cycle + fine residual + norm per cycle:
    its fields equal the oracle's after the same cycles, bit for bit, and its
    residual history agrees to rounding (the GPU sums the norm as a tree)."""
    name, P = case
    _set_omega(P, omega)
    _rows(P, True)
    H = _oracle_hierarchy(oracle, P)
    btag = P.geometry(0)["btag"]
    f = np.random.default_rng(12).standard_normal(P.info(0)["N"])
    f[btag != 0] = 0.0
    H.lv[0]["s"][:] = f
    P.zero_fields()
    P.set_field(0, "source", f)
    ncyc = 4
    r = P.solve(1e-300, ncyc)
    assert r["cycles"] == ncyc
    fn = oracle.l2_norm_gen(btag, f)
    hist = []
    for _ in range(ncyc):
        H.vcycle(oracle.VAR_H, list(omega) if isinstance(omega, tuple) else omega, 3)
        H.residual(oracle.VAR_H, 0)
        hist.append(oracle.l2_norm_gen(btag, H.lv[0]["res"]) / fn)
    np.testing.assert_array_equal(P.field(0, "p"), H.lv[0]["p"], err_msg=name)
    np.testing.assert_allclose(r["history"][1:], hist, rtol=1e-12, atol=0)
    assert hist[-1] < hist[0]


def test_tune_omega(case):
    """mgm_tune_omega keeps the factor with the fewest cycles to the tolerance:
    no more than omega = 1 needs, equal to a plain solve at the chosen factor;
    the iterate and source are restored and the factor stays set."""
    name, P = case
    P.set_smoother(1.0, 3)
    _rows(P, True)
    btag = P.geometry(0)["btag"]
    f = np.random.default_rng(13).standard_normal(P.info(0)["N"])
    f[btag != 0] = 0.0
    P.zero_fields()
    P.set_field(0, "source", f)
    r1 = P.solve(1e-8, 300)
    P.zero_fields()
    P.set_field(0, "source", f)
    t = P.tune_omega(1e-8, 300, lo=0.6, hi=1.0, step=0.1)
    # the grid 0.6 .. 1.0, then the midpoints beside the best (so 0.55 .. 1.05)
    assert 0.55 - 1e-6 <= t["omega"] <= 1.05 + 1e-6 and t["cycles"] <= r1["cycles"], (t, r1["cycles"])
    np.testing.assert_array_equal(P.field(0, "source"), f)
    assert not np.abs(P.field(0, "p")).any()
    r = P.solve(1e-8, 300)
    assert r["cycles"] == t["cycles"]
    print(f"  {name}: omega {t['omega']:.3f} {t['cycles']} cycles (omega 1: {r1['cycles']})", flush=True)
    P.set_smoother(1.0, 3)


def test_tune_omega_levels(case):
    """mgm_tune_omega_levels: coordinate descent from the single tuned factor
    never needs more cycles than it; a plain solve at the chosen per-level
    factors takes the reported count; iterate and source are restored."""
    name, P = case
    P.set_smoother(1.0, 3)
    _rows(P, True)
    btag = P.geometry(0)["btag"]
    f = np.random.default_rng(14).standard_normal(P.info(0)["N"])
    f[btag != 0] = 0.0
    P.zero_fields()
    P.set_field(0, "source", f)
    t1 = P.tune_omega(1e-8, 300, lo=0.6, hi=1.0, step=0.1)
    t = P.tune_omega_levels(1e-8, 300, lo=0.6, hi=1.4, step=0.2, sweeps=2)
    assert len(t["omega"]) == P.L and t["cycles"] <= t1["cycles"], (t, t1)
    assert P.smoother_levels() == t["omega"]
    np.testing.assert_array_equal(P.field(0, "source"), f)
    assert not np.abs(P.field(0, "p")).any()
    r = P.solve(1e-8, 300)
    assert r["cycles"] == t["cycles"]
    print(f"  {name}: per-level omega {t['omega']} {t['cycles']} cycles (single {t1['omega']:.3f}: "
          f"{t1['cycles']})", flush=True)
    P.set_smoother(1.0, 3)
    assert P.smoother_levels() == [1.0] * P.L


# weighted Jacobi on every level, and Jacobi / SOR alternating (H variant)
JACOBI = [((1, 1, 1), (0.7, 0.8, 1.6)), ((1, 0, 1), (0.8, 0.6, 1.2))]


@pytest.mark.parametrize("kinds,omega", JACOBI, ids=["all", "mixed"])
@pytest.mark.parametrize("dirichlet", [True, False])
def test_jacobi_vcycles_bitwise(case, oracle, kinds, omega, dirichlet):
    """Weighted-Jacobi levels (one full pass per sweep: LDS windows on the
    blocked level 0, slabs elsewhere, FP64) against the oracle's cycle with
    the same kinds and factors: every level's p, residual and source bit for
    bit after three cycles."""
    name, P = case
    import multigrid_meshless_amd as mgm
    P.set_smoother_levels(list(omega))
    P.set_smoother_kinds(list(kinds))
    try:
        _rows(P, dirichlet)
        H = _oracle_hierarchy(oracle, P)
        f = np.random.default_rng(15).standard_normal(P.info(0)["N"])
        if dirichlet:
            f[P.geometry(0)["btag"] != 0] = 0.0
        H.lv[0]["s"][:] = f
        P.zero_fields()
        P.set_field(0, "source", f)
        for _ in range(3):
            H.vcycle(oracle.VAR_H, list(omega), 3, kinds=list(kinds))
        P.vcycle(mgm.VAR_H, 3)
        for k in range(P.L):
            np.testing.assert_array_equal(P.field(k, "p"), H.lv[k]["p"], err_msg=f"{name} {kinds} L{k}")
            np.testing.assert_array_equal(P.field(k, "res"), H.lv[k]["res"], err_msg=f"{name} {kinds} L{k} res")
            np.testing.assert_array_equal(P.field(k, "source"), H.lv[k]["s"], err_msg=f"{name} {kinds} L{k} s")
        assert np.abs(P.field(0, "p")).max() > 0
    finally:
        P.set_smoother_kinds(None)
        P.set_smoother(1.0, 3)


def test_jacobi_graph_solve_matches_oracle(case, oracle):
    """The captured solve with Jacobi levels: fields bit for bit the oracle's
    after the same cycles, residual history to rounding."""
    name, P = case
    kinds, omega = JACOBI[0]
    P.set_smoother_levels(list(omega))
    P.set_smoother_kinds(list(kinds))
    try:
        _rows(P, True)
        H = _oracle_hierarchy(oracle, P)
        btag = P.geometry(0)["btag"]
        f = np.random.default_rng(16).standard_normal(P.info(0)["N"])
        f[btag != 0] = 0.0
        H.lv[0]["s"][:] = f
        P.zero_fields()
        P.set_field(0, "source", f)
        ncyc = 4
        r = P.solve(1e-300, ncyc)
        fn = oracle.l2_norm_gen(btag, f)
        hist = []
        for _ in range(ncyc):
            H.vcycle(oracle.VAR_H, list(omega), 3, kinds=list(kinds))
            H.residual(oracle.VAR_H, 0)
            hist.append(oracle.l2_norm_gen(btag, H.lv[0]["res"]) / fn)
        np.testing.assert_array_equal(P.field(0, "p"), H.lv[0]["p"], err_msg=name)
        np.testing.assert_allclose(r["history"][1:], hist, rtol=1e-12, atol=0)
        assert hist[-1] < hist[0]
    finally:
        P.set_smoother_kinds(None)
        P.set_smoother(1.0, 3)


def _timed(P, f):
    P.zero_fields()
    P.set_field(0, "source", f)
    return P.solve(1e-8, 300)["seconds"]


def test_tune_smoother(case):
    """mgm_tune_smoother searches SOR and Jacobi per level, ranked by solve
    time: no slower than the per-level SOR factors (10% for timing noise); a
    plain solve at its choice takes the reported count; FP32 inner cycles with
    the same choice reach 1e-10."""
    name, P = case
    P.set_smoother(1.0, 3)
    P.set_smoother_kinds(None)
    _rows(P, True)
    btag = P.geometry(0)["btag"]
    f = np.random.default_rng(17).standard_normal(P.info(0)["N"])
    f[btag != 0] = 0.0
    try:
        P.zero_fields()
        P.set_field(0, "source", f)
        ts = P.tune_omega_levels(1e-8, 300, lo=0.6, hi=1.4, step=0.2, sweeps=2)
        t_sor = min(_timed(P, f) for _ in range(3))
        P.zero_fields()          # the tuner starts from the current iterate
        P.set_field(0, "source", f)
        t = P.tune_smoother(1e-8, 300, lo=0.6, hi=1.4, step=0.2, sweeps=2)
        assert P.smoother_kinds() == t["kind"] and P.smoother_levels() == t["omega"]
        P.zero_fields()
        P.set_field(0, "source", f)
        assert P.solve(1e-8, 300)["cycles"] == t["cycles"]
        t_tuned = min(_timed(P, f) for _ in range(3))
        assert t_tuned <= 1.1 * t_sor, (t_tuned, t_sor)
        p64 = P.field(0, "p")
        P.zero_fields()
        P.set_field(0, "source", f)
        r32 = P.solve(1e-10, 300, precision=1)
        assert r32["history"][-1] <= 1e-10
        print(f"  {name}: kinds {t['kind']} omega {[round(w, 2) for w in t['omega']]} {t['cycles']} cycles "
              f"(SOR per level: {ts['cycles']}); FP32 inner to 1e-10 in {r32['cycles']}", flush=True)
        assert np.abs(P.field(0, "p") - p64).max() <= 1e-6 * np.abs(p64).max()
    finally:
        P.set_smoother_kinds(None)
        P.set_smoother(1.0, 3)


@pytest.mark.parametrize("num_relax", [1, 2, 3, 5])
def test_fused_residual_jacobi_bitwise(case, monkeypatch, num_relax):
    """With weighted Jacobi on level 0 the solve forms each fine residual and
    the next cycle's first level-0 sweep in one window pass: residual history
    and fields bit for bit those of separate passes, for 1, 2 and 3 sweeps
    (the remaining sweeps end in p with no copy)."""
    name, P = case
    kinds, omega = JACOBI[0]
    P.set_smoother(1.0, num_relax)
    P.set_smoother_levels(list(omega))
    P.set_smoother_kinds(list(kinds))
    try:
        _rows(P, True)
        f = np.random.default_rng(19).standard_normal(P.info(0)["N"])
        f[P.geometry(0)["btag"] != 0] = 0.0
        out = []
        for fuse in ("1", "0"):
            monkeypatch.setenv("MGM_FUSE_RESIDUAL", fuse)
            P.zero_fields()
            P.set_field(0, "source", f)
            r = P.solve(1e-300, 5)
            out.append((np.asarray(r["history"]), P.field(0, "p"), P.field(1, "p")))
        for a, b in zip(out[0], out[1]):
            np.testing.assert_array_equal(a, b, err_msg=f"{name} num_relax {num_relax}")
        assert out[0][0][-1] < out[0][0][0]
    finally:
        P.set_smoother_kinds(None)
        P.set_smoother(1.0, 3)


def test_fp32_zero_start_jacobi_bitwise(case, monkeypatch):
    """The FP32 inner cycle's first Jacobi sweep on each level starts from the
    zeroed correction and runs without a stencil pass: residual history and
    field bit for bit those of the full first pass (MGM_JACOBI_ZERO=0), with
    FP32 and with BF16 coarse weights."""
    name, P = case
    kinds, omega = JACOBI[0]
    P.set_smoother_levels(list(omega))
    P.set_smoother_kinds(list(kinds))
    try:
        _rows(P, True)
        f = np.random.default_rng(20).standard_normal(P.info(0)["N"])
        f[P.geometry(0)["btag"] != 0] = 0.0
        for prec in (1, 2):
            out = []
            for zero in ("1", "0"):
                monkeypatch.setenv("MGM_JACOBI_ZERO", zero)
                P.zero_fields()
                P.set_field(0, "source", f)
                r = P.solve(1e-300, 4, precision=prec)
                out.append((np.asarray(r["history"]), P.field(0, "p")))
            np.testing.assert_array_equal(out[0][0], out[1][0], err_msg=f"{name} precision {prec}")
            np.testing.assert_array_equal(out[0][1], out[1][1], err_msg=f"{name} precision {prec}")
            assert out[0][0][-1] < out[0][0][0]
    finally:
        P.set_smoother_kinds(None)
        P.set_smoother(1.0, 3)


def test_window_prolongation_bitwise(case, monkeypatch):
    """The prolongation into the blocked level 0 in the LDS-window form (the
    coarse values its block references staged once, 16-bit places) equals the
    gathering form bit for bit: residual history and field of solves in FP64,
    FP32 and FP32 + BF16 coarse weights."""
    name, P = case
    assert P.info(0)["prolong_window_max"] > 0
    if not P.info(1)["blocked"]:
        assert P.info(1)["prolong_window_max"] == 0
    kinds, omega = JACOBI[0]
    P.set_smoother_levels(list(omega))
    P.set_smoother_kinds(list(kinds))
    try:
        _rows(P, True)
        f = np.random.default_rng(21).standard_normal(P.info(0)["N"])
        f[P.geometry(0)["btag"] != 0] = 0.0
        for prec in (0, 1, 2):
            out = []
            for win in ("1", "0"):
                monkeypatch.setenv("MGM_WINDOW_PROLONG", win)
                P.zero_fields()
                P.set_field(0, "source", f)
                r = P.solve(1e-300, 4, precision=prec)
                out.append((np.asarray(r["history"]), P.field(0, "p")))
            np.testing.assert_array_equal(out[0][0], out[1][0], err_msg=f"{name} precision {prec}")
            np.testing.assert_array_equal(out[0][1], out[1][1], err_msg=f"{name} precision {prec}")
            assert out[0][0][-1] < out[0][0][0]
    finally:
        P.set_smoother_kinds(None)
        P.set_smoother(1.0, 3)
