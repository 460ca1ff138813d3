"""Benchmark: time to 1e-8 relative residual, 3D 1M-point multigrid Poisson.

Workload (one "step" = one solve to tolerance from p = 0, operators resident in
HBM): jittered 100^3 cube (seed 1; 998,816 nodes after corner/edge removal),
5 levels 100^3/50^3/25^3/13^3/7^3 (seeds 1..5), PHS r^3 + cubic polynomials,
m = 70 stencils (multiplier 3.5), Dirichlet-zero identity rows on every level,
RHS = sum of 16 Fourier modes (amplitudes N(0,1), SplitMix64 seed 2), V-cycles
of multicolour SOR (3 sweeps) until l2_norm_gen(r)/l2_norm_gen(f) <= 1e-8.  The
smoother is autotuned before the timed solves, on another source (the 16-mode
sum drawn with seed 3): one SOR factor for every level (mgm_tune_omega), then
each level's smoother -- multicolour SOR or weighted Jacobi -- and factor by
coordinate descent from it (mgm_tune_smoother); each keeps the fewest cycles
to the tolerance.  The timed solves use the per-level choice; the SOR omega = 1
and single-factor solves are timed beside them.  --omega fixes SOR factors
instead (one, or one per level); --smoother fixes the kinds (s/j per level).

Prints one JSON line (rank 0, one GPU).  `roofline` is the FP64 stencil apply
(lap_Poison over level 0) timed with HIP events on the library stream;
`cpu_baseline` is the oracle's whole solve to the same tolerance, its V-cycle in
the same colour order on the same operators, the rows of each colour class
swept in parallel on every available host thread (the 1-thread rate and a
sampled CPU weight-build rate are in its sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_FLOPS = 78.6e12
FP32_MFMA_PEAK_FLOPS = 157.3e12
COPY_BYTES = 1 << 30
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "stencil_apply_3d_1m_traffic.md")


def measured_traffic():
    """HBM bytes per level-0 stencil apply of the 100^3 workload from the
    committed rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, calibrated on
    the streaming copy), or None."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(float(line.split(":")[1]))
    except OSError:
        pass
    return None


def build_problem(mgm, n, levels, seed0=1):
    """The 3D cube Poisson hierarchy (cases.cube_poisson)."""
    from multigrid_meshless_amd import cases
    return cases.cube_poisson(n, levels, seed0)


def set_rhs(P, seed=2):
    from multigrid_meshless_amd import cases
    return cases.cube_rhs(P, seed)


def reset(P, f):
    P.zero_fields()
    P.set_field(0, "source", f)


def lu32_mfma_count(mpn, tiles):
    """v_mfma_f32_16x16x4_f32 instructions of one factorisation in the matrix-core
    weight kernel: panels of four columns, each updating every active tile
    column (T - tc of them, less the panel's own once its four columns are done)
    in each of the T tile rows."""
    count = 0
    for c in range((mpn + 3) // 4):
        tc, gc = divmod(c, 4)
        count += tiles * ((tiles - tc) - (1 if gc == 3 else 0))
    return count


def matrix_core_build(P, mpn):
    """Level-0 weight build with the matrix-core kernel forced (FP32 LU on MFMA,
    FP64 refinement): systems/s and the matrix-core flop rate of its
    factorisations against the FP32 matrix peak."""
    os.environ["MGM_LU_KERNEL"] = "mfma"
    try:
        t, n = P.rebuild_weights(0)
    finally:
        del os.environ["MGM_LU_KERNEL"]
    P.rebuild_weights(0)                # restore the autotuned kernel's rows
    tiles = -(-mpn // 16)
    flops = lu32_mfma_count(mpn, tiles) * 16 * 16 * 4 * 2
    # the flop rate counts the factorisations only; each solve adds 8 T^2 FP32
    # MFMA (block substitutions) and each refinement step 4 T^2 FP64 MFMA
    # (the residual), their number depending on how fast the refinement converges
    out = {"kernel_s": round(t, 4), "systems_per_s": round(n / t), "tiles": tiles,
           "mfma_per_system": lu32_mfma_count(mpn, tiles),
           "solve_mfma_f32_per_solve": 8 * tiles * tiles, "residual_mfma_f64_per_step": 4 * tiles * tiles,
           "mfma_tflops": round(n * flops / t / 1e12, 3),
           "frac_fp32_matrix_peak": round(n * flops / t / FP32_MFMA_PEAK_FLOPS, 4)}
    print(f"[bench] weights level 0, matrix-core kernel: {n / t:.0f} systems/s, "
          f"{out['mfma_tflops']} TFLOP/s on MFMA", file=sys.stderr, flush=True)
    return out


def uneven_build(mgm, args):
    """The uneven batched build: the cube's hierarchy with wide boundary stencils
    (boundary multiplier 5.25: m_b = 105, 125 unknowns; 70 points and 90
    unknowns inside), every level's systems built in one launch per tile class.
    Level-0 rebuild rate and the whole hierarchy's weight phase in setup."""
    from multigrid_meshless_amd import cases
    P, sizes, t_setup = cases.cube_poisson(args.n, args.levels, 1, boundary_multiplier=5.25)
    systems = sum(P.info(k)["N"] for k in range(P.L)) + sum(P.boundary_info(k)[0] for k in range(P.L))
    t_all = P.timings()["weights"]
    t0, n0 = P.rebuild_weights(0)
    nbw, mb = P.boundary_info(0)
    out = {"workload": f"{sizes[0]}^3 cube, {len(sizes)} levels, m=70 inside (90 unknowns), "
                       f"m_b={mb} on the boundary ({mb + 20} unknowns)",
           "level0_systems": n0, "level0_wide_systems": nbw, "level0_kernel_s": round(t0, 4),
           "level0_systems_per_s": round(n0 / t0),
           "hierarchy_systems": systems, "hierarchy_weight_phase_s": round(t_all, 4),
           "kernel_90": dict(zip(("tile", "waves_per_simd"), P.weight_kernel_choice(0) or (None, None))),
           "kernel_125": dict(zip(("tile", "waves_per_simd"),
                                  mgm_kernel_choice(mgm, mb + 20) or (None, None)))}
    print(f"[bench] uneven weight build: level 0 {n0} systems ({nbw} wide) in {t0:.3f}s = {n0 / t0:.0f} "
          f"systems/s; hierarchy {systems} systems, weight phase {t_all:.3f}s", file=sys.stderr, flush=True)
    del P
    return out


def mgm_kernel_choice(mgm, mpn):
    import ctypes as C
    t, w = C.c_int(), C.c_int()
    if mgm._lib.lib().mgm_weight_kernel_choice(mpn, C.byref(t), C.byref(w)) != 0:
        return None
    return t.value, w.value


def build_cavity(mgm, n, levels, Re, courant, num_relax, num_vcycles, seed0=31):
    """Lid-driven cavity (init/init_LDC.c): jittered n^2 square and coarser clouds
    (halving), PHS r^5 + degree-4 polynomials on level 0 and degree 3 below,
    multiplier 2.0, dt from the CFL number (hf/general_functions.h:34-41)."""
    from multigrid_meshless_amd import cases
    sizes = [((n - 1) >> k) + 1 for k in range(levels)]
    pts = [mgm.Points.generate(2, s, seed0 + k, 0.2) for k, s in enumerate(sizes)]
    cfg = mgm.config(dimension=2, poly_degree=4, coarse_poly_degree=3, phs_degree=5,
                     cloud_size_multiplier=2.0, omega=1.0, num_relax=num_relax,
                     transfer_rows=mgm.ROWS_BY_TAG, sweep_order=mgm.ORDER_COLOUR)
    P = mgm.Problem(cfg, pts).setup()
    cases.lid_driven_cavity(P)
    prm = cases.ns_params(Re, P.info(0)["d_avg"], courant, num_vcycles=num_vcycles)
    P.ns_set_params(**prm)
    return P, sizes, prm


def build_annulus(mgm, n, levels, seed0):
    """2D annulus r in [0.5, 1] (cases.annulus): n points per side of the
    bounding grid on level 0, `levels` clouds halving in spacing; PHS r^3 +
    cubic polynomials, multiplier 2 (m = 20); Dirichlet rows on both circles."""
    from multigrid_meshless_amd import cases
    sizes = [((n - 1) >> k) + 1 for k in range(levels)]
    pts = [cases.annulus(s, seed0 + k) for k, s in enumerate(sizes)]
    cfg = mgm.config(dimension=2, poly_degree=3, coarse_poly_degree=3, phs_degree=3, cloud_size_multiplier=2.0,
                     omega=1.0, num_relax=3, transfer_rows=mgm.ROWS_BY_TAG, sweep_order=mgm.ORDER_COLOUR)
    P = mgm.Problem(cfg, pts).setup()
    P.dirichlet_rows()
    return P, sizes


def annulus_learned(mgm, args):
    """Learned smoother on the annulus: plain FP64 SOR V-cycles against the
    learned first sweep (FP64 and FP32 inner cycles), network trained on a
    coarser annulus hierarchy (181 per side, 3 levels) and moved over."""
    from multigrid_meshless_amd import cases
    P, sizes = build_annulus(mgm, args.ann_n, args.ann_levels, 41)
    g = P.geometry(0)
    f = cases.fourier_rhs(g["x"], g["y"], np.zeros_like(g["x"]))
    f[g["btag"] != 0] = 0.0
    N = P.info(0)["N"]

    def run(prec):
        for _ in range(2):              # first: graph capture and copies; second: timed
            reset(P, f)
            r = P.solve(args.tol, args.ann_max_cycles, precision=prec)
        return r

    r = run(0)
    p0 = P.field(0, "p")
    out = {"workload": f"2D annulus r in [0.5, 1], {N} jittered points, {len(sizes)} levels {sizes} "
                       f"(grid points per side), m={P.info(0)['m']}, Dirichlet zero on both circles, "
                       f"16-mode Fourier source, tol {args.tol}",
           "plain_fp64": {"s": round(r["seconds"], 5), "cycles": r["cycles"],
                          "final_ratio": float(r["history"][-1])}}
    t_plain = r["seconds"]
    T, _ = build_annulus(mgm, 181, 3, 51)
    t0 = time.perf_counter()
    l0, l1 = T.train_smoother(snapshots=8, samples_per_snapshot=8192, epochs=30, batch=4096, lr=1e-3, seed=7)
    out["train_s"] = round(time.perf_counter() - t0, 2)
    out["train_loss"] = [round(l0, 4), round(l1, 4)]
    P.set_mlp_weights(T.mlp_weights())
    del T
    P.use_learned_smoother(True)
    for prec, tag in ((0, "learned_fp64"), (1, "learned_fp32")):
        r = run(prec)
        out[tag] = {"s": round(r["seconds"], 5), "cycles": r["cycles"], "final_ratio": float(r["history"][-1]),
                    "fell_back_to_sor": P.learned_fell_back(),
                    "max_rel_diff_vs_plain_solution": float(np.abs(P.field(0, "p") - p0).max() / np.abs(p0).max()),
                    "speedup_vs_plain": round(t_plain / r["seconds"], 3)}
        print(f"[bench] annulus {N} points: plain {out['plain_fp64']['cycles']} cycles {t_plain:.4f}s; {tag} "
              f"{r['cycles']} cycles {r['seconds']:.4f}s", file=sys.stderr, flush=True)
    P.use_learned_smoother(False)
    return out


def cavity_march(mgm, args):
    """Explicit fractional-step march of the lid-driven cavity, one whole step
    (its V-cycles included) captured as a hipGraph and replayed; FP64 and the
    FP32 pressure cycle in FP64 defect correction."""
    out = {}
    fields = {}
    for prec, tag in ((0, "fp64"), (1, "fp32_pressure_cycle")):
        P, sizes, prm = build_cavity(mgm, args.ldc_n, args.ldc_levels, args.ldc_re, 0.5, 5, 10)
        P.ns_march(mgm.VAR_F, steps=2, precision=prec)          # warm-up: graph build, FP32 copies
        P, sizes, prm = build_cavity(mgm, args.ldc_n, args.ldc_levels, args.ldc_re, 0.5, 5, 10)
        r = P.ns_march(mgm.VAR_F, steps=args.ldc_steps, precision=prec)
        fields[tag] = {k: P.field(0, k) for k in ("u", "v", "p")}
        out[tag] = {"steps": r["steps"], "s": round(r["seconds"], 4),
                    "ms_per_step": round(1e3 * r["seconds"] / max(r["steps"], 1), 4),
                    "vcycles_per_s": round(r["steps"] * prm["num_vcycles"] / r["seconds"], 1),
                    "steady_error_first_last": [float(r["errors"][0]), float(r["errors"][-1])]}
        print(f"[bench] cavity {sizes} Re {args.ldc_re} {tag}: {r['steps']} steps in {r['seconds']:.3f}s "
              f"= {out[tag]['ms_per_step']:.3f} ms/step, steady error {r['errors'][0]:.3e} -> "
              f"{r['errors'][-1]:.3e}", file=sys.stderr, flush=True)
        info = P.info(0)
        del P
    out["fp32_pressure_cycle"]["max_rel_diff_vs_fp64"] = {
        k: float(np.abs(fields["fp32_pressure_cycle"][k] - fields["fp64"][k]).max() / np.abs(fields["fp64"][k]).max())
        for k in ("u", "v", "p")}
    out["workload"] = (f"2D lid-driven cavity, fractional step, Re {args.ldc_re}, CFL 0.5 (dt {prm['dt']:.3e}), "
                       f"levels {sizes} ({info['N']} nodes, m={info['m']}), PHS r^5 + degree-4 polynomials "
                       f"(degree 3 coarse), SOR omega 1 x5 sweeps, 10 V-cycles per step")
    out["why_not_512"] = ("the reference's pressure multigrid (SOR on all-Neumann rows) diverges on this "
                          "cavity from 257^2 up, in FP64 too, so the 512^2 case cannot march")
    return out


def cpu_hierarchy(P):
    """The oracle's hierarchy on the GPU's operators, in the GPU's colour order
    (the rows of a colour class are swept in parallel, the same bits as the
    sequential sweep)."""
    import oracle
    lv = []
    for k in range(P.L):
        col = P.colours(k)
        order = np.lexsort((np.arange(len(col)), col)).astype(np.int32)
        classes = np.searchsorted(col[order], np.arange(col.max() + 2)).astype(np.int32)
        d = dict(cloud=P.cloud(k), Lp=P.matrix(k, "lap_Poison"), btag=P.geometry(k)["btag"],
                 rp=P.points(k, "restriction"), pp=P.points(k, "prolongation"), order=order, classes=classes)
        if k > 0:
            d["R"] = P.matrix(k, "restr")
        if k < P.L - 1:
            d["P"] = P.matrix(k, "prol")
        lv.append(d)
    return oracle.Hierarchy(lv), lv


def cpu_baseline(P, f, tol, max_cycles, weight_sample, omega, kinds):
    """The oracle on the host: (1) the same solve to tol on every available
    thread -- V-cycle, fine residual and l2_norm_gen norm per cycle, timed
    whole; (2) two cycles on one thread; (3) the reference's weight build
    (Gauss-Jordan inverse per node, hf/rbf.h:401-421) of the first
    weight_sample level-0 systems on every thread, as systems/s."""
    import oracle
    H, lv = cpu_hierarchy(P)
    btag = lv[0]["btag"]
    fnorm = oracle.l2_norm_gen(btag, f)
    nt = oracle.num_threads()

    def reset():
        for d in H.lv:
            d["p"][:] = 0
            d["s"][:] = 0
        H.lv[0]["s"][:] = f

    oracle.set_num_threads(nt)
    reset()
    t0 = time.perf_counter()
    cycles, ratio = 0, 1.0
    while cycles < max_cycles and ratio > tol:
        H.vcycle(oracle.VAR_H, omega, 3, kinds=kinds)
        H.residual(oracle.VAR_H, 0)
        ratio = oracle.l2_norm_gen(btag, H.lv[0]["res"]) / fnorm
        cycles += 1
    t_solve = time.perf_counter() - t0
    oracle.set_num_threads(1)
    reset()
    t0 = time.perf_counter()
    for _ in range(2):
        H.vcycle(oracle.VAR_H, omega, 3, kinds=kinds)
        H.residual(oracle.VAR_H, 0)
        oracle.l2_norm_gen(btag, H.lv[0]["res"])
    t1 = (time.perf_counter() - t0) / 2
    oracle.set_num_threads(nt)
    g = P.geometry(0)
    sub = dict(x=g["x"], y=g["y"], z=g["z"], cloud=np.ascontiguousarray(lv[0]["cloud"][:weight_sample]))
    info = P.info(0)
    t0 = time.perf_counter()
    oracle.weights(sub, 3, 3, info["poly_degree"], method=0)
    t_w = time.perf_counter() - t0
    return dict(threads=nt, solve_s=t_solve, cycles=cycles, final_ratio=ratio, s_per_cycle_1thread=t1,
                weight_systems=weight_sample, weight_s=t_w)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=100, help="points per axis on the finest level")
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--max-cycles", type=int, default=400)
    ap.add_argument("--omega", default="tune",
                    help="smoother factors of the cube solves: 'tune' (autotuned), one, or one per level (w0,w1,...)")
    ap.add_argument("--smoother", default=None,
                    help="with fixed --omega: smoother per level, s (SOR) or j (Jacobi), e.g. jjjjj")
    ap.add_argument("--learned-omega", default=None,
                    help="the learned solves' factors, one per level (default: tuned with --omega tune, "
                         "else --omega)")
    ap.add_argument("--learned-smoother", default=None,
                    help="with --learned-omega: the learned solves' smoother per level (s/j)")
    ap.add_argument("--cpu-weight-sample", type=int, default=16384,
                    help="level-0 systems of the CPU weight-build sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--no-mixed", action="store_true", help="skip the reduced-precision solves")
    ap.add_argument("--no-learned", action="store_true", help="skip training and the learned-smoother solve")
    ap.add_argument("--no-ldc", action="store_true", help="skip the lid-driven-cavity time march")
    ap.add_argument("--ldc-only", action="store_true", help="run only the lid-driven-cavity time march, print it, exit")
    ap.add_argument("--ldc-n", type=int, default=129, help="cavity points per side on the finest level")
    ap.add_argument("--ldc-levels", type=int, default=3)
    ap.add_argument("--ldc-re", type=float, default=100.0)
    ap.add_argument("--ldc-steps", type=int, default=2000)
    ap.add_argument("--no-annulus", action="store_true", help="skip the learned smoother on the 2D annulus")
    ap.add_argument("--no-uneven", action="store_true",
                    help="skip the uneven weight build (wide boundary stencils, a second hierarchy)")
    ap.add_argument("--ann-n", type=int, default=714, help="annulus grid points per side (714: 301k points)")
    ap.add_argument("--ann-levels", type=int, default=4)
    ap.add_argument("--ann-max-cycles", type=int, default=3000)
    ap.add_argument("--spmv-only", action="store_true",
                    help="set up, time the level-0 stencil apply and the streaming copy, print them, exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; runs on more than one GPU are not supported.", file=sys.stderr)
        sys.exit(2)

    import multigrid_meshless_amd as mgm
    mgm.init(0)
    if args.ldc_only:
        print(json.dumps({"cavity_time_march": cavity_march(mgm, args)}), flush=True)
        return
    P, sizes, t_setup = build_problem(mgm, args.n, args.levels)
    info0 = P.info(0)
    N, m = info0["N"], info0["m"]
    # per row: m FP64 weights and m column entries (16-bit window places on a
    # level with LDS windows, else int32 columns), x[r] of the window's own
    # rows (the halo values are the 2-3% re-read that the counters show) and y[r]
    col_bytes = 2 if info0["window_blocks"] else 4
    bytes_apply = N * (m * (8 + col_bytes) + 8 + 8)
    if args.spmv_only:
        # the roofline kernel and the copy of known size that calibrates HBM counters
        t_apply = P.time_spmv(0, args.spmv_reps)
        t_copy = mgm.stream_copy_seconds(COPY_BYTES, args.spmv_reps)
        print(json.dumps({"stencil_apply_us": round(t_apply * 1e6, 2),
                          "stencil_apply_GBps": round(bytes_apply / t_apply / 1e9, 1),
                          "copy_bytes": COPY_BYTES, "copy_us": round(t_copy * 1e6, 2),
                          "copy_GBps": round(2 * COPY_BYTES / t_copy / 1e9, 1)}), flush=True)
        return
    f = set_rhs(P)
    print(f"[bench] setup {t_setup:.2f}s levels={sizes} N0={N} m={m} colours0={info0['num_colours']} "
          f"phases={ {k: round(v, 3) for k, v in P.timings().items()} }", file=sys.stderr, flush=True)

    # SOR factors: autotuned on another source (untimed); the omega = 1 and
    # single-factor solves of the benchmark's source are timed for the record
    def timed_once():
        reset(P, f)
        P.solve(args.tol, args.max_cycles)
        reset(P, f)
        r = P.solve(args.tol, args.max_cycles)
        return {"s": round(r["seconds"], 5), "cycles": r["cycles"]}

    P.set_smoother(1.0, 3)
    omega1 = timed_once()
    single = None
    if args.omega == "tune":
        f_tune = set_rhs(P, seed=3)
        reset(P, f_tune)
        t0 = time.perf_counter()
        tuned = P.tune_omega(args.tol, args.max_cycles)
        single = dict(omega=tuned["omega"], **timed_once())
        reset(P, f_tune)
        tl = P.tune_smoother(args.tol, args.max_cycles, lo=0.5, hi=1.9, step=0.1, sweeps=3)
        omega, kinds, omega_tuned = tl["omega"], tl["kind"], True
        print(f"[bench] smoother tuned in {time.perf_counter() - t0:.1f}s on the seed-3 source: single SOR factor "
              f"{single['omega']:.3f} ({single['cycles']} cycles {single['s']:.4f}s here), per level "
              f"{''.join('sj'[k] for k in kinds)} {[round(w, 3) for w in omega]}; SOR omega 1: {omega1['cycles']} "
              f"cycles {omega1['s']:.4f}s", file=sys.stderr, flush=True)
    else:
        ws = [float(np.float32(float(v))) for v in args.omega.split(",")]
        omega, omega_tuned = (ws * P.L if len(ws) == 1 else ws), False
        kinds = [0] * P.L if not args.smoother else ["sj".index(ch) for ch in args.smoother]
    P.set_smoother_levels(omega)
    P.set_smoother_kinds(kinds)
    for _ in range(args.warmup):
        reset(P, f)
        r = P.solve(args.tol, args.max_cycles)
        print(f"[bench] warmup solve: {r['cycles']} cycles {r['seconds']:.4f}s "
              f"final {r['history'][-1]:.3e}", file=sys.stderr, flush=True)
    times, cycles, finals = [], [], []
    t_all = 0.0
    for _ in range(args.steps):
        reset(P, f)
        t0 = time.perf_counter()
        r = P.solve(args.tol, args.max_cycles)
        t_all += time.perf_counter() - t0
        times.append(r["seconds"])
        cycles.append(r["cycles"])
        finals.append(float(r["history"][-1]))
        print(f"[bench] solve: {r['cycles']} cycles {r['seconds']:.4f}s final {finals[-1]:.3e}",
              file=sys.stderr, flush=True)
    converged = all(v <= args.tol for v in finals)
    t_solve = float(np.median(times))
    p64 = P.field(0, "p")
    # a tightly converged FP64 solution: every solve's field error is measured against it
    reset(P, f)
    r_conv = P.solve(1e-12, 4 * args.max_cycles)
    p_conv = P.field(0, "p")

    def err_conv(p):
        return float(np.abs(p - p_conv).max() / np.abs(p_conv).max())

    plain_err = err_conv(p64)
    print(f"[bench] converged reference: {r_conv['cycles']} cycles to {r_conv['history'][-1]:.2e}; "
          f"plain 1e-8 solve field error {plain_err:.2e}", file=sys.stderr, flush=True)

    # reduced-precision inner V-cycles inside FP64 defect correction
    mixed = {}
    if not args.no_mixed:
        for prec, tag in ((1, "fp32"), (2, "fp32_bf16_coarse")):
            reset(P, f)
            P.solve(args.tol, args.max_cycles, precision=prec)      # builds the narrowed copies
            ts = []
            for _ in range(args.steps):
                reset(P, f)
                r = P.solve(args.tol, args.max_cycles, precision=prec)
                ts.append(r["seconds"])
            err = float(np.abs(P.field(0, "p") - p64).max() / np.abs(p64).max())
            mixed[tag] = {"s": round(float(np.median(ts)), 5), "cycles": r["cycles"],
                          "final_ratio": float(r["history"][-1]),
                          "max_rel_diff_vs_fp64_solution": err,
                          "max_rel_err_vs_converged": err_conv(P.field(0, "p"))}
            print(f"[bench] precision {prec} ({tag}): {r['cycles']} cycles {np.median(ts):.4f}s "
                  f"final {r['history'][-1]:.3e} field diff vs FP64 {err:.2e}", file=sys.stderr, flush=True)

    # learned smoother: trained on a 40^3 cloud with the same stencil settings,
    # weights moved to the 1M problem; FP64 solve, first sweep per level learned
    learned = None
    if not args.no_learned:
        Pt, _, _ = build_problem(mgm, 40, 3, seed0=21)
        t0 = time.perf_counter()
        l0, l1 = Pt.train_smoother(snapshots=8, samples_per_snapshot=8192, epochs=30, batch=4096, lr=1e-3,
                                   seed=7)
        t_train = time.perf_counter() - t0
        P.set_mlp_weights(Pt.mlp_weights())
        del Pt
        P.use_learned_smoother(True)
        learned = {"train_s": round(t_train, 2), "train_loss": [round(l0, 4), round(l1, 4)], "omega": omega}
        if args.learned_omega:
            lw = [float(np.float32(float(v))) for v in args.learned_omega.split(",")]
            lk = [0] * P.L if not args.learned_smoother else ["sj".index(ch) for ch in args.learned_smoother]
            P.set_smoother_levels(lw if len(lw) > 1 else lw * P.L)
            P.set_smoother_kinds(lk)
            learned["omega"], learned["smoother"] = P.smoother_levels(), "".join("sj"[k] for k in lk)
        elif args.omega == "tune":
            reset(P, set_rhs(P, seed=3))
            tl = P.tune_smoother(args.tol, args.max_cycles, lo=0.5, hi=1.9, step=0.1, sweeps=3)
            learned["omega"], learned["smoother"] = tl["omega"], "".join("sj"[k] for k in tl["kind"])
        for prec, tag in ((0, "fp64"), (1, "fp32"), (2, "fp32_bf16_coarse")):
            if prec and args.no_mixed:
                continue
            reset(P, f)
            P.solve(args.tol, args.max_cycles, precision=prec)
            ts = []
            for _ in range(args.steps):
                reset(P, f)
                r = P.solve(args.tol, args.max_cycles, precision=prec)
                ts.append(r["seconds"])
            fb = P.learned_fell_back()
            err = float(np.abs(P.field(0, "p") - p64).max() / np.abs(p64).max())
            learned[tag] = {"s": round(float(np.median(ts)), 5), "cycles": r["cycles"],
                            "final_ratio": float(r["history"][-1]), "fell_back_to_sor": fb,
                            "max_rel_diff_vs_fp64_solution": err,
                            "max_rel_err_vs_converged": err_conv(P.field(0, "p")),
                            "speedup_vs_fp64_sor": round(t_solve / float(np.median(ts)), 3)}
            print(f"[bench] learned smoother, precision {prec} ({tag}): {r['cycles']} cycles {np.median(ts):.4f}s "
                  f"final {r['history'][-1]:.3e} fell back {fb} field diff {err:.2e}", file=sys.stderr, flush=True)
        print(f"[bench] learned smoother trained in {t_train:.1f}s, loss {l0:.3f} -> {l1:.3f}", file=sys.stderr,
              flush=True)
        P.use_learned_smoother(False)
        P.set_smoother_levels(omega)
        P.set_smoother_kinds(kinds)

    # stencil apply roofline: lap_Poison over level 0, FP64 weights + column places
    t_apply = P.time_spmv(0, args.spmv_reps)
    achieved = bytes_apply / t_apply / 1e9
    t_copy = mgm.stream_copy_seconds(COPY_BYTES, args.spmv_reps)
    copy_gbs = 2 * COPY_BYTES / t_copy / 1e9

    # weight build of level 0 (kernel time only): systems/s and the algorithmic
    # flop rate 2/3 mpn^3 + 2 mpn^2 nrhs per system (LU + operator right-hand sides)
    t_w, n_sys = P.rebuild_weights(0)
    P.dirichlet_rows()                  # the rebuild restored the Neumann boundary rows
    mpn, nrhs = m + info0["n"], 4
    flop_sys = 2.0 / 3.0 * mpn ** 3 + 2.0 * mpn ** 2 * nrhs
    weights = {"systems": n_sys, "kernel_s": round(t_w, 4), "systems_per_s": round(n_sys / t_w),
               "mflop_per_system": round(flop_sys / 1e6, 4), "gflops": round(n_sys * flop_sys / t_w / 1e9, 1),
               "frac_fp64_peak": round(n_sys * flop_sys / t_w / FP64_PEAK_FLOPS, 4),
               "autotuned_kernel": dict(zip(("tile", "waves_per_simd"), P.weight_kernel_choice(0) or (None, None)))}
    print(f"[bench] weights level 0: {n_sys} systems in {t_w:.3f}s = {n_sys / t_w:.0f} systems/s, "
          f"{weights['gflops']} GFLOP/s", file=sys.stderr, flush=True)
    weights["matrix_core_kernel"] = matrix_core_build(P, mpn)
    P.dirichlet_rows()
    if not args.no_uneven:
        weights["uneven_wide_boundary"] = uneven_build(mgm, args)

    cavity = None if args.no_ldc else cavity_march(mgm, args)
    annulus = None if (args.no_annulus or args.no_learned) else annulus_learned(mgm, args)

    cpu = None
    if not args.no_cpu:
        c = cpu_baseline(P, f, args.tol, 2 * args.max_cycles, args.cpu_weight_sample, omega, kinds)
        cpu = {"value": round(c["solve_s"], 3), "unit": "s", "cores": c["threads"], "kind": "port",
               "sample": f"the whole solve to {args.tol} on the host: the oracle's V-cycle with the same smoothers "
                         f"({''.join('sj'[k] for k in kinds)}, factors {[round(w, 3) for w in omega]}; SOR in the "
                         f"GPU's colour order) on the same 1M-point operators (rows of a colour class in parallel on {c['threads']} threads), "
                         f"{c['cycles']} cycles to {c['final_ratio']:.2e} in {c['solve_s']:.2f} s; 1 thread: "
                         f"{c['s_per_cycle_1thread']:.3f} s/cycle (2 cycles timed) = "
                         f"{c['s_per_cycle_1thread'] * c['cycles']:.1f} s for the solve; weight build (the "
                         f"reference's Gauss-Jordan per node) of {c['weight_systems']} level-0 systems in "
                         f"{c['weight_s']:.2f} s on {c['threads']} threads = "
                         f"{c['weight_systems'] / c['weight_s']:.0f} systems/s",
               "weight_build_systems_per_s": round(c["weight_systems"] / c["weight_s"]),
               "solve_cycles": c["cycles"]}
        print(f"[bench] CPU baseline: solve {c['solve_s']:.2f} s ({c['cycles']} cycles, {c['threads']} threads), "
              f"weights {c['weight_systems'] / c['weight_s']:.0f} systems/s", file=sys.stderr, flush=True)

    out = {
        "metric": "time-to-1e-8 residual, 3D 1M-pt MG Poisson (s); stencil-apply HBM GB/s",
        "value": round(t_solve, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_solve * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D cube Poisson, {sizes[0]}^3 jittered cloud ({N} nodes), "
                               f"{len(sizes)} levels {sizes}, m={m}, FP64 V-cycle (per level: multicolour "
                               f"SOR or weighted Jacobi, 3 sweeps), tol {args.tol}",
                   "smoother_per_level": "".join("sj"[k] for k in kinds), "omega_per_level": omega,
                   "omega_tuned": omega_tuned, "num_relax": 3,
                   "omega_1": omega1, "omega_single_tuned": single,
                   "cycles": cycles[0], "converged": converged,
                   "vcycles_per_s": round(cycles[0] / t_solve, 1),
                   "final_ratio": finals[0], "setup_s": round(t_setup, 2),
                   "max_rel_err_vs_converged": plain_err,
                   "stencil_apply_GBps": round(achieved, 1),
                   "stream_copy_GBps_measured": round(copy_gbs, 1),
                   "stencil_apply_frac_of_measured_copy": round(achieved / copy_gbs, 4),
                   "reduced_precision": mixed,
                   "weight_build": weights,
                   "learned_smoother": learned,
                   "learned_smoother_annulus": annulus,
                   "cavity_time_march": cavity},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic() if (args.n, args.levels) == (100, 5) else None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
