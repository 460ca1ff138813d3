"""Write golden fixtures from the compiled reference.

Run in a container that has the reference tree (default /root/reference):

    make -C oracle ref            # builds oracle/_ref/libref_{heat,timple,fs}.so
    python tests/golden/make_golden.py

Each case writes tests/golden/<case>.npz holding the inputs (meshes are
described by the arrays the reference's reader produced) and the reference's
outputs: post-reordering geometry, stencils, transfer points, weights,
transfer rows and fields after V-cycles.  The reference runs in a subprocess
per case (it prints heavily and changes directory).  Generated clouds are
written with libmgm's generator (host code, no GPU needed), which is itself
under test: the reference reading them is the check.
"""
from __future__ import annotations

import ctypes as C
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("MGM_REFERENCE", "/root/reference")
sys.path.insert(0, ROOT)

FLOW_KEYS = ["domain_dimensions", "poly_deg", "phs_deg", "cloud_size_multiplier", "test_derivative",
             "courant_number", "steady_tolerance", "poisson_solver_tolerance", "sor_parameter",
             "num_vcycles", "num_relax", "num_time_steps", "write_interval",
             "neumann_flag_boundary", "time_step", "iter_momentum", "iter_timple", "Re", "facRe",
             "facdt"]


def flow_csv(dim, poly, phs, mult, omega=1.0, num_vcycles=2, num_relax=3, Re=100.0, dt=0.01):
    vals = [dim, poly, phs, mult, 0, 0.3, 1e-8, 1e-8, omega, num_vcycles, num_relax, 2, 10000, 1,
            dt, 2, 1, Re, 1.0, 1.0]
    return "".join(f"{k},{v}\n" for k, v in zip(FLOW_KEYS, vals))


# name: (variant lib, dim, poly, phs, mult, omega, meshes, vcycles, extra)
def cases():
    return {
        "sq2d_ref": dict(lib="heat", dim=2, poly=3, phs=3, mult=2, omega=1.0,
                         meshes=[("ref", "mesh/2d/SQ/n_2.msh"), ("ref", "mesh/2d/SQ/n_1.msh")],
                         vcycles=3),
        "gen2d": dict(lib="heat", dim=2, poly=3, phs=3, mult=2, omega=1.0,
                      meshes=[("gen", (2, 24, 0)), ("gen", (2, 12, 1)), ("gen", (2, 8, 2))],
                      vcycles=3),
        "gen2d_phs5": dict(lib="heat", dim=2, poly=4, phs=5, mult=2, omega=1.0,
                           meshes=[("gen", (2, 20, 7)), ("gen", (2, 10, 8))], vcycles=2),
        "cube3d_ref": dict(lib="heat", dim=3, poly=3, phs=3, mult=2, omega=1.0,
                           meshes=[("ref", "mesh/3d/CUBE/n_1.msh")], vcycles=0),
        "gen3d": dict(lib="heat", dim=3, poly=3, phs=3, mult=2, omega=1.0,
                      meshes=[("gen", (3, 10, 1)), ("gen", (3, 6, 2))], vcycles=2),
        "timple2d": dict(lib="timple", dim=2, poly=3, phs=3, mult=2, omega=1.0,
                         meshes=[("gen", (2, 16, 3)), ("gen", (2, 8, 4))], vcycles=2),
        "fs2d": dict(lib="fs", dim=2, poly=3, phs=3, mult=2, omega=1.0,
                     meshes=[("gen", (2, 16, 5)), ("gen", (2, 8, 6))], vcycles=2),
    }


def prepare(case: dict, workdir: str):
    from multigrid_meshless_amd import Points
    names = []
    for k, (kind, spec) in enumerate(case["meshes"]):
        name = f"l{k}.msh"
        if kind == "ref":
            shutil.copy(os.path.join(REF, spec), os.path.join(workdir, name))
        else:
            dim, n, seed = spec
            Points.generate(dim, n, seed, 0.2, with_volume_elements=True,
                            gmsh_path=os.path.join(workdir, name))
        names.append(name)
    with open(os.path.join(workdir, "flow_parameters.csv"), "w") as f:
        f.write(flow_csv(case["dim"], case["poly"], case["phs"], case["mult"], case["omega"],
                         num_vcycles=case["vcycles"] or 1))
    with open(os.path.join(workdir, "grid_filenames.csv"), "w") as f:
        f.write(f"num_levels,{len(names)}\n" + "".join(n + "\n" for n in names))


def run_case(name: str, workdir: str, out: str):
    """Child process: drive the reference harness and dump its state."""
    case = cases()[name]
    lib = C.CDLL(os.path.join(ROOT, "oracle", "_ref", f"libref_{case['lib']}.so"))
    # silence the reference's stdout
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    os.dup2(devnull, 1)
    try:
        L = lib.ref_setup(workdir.encode(), 1)
        res = {"num_levels": np.int32(L)}
        info = (C.c_int * 6)()
        dav = C.c_double()
        dims = []
        for lev in range(L):
            lib.ref_info(lev, info, C.byref(dav))
            N, nb, m, n, poly, dim = list(info)
            dims.append((N, m))
            res[f"L{lev}_info"] = np.array([N, nb, m, n, poly, dim], np.int32)
            res[f"L{lev}_d_avg"] = np.float64(dav.value)
            g = {k: np.zeros(N) for k in ("x", "y", "z", "nx", "ny", "nz")}
            bt = np.zeros(N, np.uint8)
            lib.ref_geometry(lev, *[g[k].ctypes.data_as(C.c_void_p) for k in ("x", "y", "z", "nx", "ny", "nz")],
                             bt.ctypes.data_as(C.c_void_p))
            for k, v in g.items():
                res[f"L{lev}_{k}"] = v
            res[f"L{lev}_btag"] = bt
            cl = np.zeros((N, m), np.int32)
            lib.ref_cloud(lev, cl.ctypes.data_as(C.c_void_p))
            res[f"L{lev}_cloud"] = cl
            for which, key in ((0, "rp"), (1, "pp")):
                a = np.zeros(N, np.int32)
                lib.ref_points(lev, which, a.ctypes.data_as(C.c_void_p))
                res[f"L{lev}_{key}"] = a
            for which, key in ((0, "Dx"), (1, "Dy"), (2, "Dz"), (3, "lap"), (4, "lap_Poison")):
                if key == "Dz" and dim != 3:
                    continue
                a = np.zeros((N, m))
                lib.ref_matrix(lev, which, m, a.ctypes.data_as(C.c_void_p))
                res[f"L{lev}_{key}"] = a
        for lev in range(L):
            N, m = dims[lev]
            if lev > 0:
                a = np.zeros((N, dims[lev - 1][1]))
                lib.ref_matrix(lev, 5, dims[lev - 1][1], a.ctypes.data_as(C.c_void_p))
                res[f"L{lev}_restr"] = a
            if lev < L - 1:
                a = np.zeros((N, dims[lev + 1][1]))
                lib.ref_matrix(lev, 6, dims[lev + 1][1], a.ctypes.data_as(C.c_void_p))
                res[f"L{lev}_prol"] = a

        def fields(tag, which):
            for lev in range(L):
                for w, key in which:
                    a = np.zeros(dims[lev][0])
                    lib.ref_get_field(lev, w, a.ctypes.data_as(C.c_void_p))
                    res[f"{tag}_L{lev}_{key}"] = a

        nv = case["vcycles"]
        if case["lib"] == "heat" and nv:
            # mg_heat_conduction.c:45-76 with num_time_steps = 1
            lib.ref_heat_init(1)
            fields("init", [(0, "p"), (1, "source")])
            for c in range(nv):
                lib.ref_heat_set_boundary_rhs(1)
                if c == 0:
                    a = np.zeros(dims[0])
                    lib.ref_matrix(0, 4, dims[0][1], a.ctypes.data_as(C.c_void_p))
                    res["bc_L0_lap_Poison"] = a
                lib.ref_vcycle()
                fields(f"cyc{c}", [(0, "p"), (1, "source"), (2, "res")])
        elif case["lib"] in ("timple", "fs") and nv:
            # multigrid_Poisson_solver on a seeded random source, num_vcycles cycles
            rng = np.random.default_rng(11)
            pf = 3 if case["lib"] == "timple" else 0
            for lev in range(L):
                lib.ref_set_field(lev, pf, np.zeros(dims[lev][0]).ctypes.data_as(C.c_void_p))
                lib.ref_set_field(lev, 1, np.zeros(dims[lev][0]).ctypes.data_as(C.c_void_p))
                lib.ref_set_field(lev, 2, np.zeros(dims[lev][0]).ctypes.data_as(C.c_void_p))
            src = rng.standard_normal(dims[0][0])
            lib.ref_set_field(0, 1, src.ctypes.data_as(C.c_void_p))
            res["mg_source"] = src
            lib.ref_set_omega_relax(C.c_float(case["omega"]), 3, nv)
            lib.ref_mg_poisson()
            fields("mg", [(pf, "p"), (1, "source"), (2, "res")])
            # whole time steps from the LDC initial state (mg_NS_solver.c:51-72):
            # every field of every level before the first step, level-0 fields
            # and the steady-state error after each of two steps
            lib.ref_ns_init()
            if case["lib"] == "timple":
                # time_implicit_solver starts from u_new/v_new, which the LDC
                # initialisation leaves at zero (the step would then be a no-op):
                # start from the lid velocity plus seeded interior perturbations
                rng = np.random.default_rng(13)
                N0 = dims[0][0]
                inner = res["L0_btag"] == 0
                u = np.zeros(N0)
                lib.ref_get_field(0, 4, u.ctypes.data_as(C.c_void_p))
                un = u.copy()
                un[inner] += 0.1 * rng.standard_normal(inner.sum())
                vn = np.zeros(N0)
                vn[inner] = 0.1 * rng.standard_normal(inner.sum())
                po = 0.1 * rng.standard_normal(N0)
                for fid, a in ((7, un), (8, vn), (10, po)):
                    lib.ref_set_field(0, fid, a.ctypes.data_as(C.c_void_p))
            fields("st_init", [(k, str(k)) for k in range(19)])
            d, iv, om = (C.c_double * 7)(), (C.c_int * 4)(), C.c_float()
            lib.ref_ns_params(d, iv, C.byref(om))
            res["st_params_d"] = np.array(list(d))
            res["st_params_i"] = np.array(list(iv), np.int32)
            res["st_omega"] = np.float32(om.value)
            lib.ref_step.restype = C.c_double
            for k in range(2):
                err = lib.ref_step()
                res[f"st{k}_err"] = np.float64(err)
                for fid in range(19):
                    a = np.zeros(dims[0][0])
                    lib.ref_get_field(0, fid, a.ctypes.data_as(C.c_void_p))
                    res[f"st{k}_L0_{fid}"] = a
    finally:
        os.dup2(saved, 1)
        os.close(devnull)
    np.savez_compressed(out, **res)


def main():
    only = sys.argv[1:]
    for name, case in cases().items():
        if only and name not in only:
            continue
        with tempfile.TemporaryDirectory(prefix="mgm_golden_") as wd:
            prepare(case, wd)
            out = os.path.join(HERE, f"{name}.npz")
            # inputs for the GPU side: the meshes as the reader leaves them
            from multigrid_meshless_amd import Points
            inputs = {}
            for k in range(len(case["meshes"])):
                a = Points.read_gmsh(os.path.join(wd, f"l{k}.msh"), case["dim"]).arrays()
                inf = Points.read_gmsh(os.path.join(wd, f"l{k}.msh"), case["dim"]).info()
                for key, v in a.items():
                    inputs[f"in{k}_{key}"] = v
                inputs[f"in{k}_meta"] = np.array([inf["N"], inf["nb"], inf["num_elem"]], np.int64)
                inputs[f"in{k}_d_avg"] = np.float64(inf["d_avg"])
            r = subprocess.run([sys.executable, __file__, "--child", name, wd, out + ".tmp.npz"],
                               capture_output=True, text=True)
            if r.returncode != 0:
                print(r.stdout[-2000:], r.stderr[-2000:])
                raise SystemExit(f"case {name} failed")
            ref = dict(np.load(out + ".tmp.npz"))
            os.remove(out + ".tmp.npz")
            meta = np.array([case["dim"], case["poly"], case["phs"], case["mult"], case["vcycles"]],
                            np.float64)
            np.savez_compressed(out, **ref, **inputs, case_meta=meta,
                                case_omega=np.float32(case["omega"]),
                                case_lib=np.array(case["lib"]))
            print(f"{name}: {os.path.getsize(out) / 1e6:.2f} MB")


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--child":
        run_case(sys.argv[2], sys.argv[3], sys.argv[4])
    else:
        main()
