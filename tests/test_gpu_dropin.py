"""The drop-in C surface, end to end.

tests/dropin/heat_driver.c is a heat-conduction driver written against
include/header_files the way the reference's mg_heat_conduction.c:18-86 is:
it reads flow_parameters.csv and grid_filenames.csv, reads Gmsh meshes,
builds the derivative matrices, rewrites boundary rows of the host lap_Poison
arrays from a plug-in before every V-cycle, and drives the V-cycle through
relaxation / calculate_residuals / restrict_residuals / prolongate_corrections.
tests/dropin/ns_driver.c follows mg_NS_solver.c:15-60 the same way: a
lid-driven cavity stepped with time_implicit_solver (timple_solver.h) or,
built with -DFRACTIONAL_STEP, fractional_step_explicit
(fractional_step_solver.h).

The CPU test compiles and links the three builds with gcc against
libmgm.so.  The GPU tests run them on generated meshes and repeat the same
calls through the resident API on the same meshes, with the same boundary
rows, sources and initial fields: every level-0 field the drivers write, and
every step's steady-state error, must agree bit for bit (the drop-in layer
moves host arrays in and out, and nothing else).
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import ROOT
from multigrid_meshless_amd import cases


@pytest.fixture(autouse=True)
def fixed_weight_kernel(monkeypatch):
    """The drivers run in processes of their own, each with its own weight-kernel
    autotuner; the Gauss-Jordan kernel (whose variants give the same bits) is
    pinned here and in the drivers, so both build the same operator rows."""
    monkeypatch.setenv("MGM_LU_KERNEL", "gj")

DRIVERS = os.path.join(ROOT, "tests", "dropin")
BUILDS = {"heat": ("heat_driver.c", []), "timple": ("ns_driver.c", []),
          "fs": ("ns_driver.c", ["-DFRACTIONAL_STEP"])}
LIB = os.path.join(ROOT, "multigrid_meshless_amd")
HAVE_LIB = os.path.isfile(os.path.join(LIB, "libmgm.so"))

# flow_parameters.csv is read positionally (hf/mesh_functions.h:37-100)
PARAM_NAMES = ["domain_dimensions", "poly_deg", "phs_deg", "cloud_size_multiplier", "test_derivative",
               "courant_number", "steady_tolerance", "poisson_solver_tolerance", "sor_parameter",
               "num_vcycles", "num_relax", "num_time_steps", "write_interval", "neumann_flag_boundary",
               "time_step", "iter_momentum", "iter_timple", "Re", "facRe", "facdt"]


def compile_driver(d, build="heat"):
    src, defines = BUILDS[build]
    exe = os.path.join(d, build)
    cmd = ["gcc", "-std=gnu11", "-O2", "-Wall", "-Werror", *defines, f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(DRIVERS, src), "-o", exe, f"-L{LIB}", "-lmgm", f"-Wl,-rpath,{LIB}", "-lm"]
    r = subprocess.run(cmd, cwd=d, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def run_driver(exe, d, order="natural"):
    env = dict(os.environ)
    if order == "colour":
        env["MGM_SWEEP_ORDER"] = "colour"
    else:
        env.pop("MGM_SWEEP_ORDER", None)
    r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def write_case(mgm, d, sizes, seed, omega, num_vcycles, num_relax, **kw):
    values = dict(zip(PARAM_NAMES, [2, 3, 3, 2, 0, 0.3, 1e-8, 1e-8, omega, num_vcycles, num_relax, 1, 10000,
                                    1, 0.1, 5, 1, 10, 1.0, 1.0]))
    values.update(kw)
    with open(os.path.join(d, "flow_parameters.csv"), "w") as f:
        for name in PARAM_NAMES:
            f.write(f"{name},{values[name]!r}\n")
    meshes = [f"L{k}.msh" for k in range(len(sizes))]
    for k, (name, n) in enumerate(zip(meshes, sizes)):
        mgm.Points.generate(2, n, seed + k, 0.2, with_volume_elements=True, gmsh_path=os.path.join(d, name))
    with open(os.path.join(d, "grid_filenames.csv"), "w") as f:
        f.write(f"num_levels,{len(sizes)}\n")
        for name in meshes:
            f.write(name + "\n")
    return [os.path.join(d, m) for m in meshes]


@pytest.mark.skipif(not HAVE_LIB or shutil.which("gcc") is None, reason="libmgm.so or gcc missing")
@pytest.mark.parametrize("build", sorted(BUILDS))
def test_dropin_driver_compiles_and_links(build):
    with tempfile.TemporaryDirectory() as d:
        assert os.path.isfile(compile_driver(d, build))


def resident_problem(mgm, paths, omega, num_relax, order="natural"):
    """The hierarchy read_complete_mesh_data builds from flow_parameters.csv
    (degree 3 polynomials and PHS, multiplier 2; coarse levels degree 3; the
    reference's transfer rows; the natural sweep order, or the multicolour one
    under MGM_SWEEP_ORDER=colour)."""
    cfg = mgm.config(dimension=2, poly_degree=3, coarse_poly_degree=3, phs_degree=3, cloud_size_multiplier=2.0,
                     omega=omega, num_relax=num_relax, transfer_rows=mgm.ROWS_REFERENCE,
                     sweep_order=mgm.ORDER_COLOUR if order == "colour" else mgm.ORDER_NATURAL)
    return mgm.Problem(cfg, [mgm.Points.read_gmsh(p, 2) for p in paths]).setup()


def resident_cycles(mgm, paths, omega, num_vcycles, num_relax, source, order="natural"):
    P = resident_problem(mgm, paths, omega, num_relax, order)
    geo = P.geometry(0)
    A = P.matrix(0, "lap_Poison")
    wall = (geo["btag"] != 0) & (geo["x"] == 0)
    A[wall] = 0.0
    A[wall, 0] = 1.0
    P.set_matrix(0, "lap_Poison", A)
    P.zero_fields()
    P.set_field(0, "source", source)
    L = P.L
    for _ in range(num_vcycles):
        for lev in range(L):
            P.relax(lev)
            P.residual(lev)
            if lev != L - 1:
                P.restrict(lev)
        for lev in range(L - 1, 0, -1):
            P.prolong(lev)
            if lev != 1:
                P.relax(lev - 1)
    return P, geo, wall


@pytest.mark.gpu
@pytest.mark.parametrize("order", ["natural", "colour"])
@pytest.mark.parametrize("sizes,seed,num_vcycles,num_relax", [((33, 17, 9), 5, 6, 3),
                                                              ((65, 33, 17, 9), 11, 4, 2)])
def test_dropin_heat_driver_matches_resident_api(gpu, sizes, seed, num_vcycles, num_relax, order):
    mgm = gpu
    omega = 1.6
    with tempfile.TemporaryDirectory() as d:
        exe = compile_driver(d)
        paths = write_case(mgm, d, sizes, seed, omega, num_vcycles, num_relax)
        print(f"[dropin] heat driver, levels {sizes}, seed {seed}, {order} order", flush=True)
        stdout = run_driver(exe, d, order)
        assert stdout.count("Pressure Residual") == num_vcycles * (2 * len(sizes) - 2)
        out = np.fromfile(os.path.join(d, "out.bin"), dtype=np.float64).reshape(4, -1)
        x, y, source, p = out
        P, geo, wall = resident_cycles(mgm, paths, omega, num_vcycles, num_relax, source, order)

    np.testing.assert_array_equal(x, geo["x"])
    np.testing.assert_array_equal(y, geo["y"])
    assert wall.any() and np.all(source[wall] == 1.0)
    assert np.all(np.isfinite(p)) and np.abs(p).max() > 0
    np.testing.assert_array_equal(p, P.field(0, "p"))


NS_FIELDS = ("u", "v", "p", "u_new", "v_new", "p_old", "pprime")


@pytest.mark.gpu
@pytest.mark.parametrize("order", ["natural", "colour"])
@pytest.mark.parametrize("scheme", ["timple", "fs"])
def test_dropin_ns_driver_matches_resident_api(gpu, scheme, order):
    mgm = gpu
    sizes, seed, omega, num_relax, num_vcycles, steps = (33, 17, 9), 23, 1.2, 3, 4, 3
    ns = dict(Re=100, courant_number=0.5, iter_momentum=3, iter_timple=2)
    with tempfile.TemporaryDirectory() as d:
        exe = compile_driver(d, scheme)
        paths = write_case(mgm, d, sizes, seed, omega, num_vcycles, num_relax, num_time_steps=steps, **ns)
        print(f"[dropin] ns driver {scheme}, levels {sizes}, seed {seed}, {steps} steps, {order} order", flush=True)
        stdout = run_driver(exe, d, order)
        assert stdout.count("Steady state error") == steps
        rec = np.fromfile(os.path.join(d, "steps.bin"), dtype=np.float64)
        out = np.fromfile(os.path.join(d, "out.bin"), dtype=np.float64).reshape(9, -1)
        P = resident_problem(mgm, paths, omega, num_relax, order)

    dt, errors = rec[0], rec[1:]
    geo = P.geometry(0)
    np.testing.assert_array_equal(out[0], geo["x"])
    np.testing.assert_array_equal(out[1], geo["y"])
    # dt from calculate_dt (hf/general_functions.h:34-41) on the level-0 spacing
    assert dt == cases.calculate_dt(P.info(0)["d_avg"], 1.0 / ns["Re"], ns["courant_number"])
    lid = np.where(np.abs(geo["y"] - 1.0) < 1e-9, 1.0, 0.0)
    P.set_field(0, "u", lid)
    if scheme == "timple":
        P.set_field(0, "u_new", lid)
    mu = 1.0 / ns["Re"]
    P.ns_set_params(dt=dt, rho=1.0, mu=mu, nu=mu, facRe=1.0, facdt=1.0, field_rho=1.0,
                    iter_momentum=ns["iter_momentum"], iter_timple=ns["iter_timple"],
                    num_vcycles=num_vcycles)
    mine = [P.timple_step() if scheme == "timple" else P.fs_step() for _ in range(steps)]
    assert np.all(np.isfinite(errors)) and min(errors) > 0
    np.testing.assert_array_equal(errors, mine)
    for k, name in enumerate(NS_FIELDS):
        np.testing.assert_array_equal(out[2 + k], P.field(0, name), err_msg=name)
