"""CPU-only checks of libmgm: ABI exports, host I/O, generator (no GPU calls)."""
import os
import re
import tempfile

import numpy as np
import pytest

from conftest import REFERENCE, ROOT


def declared_functions():
    """Every non-inline function declared under include/ (mgm.h and the drop-in headers)."""
    names = set()
    inc = os.path.join(ROOT, "include")
    for dirpath, _, files in os.walk(inc):
        for fn in files:
            if not fn.endswith(".h"):
                continue
            text = open(os.path.join(dirpath, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            text = re.sub(r"static inline[^{]*\{[^}]*\}", "", text)
            for m in re.finditer(r"^\s*(?:const\s+|struct\s+)?[A-Za-z_0-9]+\s*\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                                 text, re.M):
                if m.group(1) not in ("if", "while", "for", "return", "sizeof"):
                    names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from multigrid_meshless_amd._lib import SIGNATURES, lib
    L = lib()
    decl = declared_functions()
    assert len(decl) > 60
    for name in decl:
        assert hasattr(L, name), f"{name} declared in include/ but not exported"
    bound = {s[0] for s in SIGNATURES}
    mgm_decl = {n for n in decl if n.startswith("mgm_") and not n.startswith("mgm_dropin_")}
    assert mgm_decl == bound, mgm_decl ^ bound


def test_generator_roundtrips_through_both_readers(oracle):
    """The generator's in-memory result equals what the Gmsh reader (ours and the
    oracle's restatement of the reference reader) reads back from its file."""
    from multigrid_meshless_amd import Points
    for dim, n in ((2, 17), (3, 8)):
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "c.msh")
            P = Points.generate(dim, n, 42, 0.2, with_volume_elements=True, gmsh_path=path)
            a, ia = P.arrays(), P.info()
            b = Points.read_gmsh(path, dim)
            bb, ib = b.arrays(), b.info()
            o = oracle.read_gmsh(path, dim)
            for k in ("x", "y", "z", "nx", "ny", "nz", "btag", "ctag"):
                np.testing.assert_array_equal(a[k], bb[k], err_msg=k)
                np.testing.assert_array_equal(a[k], o[k], err_msg=k)
            assert ia == ib
            assert (o["N"], o["nb"], o["num_elem"], o["d_avg"]) == (ia["N"], ia["nb"], ia["num_elem"], ia["d_avg"])
            if dim == 3:
                assert ia["N"] == (n - 2) ** 3 + 6 * (n - 2) ** 2 and ia["nb"] == 6 * (n - 2) ** 2
            else:
                assert ia["N"] == (n - 2) ** 2 + 4 * (n - 2) and ia["nb"] == 4 * (n - 2)
            # boundary nodes first, then interior (the reader's index convention)
            assert a["btag"][:ia["nb"]].all() and not a["btag"][ia["nb"]:].any()


@pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "mesh")), reason="reference meshes absent")
@pytest.mark.parametrize("rel,dim", [("mesh/2d/SQ/n_2.msh", 2), ("mesh/2d/HOLE_SQ/n_1.msh", 2),
                                     ("mesh/3d/CUBE/n_1.msh", 3), ("mesh/3d/SP_in_CUBE/n_1.msh", 3)])
def test_reader_matches_oracle_on_reference_meshes(rel, dim, oracle):
    from multigrid_meshless_amd import Points
    path = os.path.join(REFERENCE, rel)
    a = Points.read_gmsh(path, dim).arrays()
    o = oracle.read_gmsh(path, dim)
    for k in ("x", "y", "z", "nx", "ny", "nz", "btag", "ctag"):
        np.testing.assert_array_equal(a[k], o[k], err_msg=k)


def test_reader_rejects_malformed_files():
    from multigrid_meshless_amd import MgmError, Points
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "bad.msh")
        open(p, "w").write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n3\n1 0 0\n")
        with pytest.raises(MgmError):
            Points.read_gmsh(p, 2)
        with pytest.raises(MgmError):
            Points.read_gmsh(os.path.join(d, "missing.msh"), 2)


def test_poly_terms_closed_form():
    from multigrid_meshless_amd._lib import lib
    for dim in (2, 3):
        for deg in range(0, 6):
            n = lib().mgm_poly_terms(dim, deg, None, None, None)
            exp = (deg + 1) * (deg + 2) // 2 if dim == 2 else (deg + 1) * (deg + 2) * (deg + 3) // 6
            assert n == exp


def test_splitmix_matches_oracle(oracle):
    import ctypes as C
    from multigrid_meshless_amd._lib import lib
    s = C.c_uint64(123)
    ours = [lib().mgm_splitmix64(C.byref(s)) for _ in range(5)]
    assert ours == [int(v) for v in oracle.splitmix64(123, 5)]


def test_setup_without_gpu_fails_loudly():
    """With no device initialised the GPU entry points return an error (no fallback)."""
    import multigrid_meshless_amd as mgm
    if mgm.device_count() > 0:
        pytest.skip("a GPU is visible")
    pts = mgm.Points.generate(2, 10, 0)
    P = mgm.Problem(mgm.config(), [pts])
    with pytest.raises(mgm.MgmError):
        P.setup()


def test_annulus_cloud_geometry():
    """cases.annulus: boundary nodes first, on the two circles, with unit normals
    pointing out of the domain; interior nodes strictly between the circles."""
    from multigrid_meshless_amd import cases
    P = cases.annulus(91, 5)
    info, a = P.info(), P.arrays()
    nb, N = info["nb"], info["N"]
    assert np.all(a["btag"][:nb] == 1) and np.all(a["btag"][nb:] == 0) and not a["ctag"].any()
    r = np.hypot(a["x"], a["y"])
    outer = np.abs(r[:nb] - 1.0) < 1e-12
    inner = np.abs(r[:nb] - 0.5) < 1e-12
    assert np.all(outer | inner) and outer.sum() > inner.sum() > 0
    n = np.stack([a["nx"][:nb], a["ny"][:nb]], 1)
    radial = np.stack([a["x"][:nb], a["y"][:nb]], 1) / r[:nb, None]
    np.testing.assert_allclose(np.hypot(n[:, 0], n[:, 1]), 1.0, atol=1e-14)
    np.testing.assert_allclose(n[outer], radial[outer], atol=1e-14)
    np.testing.assert_allclose(n[inner], -radial[inner], atol=1e-14)
    h = 2.0 / 90
    assert np.all(r[nb:] > 0.5 + 0.5 * h) and np.all(r[nb:] < 1.0 - 0.5 * h)
    assert info["d_avg"] == h and N - nb > 10 * nb
    # the same seed gives the same cloud, another seed another one
    np.testing.assert_array_equal(cases.annulus(91, 5).arrays()["x"], a["x"])
    assert not np.array_equal(cases.annulus(91, 6).arrays()["x"], a["x"])


def test_write_history_formats(tmp_path):
    """The convergence file in the drivers' two formats ("%d, %e" and "%d, %lf")."""
    import multigrid_meshless_amd as mgm
    v = [6.398105703570e-03, 3.2277e-05, 0.0, 1.5]
    for sci, fmt in ((True, "%d, %e\n"), (False, "%d, %f\n")):
        p = tmp_path / f"conv{int(sci)}.csv"
        mgm.write_history(str(p), v, scientific=sci)
        assert p.read_text() == "".join(fmt % (i, x) for i, x in enumerate(v))
    with pytest.raises(mgm.MgmError):
        mgm.write_history(str(tmp_path / "missing" / "conv.csv"), v)


def test_c_abi_rejects_bad_arguments(tmp_path):
    """Every entry point checks its arguments before any device work: bad
    configurations, calls on a hierarchy that is not set up, out-of-range
    levels and fields come back as MgmError with the library's message."""
    import ctypes as C

    import multigrid_meshless_amd as mgm
    from multigrid_meshless_amd._lib import lib
    pts = mgm.Points.generate(2, 12, 0)
    for bad in (dict(dimension=4), dict(phs_degree=0), dict(poly_degree=0), dict(coarse_poly_degree=0)):
        with pytest.raises(mgm.MgmError):
            mgm.Problem(mgm.config(**bad), [pts])
    with pytest.raises(mgm.MgmError):
        mgm.Problem(mgm.config(), [])
    with pytest.raises(KeyError):
        mgm.config(no_such_field=1)
    P = mgm.Problem(mgm.config(), [pts])          # created, never set up
    calls = [lambda: P.relax(0), lambda: P.residual(0), lambda: P.vcycle(), lambda: P.solve(1e-8, 5),
             lambda: P.field(0, "p"), lambda: P.set_field(0, "p", np.zeros(pts.info()["N"])),
             lambda: P.info(0), lambda: P.matrix(0, "lap"), lambda: P.fs_step(),
             lambda: P.write_solution(str(tmp_path / "s.csv"))]
    for call in calls:
        with pytest.raises(mgm.MgmError, match="not set up|level out of range"):
            call()
    assert not (tmp_path / "s.csv").exists()
    h = C.c_void_p()
    z = np.zeros(4)
    t = np.zeros(4, np.uint8)
    assert lib().mgm_points_from_arrays(4, 5, 0, 0.1, *[z.ctypes.data] * 6, t.ctypes.data, t.ctypes.data,
                                        C.byref(h)) != 0
    assert b"bad sizes" in lib().mgm_last_error()
    assert lib().mgm_set_smoother(P._h, 2.5, 3) != 0 and b"omega" in lib().mgm_last_error()
    assert lib().mgm_set_smoother(P._h, 1.0, -1) != 0
