"""ctypes binding of the CPU parity oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke() entry point as the
checker and by bench.py's cpu_baseline leg.  The product package
(multigrid_meshless_amd) never imports it.

Each wrapper names the reference function it restates; the C source cites the
reference file:line.  Arrays are numpy, row-major, reference index order.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

VAR_H, VAR_T, VAR_F = 0, 1, 2

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_sp = np.ctypeslib.ndpointer(dtype=np.int16, flags="C_CONTIGUOUS")


class _Mesh(C.Structure):
    _fields_ = [("N", C.c_int), ("nb", C.c_int), ("num_elem", C.c_int),
                ("num_corners_removed", C.c_int), ("d_avg", C.c_double)] + \
               [(k, C.POINTER(C.c_double)) for k in ("x", "y", "z", "nx", "ny", "nz")] + \
               [("btag", C.POINTER(C.c_uint8)), ("ctag", C.POINTER(C.c_uint8))]


class Level(C.Structure):
    """orc_level: one multigrid level's arrays for the V-cycle primitives."""
    _fields_ = [("N", C.c_int), ("m", C.c_int),
                ("cloud", C.c_void_p), ("Lp", C.c_void_p), ("lap", C.c_void_p),
                ("btag", C.c_void_p), ("order", C.c_void_p),
                ("p", C.c_void_p), ("s", C.c_void_p), ("res", C.c_void_p),
                ("rp", C.c_void_p), ("R", C.c_void_p), ("pp", C.c_void_p), ("P", C.c_void_p),
                ("cls", C.c_void_p), ("ncls", C.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built: run `make -C oracle`")
        L = C.CDLL(path)
        L.orc_read_gmsh.argtypes = [C.c_char_p, C.c_int, C.POINTER(_Mesh)]
        L.orc_free_mesh.argtypes = [C.POINTER(_Mesh)]
        L.orc_poly_terms.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_find_cloud_index.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _u8, _u8,
                                           C.c_double, C.c_int, _ip]
        L.orc_find_nearest_point.argtypes = [C.c_int, _dp, _dp, _dp, _u8,
                                             C.c_int, _dp, _dp, _dp, _u8, C.c_double, _ip]
        L.orc_rcm.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _u8, _ip, _ip]
        L.orc_correct_normals.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _u8, _ip]
        L.orc_avg_dx.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _u8, _ip]
        L.orc_avg_dx.restype = C.c_double
        L.orc_weights.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _sp, _sp, _sp,
                                  _dp, _dp, _dp, C.c_void_p, _ip, C.c_int,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_lap_poisson.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_void_p, _dp,
                                      _dp, _dp, _dp, _u8, C.c_void_p, _dp]
        L.orc_restriction.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _sp, _sp, _sp,
                                      _dp, _dp, _dp, _ip, C.c_int, C.c_int, _dp, _dp, _dp, _u8,
                                      _ip, C.c_int, C.c_int, _dp]
        L.orc_prolongation.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _u8, _ip, C.c_int,
                                       C.c_int, C.c_int, C.c_int, C.c_int, _sp, _sp, _sp,
                                       _dp, _dp, _dp, _ip, C.c_int, C.c_int, _dp]
        for fn in ("orc_relax",):
            getattr(L, fn).argtypes = [C.c_int, C.POINTER(Level), C.c_float, C.c_int]
        L.orc_residual.argtypes = [C.c_int, C.POINTER(Level)]
        L.orc_restrict.argtypes = [C.c_int, C.POINTER(Level), C.POINTER(Level)]
        L.orc_prolong.argtypes = [C.c_int, C.POINTER(Level), C.POINTER(Level)]
        L.orc_vcycle.argtypes = [C.c_int, C.c_int, C.POINTER(Level), C.c_float, C.c_int]
        L.orc_vcycle_levels.argtypes = [C.c_int, C.c_int, C.POINTER(Level), C.c_void_p, C.c_int]
        L.orc_vcycle_smooth.argtypes = [C.c_int, C.POINTER(Level), C.c_void_p, C.c_void_p, C.c_int]
        L.orc_l2_norm_gen.argtypes = [C.c_int, _u8, _dp, C.c_void_p]
        L.orc_l2_norm_gen.restype = C.c_double
        L.orc_spmv.argtypes = [C.c_int, C.c_int, _dp, _ip, _dp, _dp]
        L.orc_greedy_colour.argtypes = [C.c_int, C.c_int, _ip, _ip]
        L.orc_splitmix64.argtypes = [C.POINTER(C.c_uint64)]
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_num_threads.restype = C.c_int
        _LIB = L
    return _LIB


def _np(ptr, n, dt):
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)


def read_gmsh(path: str, dim: int) -> dict:
    """read_PointStructure (hf/mesh_functions.h:172-476)."""
    m = _Mesh()
    rc = lib().orc_read_gmsh(path.encode(), dim, C.byref(m))
    if rc:
        raise IOError(f"orc_read_gmsh({path}) failed: {rc}")
    N = m.N
    out = dict(N=N, nb=m.nb, num_elem=m.num_elem, d_avg=m.d_avg,
               num_corners_removed=m.num_corners_removed)
    for k in ("x", "y", "z", "nx", "ny", "nz"):
        out[k] = _np(getattr(m, k), N, np.float64)
    out["btag"] = _np(m.btag, N, np.uint8)
    out["ctag"] = _np(m.ctag, N, np.uint8)
    lib().orc_free_mesh(C.byref(m))
    return out


def poly_terms(dim: int, deg: int):
    """calculate_parameters monomial table (hf/mesh_functions.h:131-170)."""
    n = lib().orc_poly_terms(dim, deg, None, None, None)
    px, py, pz = (np.zeros(n, np.int16) for _ in range(3))
    lib().orc_poly_terms(dim, deg, px.ctypes.data, py.ctypes.data, pz.ctypes.data)
    return px, py, pz


def find_cloud_index(x, y, z, btag, ctag, nb, d_avg, m):
    N = len(x)
    cloud = np.zeros(N * m, np.int32)
    lib().orc_find_cloud_index(N, nb, x, y, z, btag, ctag, d_avg, m, cloud)
    return cloud.reshape(N, m)


def find_nearest_point(a: dict, b: dict, d_avg_b: float):
    out = np.zeros(len(a["x"]), np.int32)
    lib().orc_find_nearest_point(len(a["x"]), a["x"], a["y"], a["z"], a["ctag"],
                                 len(b["x"]), b["x"], b["y"], b["z"], b["ctag"], d_avg_b, out)
    return out


def rcm(lv: dict):
    """rcm_reordering_with_boundarynodes, in place on lv; returns queue1."""
    N, m = lv["cloud"].shape
    q1 = np.zeros(N, np.int32)
    cl = np.ascontiguousarray(lv["cloud"].reshape(-1))
    rc = lib().orc_rcm(N, m, lv["x"], lv["y"], lv["z"], lv["nx"], lv["ny"], lv["nz"],
                       lv["btag"], cl, q1)
    if rc:
        raise RuntimeError("BFS did not reach every node")
    lv["cloud"] = cl.reshape(N, m)
    return q1


def correct_normals(lv: dict):
    N, m = lv["cloud"].shape
    lib().orc_correct_normals(N, m, lv["x"], lv["y"], lv["z"], lv["nx"], lv["ny"], lv["nz"],
                              lv["btag"], np.ascontiguousarray(lv["cloud"].reshape(-1)))


def avg_dx(lv: dict, dim: int) -> float:
    N, m = lv["cloud"].shape
    return lib().orc_avg_dx(N, m, dim, lv["x"], lv["y"], lv["z"], lv["btag"],
                            np.ascontiguousarray(lv["cloud"].reshape(-1)))


def weights(lv: dict, dim: int, phs: int, poly: int, method: int = 0):
    """create_derivative_matrices (hf/rbf.h:401-421); returns Dx, Dy, Dz, lap (N x m)."""
    N, m = lv["cloud"].shape
    px, py, pz = poly_terms(dim, poly)
    n = len(px)
    outs = [np.zeros(N * m) for _ in range(4)]
    ctag = lv.get("ctag")
    rc = lib().orc_weights(N, m, n, dim, phs, px, py, pz, lv["x"], lv["y"], lv["z"],
                           None if ctag is None else ctag.ctypes.data,
                           np.ascontiguousarray(lv["cloud"].reshape(-1)), method,
                           outs[0].ctypes.data, outs[1].ctypes.data,
                           outs[2].ctypes.data if dim == 3 else None, outs[3].ctypes.data)
    if rc:
        raise RuntimeError("weights failed")
    return [o.reshape(N, m) for o in outs]


def lap_poisson(lv: dict, dim: int, Dx, Dy, Dz, lap):
    N, m = lap.shape
    out = np.zeros(N * m)
    lib().orc_lap_poisson(N, m, dim, np.ascontiguousarray(Dx.reshape(-1)),
                          np.ascontiguousarray(Dy.reshape(-1)),
                          np.ascontiguousarray(Dz.reshape(-1)).ctypes.data if dim == 3 else None,
                          np.ascontiguousarray(lap.reshape(-1)), lv["nx"], lv["ny"], lv["nz"],
                          lv["btag"], None, out)
    return out.reshape(N, m)


def restriction(f: dict, c: dict, dim: int, phs: int, poly_f: int, row_mode=0, method=0):
    """create_restriction_matrix (hf/mesh_functions.h:533-600): rows on c, width m_f."""
    Nf, mf = f["cloud"].shape
    px, py, pz = poly_terms(dim, poly_f)
    R = np.zeros(len(c["x"]) * mf)
    lib().orc_restriction(Nf, mf, len(px), dim, phs, px, py, pz, f["x"], f["y"], f["z"],
                          np.ascontiguousarray(f["cloud"].reshape(-1)), len(c["x"]), c["nb"],
                          c["x"], c["y"], c["z"], c["btag"], c["rp"], row_mode, method, R)
    return R.reshape(len(c["x"]), mf)


def prolongation(f: dict, c: dict, dim: int, phs: int, poly_c: int, row_mode=0, method=0):
    """create_prolongation_matrix (hf/mesh_functions.h:602-670): rows on f, width m_c."""
    Nc, mc = c["cloud"].shape
    px, py, pz = poly_terms(dim, poly_c)
    P = np.zeros(len(f["x"]) * mc)
    lib().orc_prolongation(len(f["x"]), f["nb"], f["x"], f["y"], f["z"], f["btag"], f["pp"],
                           Nc, mc, len(px), dim, phs, px, py, pz, c["x"], c["y"], c["z"],
                           np.ascontiguousarray(c["cloud"].reshape(-1)), row_mode, method, P)
    return P.reshape(len(f["x"]), mc)


def l2_norm_gen(btag, a, b=None) -> float:
    """l2_norm_gen (hf/general_functions.h:181-192)."""
    return lib().orc_l2_norm_gen(len(a), btag, np.ascontiguousarray(a, dtype=np.float64),
                                 None if b is None else np.ascontiguousarray(b).ctypes.data)


def spmv(W, cloud, f):
    N, m = W.shape
    out = np.zeros(N)
    lib().orc_spmv(N, m, np.ascontiguousarray(W.reshape(-1)),
                   np.ascontiguousarray(cloud.reshape(-1), dtype=np.int32),
                   np.ascontiguousarray(f, dtype=np.float64), out)
    return out


def greedy_colour(cloud):
    N, m = cloud.shape
    col = np.zeros(N, np.int32)
    nc = lib().orc_greedy_colour(N, m, np.ascontiguousarray(cloud.reshape(-1), dtype=np.int32), col)
    return nc, col


def splitmix64(seed: int, n: int):
    s = C.c_uint64(seed)
    return np.array([lib().orc_splitmix64(C.byref(s)) for _ in range(n)], dtype=np.uint64)


def num_threads() -> int:
    return lib().orc_num_threads()


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(n)


class Hierarchy:
    """Holds numpy arrays for each level and exposes the V-cycle primitives.

    levels: list of dicts with keys cloud (N,m) int32, Lp (N,m), lap (N,m) optional,
    btag (N,) uint8, and optionally order (N,) int32 sweep order, classes
    (ncls+1,) int32 offsets of the colour classes in order (their rows are then
    swept in parallel), rp/R (coarse side), pp/P (fine side).  p/s/res are
    created zero if absent.
    """

    def __init__(self, levels):
        self.lv = levels
        self._keep = []
        arr = (Level * len(levels))()
        for k, d in enumerate(levels):
            N, m = d["cloud"].shape
            for key in ("p", "s", "res"):
                d.setdefault(key, np.zeros(N))
            L = arr[k]
            L.N, L.m = N, m

            def ptr(key, dt):
                if key not in d or d[key] is None:
                    return None
                a = np.ascontiguousarray(d[key], dtype=dt)
                if a is not d[key]:
                    d[key] = a
                self._keep.append(a)
                return a.ctypes.data

            L.cloud = ptr("cloud", np.int32)
            L.Lp = ptr("Lp", np.float64)
            L.lap = ptr("lap", np.float64)
            L.btag = ptr("btag", np.uint8)
            L.order = ptr("order", np.int32)
            L.p = ptr("p", np.float64)
            L.s = ptr("s", np.float64)
            L.res = ptr("res", np.float64)
            L.rp = ptr("rp", np.int32)
            L.R = ptr("R", np.float64)
            L.pp = ptr("pp", np.int32)
            L.P = ptr("P", np.float64)
            L.cls = ptr("classes", np.int32)
            L.ncls = len(d["classes"]) - 1 if d.get("classes") is not None else 0
        self.arr = arr

    def relax(self, var, k, omega, num_relax):
        lib().orc_relax(var, C.byref(self.arr[k]), omega, num_relax)

    def residual(self, var, k):
        lib().orc_residual(var, C.byref(self.arr[k]))

    def restrict(self, var, k):
        lib().orc_restrict(var, C.byref(self.arr[k]), C.byref(self.arr[k + 1]))

    def prolong(self, var, k):
        lib().orc_prolong(var, C.byref(self.arr[k - 1]), C.byref(self.arr[k]))

    def vcycle(self, var, omega, num_relax, kinds=None):
        """One V-cycle; omega is one factor, or a sequence of one per level;
        kinds (H variant only): 0 SOR or 1 weighted Jacobi per level."""
        if kinds is not None:
            assert var == VAR_H and len(kinds) == len(self.lv)
            w = np.ascontiguousarray(np.broadcast_to(np.asarray(omega, np.float32), (len(self.lv),)))
            k = np.ascontiguousarray(kinds, dtype=np.int32)
            lib().orc_vcycle_smooth(len(self.lv), self.arr, w.ctypes.data, k.ctypes.data, num_relax)
            return
        if np.ndim(omega):
            w = np.ascontiguousarray(omega, dtype=np.float32)
            assert len(w) == len(self.lv)
            lib().orc_vcycle_levels(var, len(self.lv), self.arr, w.ctypes.data, num_relax)
        else:
            lib().orc_vcycle(var, len(self.lv), self.arr, omega, num_relax)
