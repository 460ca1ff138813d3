"""multigrid_meshless_amd -- MI355X-native RBF-FD meshless multigrid hot path.

Python host mirror of the C ABI in include/mgm.h (libmgm.so, HIP kernels for
gfx950).  Names follow the reference's operator interface:
read_PointStructure/read_complete_mesh_data -> Points/Problem.setup,
relaxation/calculate_residuals/restrict_residuals/prolongate_corrections ->
Problem.relax/residual/restrict/prolong, the V-cycle loop -> Problem.vcycle,
and the resident solve -> Problem.solve.  Arrays are numpy, in the reference's
node order (after its BFS reordering), row-major.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import Config, LevelInfo, MgmError, check, lib

VAR_H, VAR_T, VAR_F = 0, 1, 2
SMOOTH_SOR, SMOOTH_JACOBI = 0, 1
ROWS_REFERENCE, ROWS_BY_TAG = 0, 1
ORDER_COLOUR, ORDER_NATURAL = 0, 1
MAT = dict(Dx=0, Dy=1, Dz=2, lap=3, lap_Poison=4, restr=5, prol=6)
FIELD = dict(p=0, source=1, res=2, pprime=3, u=4, v=5, w=6, u_new=7, v_new=8, w_new=9,
             p_old=10, T=11, dpdn=12, dpdx=13, dpdy=14, dpdz=15, u_old=16, v_old=17, w_old=18)

__all__ = ["Config", "Points", "Problem", "init", "device_count", "MgmError", "config", "stream_copy_seconds",
           "write_history",
           "VAR_H", "VAR_T", "VAR_F", "SMOOTH_SOR", "SMOOTH_JACOBI", "MAT", "FIELD"]


def device_count() -> int:
    return lib().mgm_device_count()


def init(device: int = 0):
    check(lib().mgm_init(device), "mgm_init")


def stream_copy_seconds(nbytes: int, reps: int = 20) -> float:
    """Seconds per device-to-device streaming copy of nbytes (measured HBM reference)."""
    s = C.c_double()
    check(lib().mgm_stream_copy(int(nbytes), int(reps), C.byref(s)), "stream_copy")
    return s.value


def write_history(path: str, values, scientific: bool = True) -> None:
    """Convergence file: line i is "i, value", printed "%e" (mg_NS_solver.c:77-78)
    or "%lf" (mg_heat_conduction.c:77)."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    check(lib().mgm_write_history(path.encode(), int(v.size), v.ctypes.data, int(scientific)), "write_history")


def config(**kw) -> Config:
    c = Config()
    lib().mgm_config_default(C.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


class Points:
    """A level's points as the Gmsh reader leaves them (hf/mesh_functions.h:172-476)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def read_gmsh(cls, path: str, dim: int) -> "Points":
        h = C.c_void_p()
        check(lib().mgm_points_read_gmsh(path.encode(), dim, C.byref(h)), "read_gmsh")
        return cls(h)

    @classmethod
    def generate(cls, dim: int, n: int, seed: int, jitter: float = 0.2,
                 with_volume_elements: bool = False, gmsh_path: str | None = None) -> "Points":
        h = C.c_void_p()
        check(lib().mgm_points_generate(dim, n, seed, jitter, int(with_volume_elements),
                                        gmsh_path.encode() if gmsh_path else None, C.byref(h)),
              "generate")
        return cls(h)

    @classmethod
    def from_arrays(cls, N, nb, num_elem, d_avg, x, y, z, nx, ny, nz, btag, ctag) -> "Points":
        h = C.c_void_p()
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, nx, ny, nz)]
        tags = [np.ascontiguousarray(a, dtype=np.uint8) for a in (btag, ctag)]
        check(lib().mgm_points_from_arrays(int(N), int(nb), int(num_elem), float(d_avg),
                                           *[a.ctypes.data for a in arrs + tags], C.byref(h)),
              "points_from_arrays")
        return cls(h)

    def info(self):
        N, nb, ne, d = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        check(lib().mgm_points_info(self._h, C.byref(N), C.byref(nb), C.byref(ne), C.byref(d)))
        return dict(N=N.value, nb=nb.value, num_elem=ne.value, d_avg=d.value)

    def arrays(self) -> dict:
        N = self.info()["N"]
        out = {k: np.zeros(N) for k in ("x", "y", "z", "nx", "ny", "nz")}
        out["btag"] = np.zeros(N, np.uint8)
        out["ctag"] = np.zeros(N, np.uint8)
        check(lib().mgm_points_copy(self._h, *[out[k].ctypes.data for k in
                                                ("x", "y", "z", "nx", "ny", "nz", "btag", "ctag")]))
        return out

    def __del__(self):
        try:
            if self._h:
                lib().mgm_points_free(self._h)
                self._h = None
        except Exception:
            pass


class Problem:
    """A multigrid hierarchy resident on the GPU."""

    def __init__(self, cfg: Config, levels: list):
        self.cfg = cfg
        self.L = len(levels)
        h = C.c_void_p()
        check(lib().mgm_problem_create(C.byref(cfg), self.L, C.byref(h)), "problem_create")
        self._h = h
        self._levels = levels
        for k, pts in enumerate(levels):
            check(lib().mgm_problem_set_level(h, k, pts._h), "set_level")

    def setup(self):
        check(lib().mgm_problem_setup(self._h), "setup")
        return self

    def info(self, lev: int) -> dict:
        o = LevelInfo()
        check(lib().mgm_level_info_get(self._h, lev, C.byref(o)), "level_info")
        return {k: getattr(o, k) for k, _ in o._fields_}

    def timings(self) -> dict:
        t = np.zeros(6)
        check(lib().mgm_setup_timings(self._h, t.ctypes.data))
        return dict(zip(("knn", "reorder_colour", "transfer_points", "weights", "slabs", "total"), t))

    def rebuild_weights(self, lev: int):
        s, n = C.c_double(), C.c_long()
        check(lib().mgm_rebuild_weights(self._h, lev, C.byref(s), C.byref(n)), "rebuild_weights")
        return s.value, n.value

    def weight_kernel_choice(self, lev: int = 0):
        """(tile, waves_per_simd) of the weight kernel that ran the last build of
        this level's system size (the autotuner's choice unless forced), or None.
        waves_per_simd 0 is the matrix-core kernel (FP32 LU on MFMA + FP64
        refinement, tile = its 16 x 16 tile count); otherwise the FP64
        Gauss-Jordan register-tile kernel."""
        info = self.info(lev)
        t, w = C.c_int(), C.c_int()
        if lib().mgm_weight_kernel_choice(info["m"] + info["n"], C.byref(t), C.byref(w)) != 0:
            return None
        return t.value, w.value

    def geometry(self, lev: int) -> dict:
        N = self.info(lev)["N"]
        out = {k: np.zeros(N) for k in ("x", "y", "z", "nx", "ny", "nz")}
        out["btag"] = np.zeros(N, np.uint8)
        check(lib().mgm_get_geometry(self._h, lev, *[out[k].ctypes.data for k in
                                                     ("x", "y", "z", "nx", "ny", "nz", "btag")]))
        return out

    def cloud(self, lev: int) -> np.ndarray:
        i = self.info(lev)
        out = np.zeros((i["N"], i["m"]), np.int32)
        check(lib().mgm_get_cloud(self._h, lev, out.ctypes.data))
        return out

    def points(self, lev: int, which: str) -> np.ndarray:
        out = np.zeros(self.info(lev)["N"], np.int32)
        check(lib().mgm_get_points(self._h, lev, 0 if which == "restriction" else 1, out.ctypes.data))
        return out

    def colours(self, lev: int) -> np.ndarray:
        out = np.zeros(self.info(lev)["N"], np.int32)
        check(lib().mgm_get_colours(self._h, lev, out.ctypes.data))
        return out

    def _width(self, lev, which):
        if which == MAT["restr"]:
            return self.info(lev - 1)["m"]
        if which == MAT["prol"]:
            return self.info(lev + 1)["m"]
        return self.info(lev)["m"]

    def matrix(self, lev: int, name: str) -> np.ndarray:
        w = MAT[name]
        out = np.zeros((self.info(lev)["N"], self._width(lev, w)))
        check(lib().mgm_get_matrix(self._h, lev, w, out.ctypes.data), f"get_matrix {name}")
        return out

    def set_matrix(self, lev: int, name: str, a: np.ndarray):
        w = MAT[name]
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (self.info(lev)["N"], self._width(lev, w))
        check(lib().mgm_set_matrix(self._h, lev, w, a.ctypes.data), f"set_matrix {name}")

    def dirichlet_rows(self):
        check(lib().mgm_dirichlet_rows(self._h), "dirichlet_rows")

    def field(self, lev: int, name: str) -> np.ndarray:
        out = np.zeros(self.info(lev)["N"])
        check(lib().mgm_get_field(self._h, lev, FIELD[name], out.ctypes.data), f"get_field {name}")
        return out

    def set_field(self, lev: int, name: str, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (self.info(lev)["N"],)
        check(lib().mgm_set_field(self._h, lev, FIELD[name], a.ctypes.data), f"set_field {name}")

    def write_solution(self, path: str, fields=("p",), with_z: bool = False, lev: int = 0) -> None:
        """One line per node in reference order, "x, y[, z], fields..." printed
        "%lf": ("p",) without z is the heat driver's Solution.csv
        (mg_heat_conduction.c:89-93), ("u", "v", "w", "p_old") with z the NS
        driver's (mg_NS_solver.c:84-90)."""
        ids = (C.c_int * len(fields))(*[FIELD[k] for k in fields])
        check(lib().mgm_write_solution(self._h, lev, path.encode(), int(with_z), len(fields), ids),
              "write_solution")

    def zero_fields(self):
        check(lib().mgm_zero_fields(self._h))

    def set_smoother(self, omega: float, num_relax: int):
        check(lib().mgm_set_smoother(self._h, omega, num_relax), "set_smoother")

    def tune_omega(self, tol=1e-8, max_cycles=400, variant=VAR_H, precision=0, lo=0.5, hi=1.0, step=0.05):
        """Pick the SOR factor that solves the current level-0 source (from the
        current iterate) in the fewest cycles: a grid lo..hi by step, then the
        midpoints beside the best.  Sets it as the smoother's factor; the
        iterate and source are restored, every other field left zero."""
        w, n = C.c_float(), C.c_int()
        check(lib().mgm_tune_omega(self._h, variant, precision, tol, max_cycles, lo, hi, step, C.byref(w),
                                   C.byref(n)), "tune_omega")
        return dict(omega=w.value, cycles=n.value)

    def tune_omega_levels(self, tol=1e-8, max_cycles=400, variant=VAR_H, precision=0, lo=0.5, hi=1.5, step=0.1,
                          sweeps=2):
        """Per-level SOR factors by coordinate descent from the current ones
        (fewest cycles to tol on the current level-0 source); they stay set."""
        w = np.zeros(self.L, np.float32)
        n = C.c_int()
        check(lib().mgm_tune_omega_levels(self._h, variant, precision, tol, max_cycles, lo, hi, step, sweeps,
                                          w.ctypes.data, C.byref(n)), "tune_omega_levels")
        return dict(omega=[float(v) for v in w], cycles=n.value)

    def set_smoother_levels(self, omega):
        """A SOR factor per level (None or [] returns to the single factor)."""
        w = np.ascontiguousarray(omega if omega is not None else [], dtype=np.float32)
        check(lib().mgm_set_smoother_levels(self._h, w.ctypes.data if len(w) else None, len(w)),
              "set_smoother_levels")

    def smoother_levels(self):
        w = np.zeros(self.L, np.float32)
        check(lib().mgm_get_smoother_levels(self._h, w.ctypes.data), "smoother_levels")
        return [float(v) for v in w]

    def set_smoother_kinds(self, kinds):
        """SMOOTH_SOR or SMOOTH_JACOBI per level (None or [] returns to SOR everywhere)."""
        k = np.ascontiguousarray(kinds if kinds is not None else [], dtype=np.int32)
        check(lib().mgm_set_smoother_kinds(self._h, k.ctypes.data if len(k) else None, len(k)),
              "set_smoother_kinds")

    def set_learned_levels(self, use):
        """Levels whose first sweep the learned network replaces (None: all that fit)."""
        u = np.ascontiguousarray([1 if x else 0 for x in use] if use is not None else [], dtype=np.int32)
        check(lib().mgm_set_learned_levels(self._h, u.ctypes.data if len(u) else None, len(u)),
              "set_learned_levels")

    def smoother_kinds(self):
        k = np.zeros(self.L, np.int32)
        check(lib().mgm_get_smoother_kinds(self._h, k.ctypes.data), "smoother_kinds")
        return [int(v) for v in k]

    def tune_smoother(self, tol=1e-8, max_cycles=400, variant=VAR_H, precision=0, lo=0.5, hi=1.5, step=0.1,
                      sweeps=3, kinds=(SMOOTH_SOR, SMOOTH_JACOBI)):
        """Per-level smoother kind and factor by coordinate descent from the
        current ones (fewest cycles to tol on the current level-0 source); they
        stay set."""
        w = np.zeros(self.L, np.float32)
        k = np.zeros(self.L, np.int32)
        n = C.c_int()
        mask = sum(1 << kd for kd in kinds)
        check(lib().mgm_tune_smoother(self._h, variant, precision, tol, max_cycles, lo, hi, step, sweeps, mask,
                                      w.ctypes.data, k.ctypes.data, C.byref(n)), "tune_smoother")
        return dict(omega=[float(v) for v in w], kind=[int(v) for v in k], cycles=n.value)

    # ---- operators (reference variant semantics) ----
    def relax(self, lev, variant=VAR_H):
        check(lib().mgm_relax(self._h, lev, variant), "relax")

    def residual(self, lev, variant=VAR_H):
        check(lib().mgm_residual(self._h, lev, variant), "residual")

    def restrict(self, lev, variant=VAR_H):
        check(lib().mgm_restrict(self._h, lev, variant), "restrict")

    def prolong(self, lev, variant=VAR_H):
        check(lib().mgm_prolong(self._h, lev, variant), "prolong")

    def vcycle(self, variant=VAR_H, ncycles=1):
        check(lib().mgm_vcycle(self._h, variant, ncycles), "vcycle")

    def spmv(self, lev, matrix="lap_Poison", src="p", dst="dpdn"):
        check(lib().mgm_spmv(self._h, lev, MAT[matrix], FIELD[src], FIELD[dst]), "spmv")

    def spmv_wide(self, lev, matrix="Dx", src="p", dst="dpdx"):
        """spmv with the boundary rows taken from the wide boundary stencils."""
        check(lib().mgm_spmv_wide(self._h, lev, MAT[matrix], FIELD[src], FIELD[dst]), "spmv_wide")

    def boundary_info(self, lev: int):
        """(number of boundary nodes with a wide stencil, its width m_b); (0, 0) if none."""
        nbw, mb = C.c_int(), C.c_int()
        check(lib().mgm_boundary_info(self._h, lev, C.byref(nbw), C.byref(mb)), "boundary_info")
        return nbw.value, mb.value

    def boundary_rows(self, lev: int, name: str = "Dx") -> dict:
        """The wide boundary rows of Dx, Dy, Dz or lap: nodes (reference order),
        cloud and weights (nbw x m_b)."""
        nbw, mb = self.boundary_info(lev)
        nodes = np.zeros(nbw, np.int32)
        cloud = np.zeros((nbw, mb), np.int32)
        w = np.zeros((nbw, mb))
        check(lib().mgm_get_boundary_rows(self._h, lev, MAT[name], nodes.ctypes.data, cloud.ctypes.data,
                                          w.ctypes.data), "get_boundary_rows")
        return dict(nodes=nodes, cloud=cloud, W=w)

    def l2_norm_gen(self, lev, a, b=None) -> float:
        out = C.c_double()
        check(lib().mgm_l2_norm_gen(self._h, lev, FIELD[a], -1 if b is None else FIELD[b],
                                    C.byref(out)), "l2_norm_gen")
        return out.value

    def solve(self, tol=1e-8, max_cycles=200, variant=VAR_H, precision=0):
        hist = np.full(max_cycles + 1, np.nan)
        nc, sec = C.c_int(), C.c_double()
        check(lib().mgm_solve(self._h, variant, precision, tol, max_cycles, C.byref(nc),
                              hist.ctypes.data, C.byref(sec)), "solve")
        return dict(cycles=nc.value, seconds=sec.value, history=hist[:nc.value + 1])

    # ---- time steps (level-0 fields, resident) ----
    def ns_set_params(self, **kw):
        """dt, rho, mu, nu, facRe, facdt, field_rho, iter_momentum, iter_timple, num_vcycles."""
        from ._lib import NsParams
        check(lib().mgm_ns_set_params(self._h, C.byref(NsParams(**kw))), "ns_set_params")

    def timple_step(self) -> float:
        err = C.c_double()
        check(lib().mgm_timple_step(self._h, C.byref(err)), "timple_step")
        return err.value

    def fs_step(self) -> float:
        err = C.c_double()
        check(lib().mgm_fs_step(self._h, C.byref(err)), "fs_step")
        return err.value

    def ns_march(self, scheme=VAR_F, steps=1, steady_tol=0.0, precision=0) -> dict:
        """The mg_NS_solver.c time loop: up to `steps` steps of the fractional step
        (VAR_F) or TIMPLE (VAR_T), one step captured as a graph and replayed."""
        errs = np.zeros(max(steps, 1))
        done, sec = C.c_int(), C.c_double()
        check(lib().mgm_ns_march(self._h, scheme, precision, steps, steady_tol, errs.ctypes.data,
                                 C.byref(done), C.byref(sec)), "ns_march")
        return {"steps": done.value, "errors": errs[:done.value], "seconds": sec.value}

    # ---- learned smoother ----
    def mlp_inputs(self) -> int:
        n = lib().mgm_mlp_inputs(self._h)
        check(min(n, 0), "mlp_inputs")
        return n

    def mlp_weights(self) -> np.ndarray:
        w = np.zeros(lib().mgm_mlp_weight_count(self.mlp_inputs()), np.float32)
        check(lib().mgm_mlp_get_weights(self._h, w.ctypes.data), "mlp_get_weights")
        return w

    def set_mlp_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == lib().mgm_mlp_weight_count(self.mlp_inputs())
        check(lib().mgm_mlp_set_weights(self._h, w.ctypes.data), "mlp_set_weights")

    def mlp_forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.zeros(X.shape[0], np.float32)
        check(lib().mgm_mlp_forward(self._h, X.shape[0], X.ctypes.data, y.ctypes.data), "mlp_forward")
        return y

    def mlp_gradient(self, X, t):
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(lib().mgm_mlp_weight_count(self.mlp_inputs()), np.float32)
        loss = C.c_double()
        check(lib().mgm_mlp_gradient(self._h, X.shape[0], X.ctypes.data, t.ctypes.data, C.byref(loss),
                                     g.ctypes.data), "mlp_gradient")
        return loss.value, g

    def mlp_features(self, nodes, res, err):
        """Network inputs and (unclipped) targets of level-0 nodes from
        reference-order residual and error fields."""
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        res = np.ascontiguousarray(res, dtype=np.float64)
        err = np.ascontiguousarray(err, dtype=np.float64)
        N = self.info(0)["N"]
        assert res.size == N and err.size == N
        X = np.zeros((nodes.size, self.mlp_inputs()), np.float32)
        t = np.zeros(nodes.size, np.float32)
        check(lib().mgm_mlp_features(self._h, nodes.size, nodes.ctypes.data, res.ctypes.data, err.ctypes.data,
                                     X.ctypes.data, t.ctypes.data), "mlp_features")
        return X, t

    def mlp_fit(self, X, t, epochs=1, batch=4096, lr=1e-3, seed=7):
        """Adam on the current weights over rows (X, t); (loss before, loss after)."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        assert X.shape == (t.size, self.mlp_inputs())
        l0, l1 = C.c_double(), C.c_double()
        check(lib().mgm_mlp_fit(self._h, t.size, X.ctypes.data, t.ctypes.data, epochs, batch, lr, seed,
                                C.byref(l0), C.byref(l1)), "mlp_fit")
        return l0.value, l1.value

    def train_smoother(self, snapshots=8, samples_per_snapshot=8192, epochs=30, batch=4096, lr=1e-3, seed=7):
        from ._lib import TrainConfig
        cfg = TrainConfig(snapshots, samples_per_snapshot, epochs, batch, lr, seed)
        l0, l1 = C.c_double(), C.c_double()
        check(lib().mgm_train_smoother(self._h, C.byref(cfg), C.byref(l0), C.byref(l1)), "train_smoother")
        return l0.value, l1.value

    def use_learned_smoother(self, enable=True):
        check(lib().mgm_use_learned_smoother(self._h, int(bool(enable))), "use_learned_smoother")

    def learned_fell_back(self) -> bool:
        return bool(lib().mgm_learned_fell_back(self._h))

    def time_spmv(self, lev=0, reps=20) -> float:
        s = C.c_double()
        check(lib().mgm_time_spmv(self._h, lev, reps, C.byref(s)), "time_spmv")
        return s.value

    def __del__(self):
        try:
            if self._h:
                lib().mgm_problem_destroy(self._h)
                self._h = None
        except Exception:
            pass
