"""Problem plug-ins of the time steps: the reference's init/ cases as functions
of a set-up Problem.

lid_driven_cavity follows init/init_LDC.c:16-44; calculate_dt follows
hf/general_functions.h:34-41.  annulus builds the point cloud of the
annular Poisson case the learned smoother is also measured on (the reference
reads such clouds from Gmsh files, e.g. its mesh/2d/TC annuli).
"""
from __future__ import annotations

import numpy as np


def calculate_dt(d_avg: float, nu: float, courant: float) -> float:
    """dt = courant * 0.5 / (2 nu / d^2 + 1 / d), d the level-0 mean spacing."""
    termd = 2 * nu / (d_avg * d_avg)
    termc = 1 / d_avg
    return 0.5 / (termd + termc) * courant


def lid_driven_cavity(P, lid_speed: float = 1.0, timple: bool = False) -> None:
    """Initial and boundary conditions of the lid-driven cavity on every level:
    u = v = w = p = p_old = 0, then u = lid_speed and v = w = p = 0 on the nodes
    with |y - 1| < 1e-9 (init/init_LDC.c:16-44).  The fractional step reads the
    velocity from u; the TIMPLE step reads it from u_new (its own plug-in,
    init/init_TC.c:19-50, sets the wall velocity there), so with timple=True the
    same conditions go on u_new, v_new, w_new too."""
    names = ("u", "u_new") if timple else ("u",)
    for lev in range(P.L):
        g = P.geometry(lev)
        lid = np.abs(g["y"] - 1.0) < 1e-9
        zero = np.zeros(g["y"].size)
        for name in ("v", "w", "p", "p_old") + (("v_new", "w_new") if timple else ()):
            P.set_field(lev, name, zero)
        for name in names:
            P.set_field(lev, name, np.where(lid, lid_speed, 0.0))


def ns_params(Re: float, d_avg: float, courant: float, rho: float = 1.0, facRe: float = 1.0,
              facdt: float = 1.0, iter_momentum: int = 5, iter_timple: int = 1, num_vcycles: int = 10) -> dict:
    """Keyword arguments of Problem.ns_set_params as read_flow_parameters derives
    them (hf/mesh_functions.h:86-99: rho = 1, mu = rho / Re, nu = mu / rho) and
    mg_NS_solver.c:51 sets dt."""
    mu = rho / Re
    nu = mu / rho
    return dict(dt=calculate_dt(d_avg, nu, courant), rho=rho, mu=mu, nu=nu, facRe=facRe, facdt=facdt,
                field_rho=rho, iter_momentum=iter_momentum, iter_timple=iter_timple, num_vcycles=num_vcycles)


def annulus(n: int, seed: int, jitter: float = 0.2, r_in: float = 0.5, r_out: float = 1.0):
    """Jittered point cloud on the annulus r_in <= r <= r_out, laid out as the
    Gmsh reader leaves a level (boundary nodes first, no corners).

    Interior: the nodes of an n x n Cartesian grid of spacing h = 2 r_out / (n - 1)
    over [-r_out, r_out]^2, each moved by U(-jitter h, jitter h) per axis, kept
    where r_in + h/2 < r < r_out - h/2.  Boundary: round(2 pi r / h) nodes on
    each circle, equally spaced in angle with each angle moved by
    U(-jitter, jitter) of a spacing; unit normals point out of the domain (+r
    on the outer circle, -r on the inner one).  numpy's default generator,
    seeded with `seed`."""
    from . import Points
    rng = np.random.default_rng(seed)
    h = 2.0 * r_out / (n - 1)
    bx, by, bnx, bny = [], [], [], []
    for r, sgn in ((r_out, 1.0), (r_in, -1.0)):
        k = max(8, int(round(2 * np.pi * r / h)))
        t = (np.arange(k) + jitter * rng.uniform(-1.0, 1.0, k)) * (2 * np.pi / k)
        bx.append(r * np.cos(t)); by.append(r * np.sin(t))
        bnx.append(sgn * np.cos(t)); bny.append(sgn * np.sin(t))
    g = -r_out + h * np.arange(n)
    X, Y = np.meshgrid(g, g, indexing="ij")
    X = X.ravel() + jitter * h * rng.uniform(-1.0, 1.0, X.size)
    Y = Y.ravel() + jitter * h * rng.uniform(-1.0, 1.0, Y.size)
    R = np.hypot(X, Y)
    keep = (R > r_in + 0.5 * h) & (R < r_out - 0.5 * h)
    x = np.concatenate(bx + [X[keep]])
    y = np.concatenate(by + [Y[keep]])
    nb = sum(len(b) for b in bx)
    N = x.size
    nx = np.concatenate(bnx + [np.zeros(N - nb)])
    ny = np.concatenate(bny + [np.zeros(N - nb)])
    btag = np.zeros(N, np.uint8)
    btag[:nb] = 1
    zero = np.zeros(N)
    return Points.from_arrays(N, nb, 2 * N, h, x, y, zero, nx, ny, zero, btag, np.zeros(N, np.uint8))


def fourier_rhs(x, y, z, seed: int = 2, modes: int = 16):
    """f = sum_q a_q sin(pi (k_q . x)) with k_q in {1..4}^3 and a_q ~ N(0, 1): the
    random smooth right-hand side of the 3D cube Poisson workload.  The draws
    come from SplitMix64 seeded with `seed`: three wave numbers 1 + (u mod 4),
    then a_q by Box-Muller from two uniforms (top 53 bits of a draw)."""
    import ctypes as C
    import math

    from ._lib import lib
    s = C.c_uint64(seed)

    def u01():
        return (lib().mgm_splitmix64(C.byref(s)) >> 11) * 2.0 ** -53

    f = np.zeros_like(np.asarray(x, dtype=np.float64))
    for _ in range(modes):
        k = [1 + (lib().mgm_splitmix64(C.byref(s)) % 4) for _ in range(3)]
        u1, u2 = u01(), u01()
        a = math.sqrt(-2.0 * math.log(max(u1, 1e-300))) * math.cos(2 * math.pi * u2)
        f += a * np.sin(math.pi * (k[0] * x + k[1] * y + k[2] * z))
    return f


def cube_levels(n: int, levels: int):
    """Points per axis of a hierarchy that halves its spacing: n, (n+1)//2, ..."""
    sizes, v = [], n
    for _ in range(levels):
        sizes.append(v)
        v = (v + 1) // 2
    return sizes


def cube_poisson(n: int = 100, levels: int = 5, seed0: int = 1, setup: bool = True,
                 boundary_multiplier: float = 0.0):
    """The 3D cube Poisson hierarchy: a jittered n^3 cloud on [0,1]^3 (jitter
    0.2 h, seeds seed0, seed0+1, ... per level) and `levels` clouds halving in
    spacing; PHS r^3 + cubic polynomials, multiplier 3.5 (m = 70, mpn = 90),
    multicolour SOR with omega 1 and 3 sweeps, transfer rows on every interior
    node, Dirichlet identity rows on every boundary node of every level.
    boundary_multiplier > 3.5 adds wide boundary stencils (5.25: m_b = 105,
    125-unknown systems).  Returns (problem, sizes, setup seconds)."""
    import time

    from . import ORDER_COLOUR, ROWS_BY_TAG, Points, Problem, config
    sizes = cube_levels(n, levels)
    pts = [Points.generate(3, s, seed0 + k, 0.2) for k, s in enumerate(sizes)]
    cfg = config(dimension=3, poly_degree=3, coarse_poly_degree=3, phs_degree=3, cloud_size_multiplier=3.5,
                 omega=1.0, num_relax=3, transfer_rows=ROWS_BY_TAG, sweep_order=ORDER_COLOUR,
                 boundary_cloud_size_multiplier=boundary_multiplier)
    P = Problem(cfg, pts)
    t = 0.0
    if setup:
        t0 = time.perf_counter()
        P.setup()
        t = time.perf_counter() - t0
        P.dirichlet_rows()
    return P, sizes, t


def cube_rhs(P, seed: int = 2):
    """The 16-mode Fourier source on level 0's interior nodes, 0 on its boundary
    (seed 2: the benchmark's source)."""
    g = P.geometry(0)
    f = fourier_rhs(g["x"], g["y"], g["z"], seed=seed)
    f[g["btag"] != 0] = 0.0
    return f
