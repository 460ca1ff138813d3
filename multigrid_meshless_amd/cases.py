"""Problem plug-ins of the time steps: the reference's init/ cases as functions
of a set-up Problem.

lid_driven_cavity follows init/init_LDC.c:16-44; calculate_dt follows
hf/general_functions.h:34-41.
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
