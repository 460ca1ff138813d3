"""Shared helpers: run the oracle pipeline on a golden case's inputs.

Mirrors read_complete_mesh_data (hf/mesh_functions.h:914-928) and
create_derivative_matrices (hf/rbf.h:401-421) step by step, on the arrays the
Gmsh reader produced (stored in each fixture as in<k>_*).
"""
from __future__ import annotations

import numpy as np


def case_meta(g):
    dim, poly, phs, mult, nv = g["case_meta"]
    return int(dim), int(poly), int(phs), float(mult), int(nv)


def level_inputs(g, k):
    d = {key: g[f"in{k}_{key}"].copy() for key in ("x", "y", "z", "nx", "ny", "nz", "btag", "ctag")}
    N, nb, ne = g[f"in{k}_meta"]
    d.update(N=int(N), nb=int(nb), num_elem=int(ne), d_avg=float(g[f"in{k}_d_avg"]))
    return d


def num_levels(g):
    k = 0
    while f"in{k}_x" in g:
        k += 1
    return k


def oracle_setup(orc, g, weights=True, row_mode=0, method=0):
    """Full oracle setup on a golden case; returns list of level dicts."""
    dim, poly, phs, mult, _ = case_meta(g)
    L = num_levels(g)
    levels = []
    for k in range(L):
        lv = level_inputs(g, k)
        lv["poly"] = poly if k == 0 else 3
        px, _, _ = orc.poly_terms(dim, lv["poly"])
        lv["n"] = len(px)
        lv["m"] = int(np.floor(mult * lv["n"] + 1e-9))
        lv["cloud"] = orc.find_cloud_index(lv["x"], lv["y"], lv["z"], lv["btag"], lv["ctag"],
                                           lv["nb"], lv["d_avg"], lv["m"])
        lv["q1"] = orc.rcm(lv)
        orc.correct_normals(lv)
        lv["d_avg"] = orc.avg_dx(lv, dim)
        levels.append(lv)
    for k, lv in enumerate(levels):
        other = levels[k] if k == L - 1 else levels[k + 1]
        lv["pp"] = orc.find_nearest_point(lv, other, other["d_avg"])
        other = levels[k] if k == 0 else levels[k - 1]
        lv["rp"] = orc.find_nearest_point(lv, other, other["d_avg"])
    if weights:
        for k, lv in enumerate(levels):
            Dx, Dy, Dz, lap = orc.weights(lv, dim, phs, lv["poly"], method)
            lv.update(Dx=Dx, Dy=Dy, Dz=Dz, lap=lap)
            lv["lap_Poison"] = orc.lap_poisson(lv, dim, Dx, Dy, Dz, lap)
        for k in range(L):
            if k > 0:
                levels[k]["restr"] = orc.restriction(levels[k - 1], levels[k], dim, phs,
                                                     levels[k - 1]["poly"], row_mode, method)
            if k < L - 1:
                levels[k]["prol"] = orc.prolongation(levels[k], levels[k + 1], dim, phs,
                                                     levels[k + 1]["poly"], row_mode, method)
    return levels


def rel_row_err(a, b):
    """max over rows of |a-b| / max|b| (row-wise relative error)."""
    a = np.asarray(a)
    b = np.asarray(b)
    scale = np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-300)
    return float((np.abs(a - b) / scale).max())


def exact_weight_row(orc, lv, i, dim, phs, op="Dx"):
    """Weights of row i solved in x87 extended precision (64-bit mantissa) by
    pivoted elimination: an independent check of which FP64 result is closer
    to the exact weights (hf/rbf.h:42-77 matrix, row-0 right-hand side)."""
    px, py, pz = orc.poly_terms(dim, lv["poly"])
    m, n = lv["m"], len(px)
    ld = np.longdouble
    cl = lv["cloud"][i]
    X = np.array([lv["x"][cl], lv["y"][cl], lv["z"][cl]], dtype=ld).T
    N = m + n
    A = np.zeros((N, N), dtype=ld)
    half = ld(phs) / 2
    for a in range(m):
        d2 = ((X[a, :dim] - X[:, :dim]) ** 2).sum(axis=1)
        A[a, :m] = np.where(d2 == 0, ld(0), d2 ** half)
        for k in range(n):
            v = ld(1)
            for d, pw in enumerate((px[k], py[k], pz[k])):
                v *= (X[a, d] - X[0, d]) ** int(pw)
            A[a, m + k] = v
            A[m + k, a] = v
    b = np.zeros(N, dtype=ld)
    d2 = ((X[0, :dim] - X[:, :dim]) ** 2).sum(axis=1)
    rb = np.where(d2 == 0, ld(0), d2 ** ((ld(phs) - 2) / 2))
    if op == "Dx":
        b[:m] = phs * (X[0, 0] - X[:, 0]) * rb
        for k in range(n):
            if px[k] == 1 and py[k] == 0 and pz[k] == 0:
                b[m + k] = 1
    else:
        b[:m] = phs * phs * rb
        for k in range(n):
            if sorted((px[k], py[k], pz[k])) == [0, 0, 2]:
                b[m + k] = 2
    M, r = A.copy(), b.copy()
    for k in range(N):
        p = k + int(np.argmax(np.abs(M[k:, k])))
        M[[k, p]] = M[[p, k]]
        r[[k, p]] = r[[p, k]]
        l = M[k + 1:, k] / M[k, k]
        M[k + 1:, k:] -= l[:, None] * M[k, k:]
        r[k + 1:] -= l * r[k]
    xs = np.zeros(N, dtype=ld)
    for k in range(N - 1, -1, -1):
        xs[k] = (r[k] - (M[k, k + 1:] * xs[k + 1:]).sum()) / M[k, k]
    return xs[:m].astype(np.float64)
