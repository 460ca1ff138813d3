// rbf_entry.hpp -- entries of the local RBF-FD systems, shared by the weight
// kernels (weights.hip, lu32.hip).  FP64, no contraction: every kernel forms
// the same matrix entries bit for bit.
//
// For a stencil x_0 (the seed) .. x_{m-1} and n monomials p_k:
//   A = [[Phi, P], [P^T, 0]],  Phi_ij = |x_i - x_j|^phs,  P_ik = p_k(x_i - x_0)
// (hf/rbf.h:42-77).  The right-hand side of an operator is row 0 of its
// right-hand-side matrix (the operator applied to phi(|x - x_j|) and to the
// monomials at x = x_0: hf/rbf.h:79-270); that of an interpolation target t
// is [phi(|t - x_j|); p_k(t - x_0)] (hf/mesh_functions.h:533-670).
#pragma once

#include <hip/hip_runtime.h>

namespace mgm {

__device__ inline double rbf_phs(double sum, int phs)
{
    if (sum == 0) return 0;
    if (phs & 1) {  // r^phs = sum^((phs-1)/2) * sqrt(sum)
        double r = sqrt(sum), v = r;
        for (int k = 0; k < (phs - 1) / 2; k++) v = v * sum;
        return v;
    }
    return pow(sum, phs * 0.5);
}

__device__ inline double ipow(double d, int p)
{
    double v = 1;
    for (int k = 0; k < p; k++) v = (k == 0) ? d : v * d;
    return v;
}

__device__ inline double dist_sum(double ax, double ay, double az, double bx, double by, double bz, int dim)
{
    double s = 0, d;
    d = ax - bx; s += d * d;
    d = ay - by; s += d * d;
    if (dim == 3) { d = az - bz; s += d * d; }
    return s;
}

// p_k(dx, dy, dz) for exponents (ex, ey, ez); factors of exponent 0 are
// skipped, so a factor 0^0 never appears.
__device__ inline double monomial(double dx, double dy, double dz, int ex, int ey, int ez)
{
    double w = 1;
    if (ex) w *= ipow(dx, ex);
    if (ey) w *= ipow(dy, ey);
    if (ez) w *= ipow(dz, ez);
    return w;
}

// Entry i (i < m + n) of the right-hand side of operator op (0 Dx, 1 Dy,
// 2 Dz, 3 Laplacian) at the seed (sx, sy, sz).
__device__ inline double operator_rhs(int op, int i, int m, int dim, int phs, const short *px, const short *py,
                                      const short *pz, const double *cx, const double *cy, const double *cz,
                                      double sx, double sy, double sz)
{
    if (i < m) {
        const double s2 = dist_sum(sx, sy, sz, cx[i], cy[i], cz[i], dim);
        const double rb = rbf_phs(s2, phs - 2);
        if (op < 3) {
            const double d = (op == 0) ? sx - cx[i] : (op == 1 ? sy - cy[i] : sz - cz[i]);
            return phs * d * rb;
        }
        return (phs * phs) * rb;
    }
    const int k = i - m, ex = px[k], ey = py[k], ez = pz[k];
    if (op == 0 && ex == 1 && ey == 0 && ez == 0) return 1;
    if (op == 1 && ex == 0 && ey == 1 && ez == 0) return 1;
    if (op == 2 && ex == 0 && ey == 0 && ez == 1) return 1;
    if (op == 3 && ((ex == 2 && ey == 0 && ez == 0) || (ex == 0 && ey == 2 && ez == 0) ||
                    (ex == 0 && ey == 0 && ez == 2)))
        return 2;
    return 0;
}

// Entry i of the right-hand side of an interpolation target t.
__device__ inline double target_rhs(int i, int m, int dim, int phs, const short *px, const short *py,
                                    const short *pz, const double *cx, const double *cy, const double *cz,
                                    double sx, double sy, double sz, double tx, double ty, double tz)
{
    if (i < m) return rbf_phs(dist_sum(tx, ty, tz, cx[i], cy[i], cz[i], dim), phs);
    const int k = i - m;
    return ipow(tx - sx, px[k]) * ipow(ty - sy, py[k]) * ipow(tz - sz, pz[k]);
}

}  // namespace mgm
