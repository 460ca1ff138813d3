// weights.hip -- batched RBF-FD local solves (operator weights and transfer rows).
//
// For every node s the reference builds the augmented PHS + polynomial matrix
// A = [[Phi, P], [P^T, 0]] over its stencil (hf/rbf.h:42-77), inverts it and
// keeps the first m entries of  L0 * A^{-1}  for L0 = row 0 of the operator's
// right-hand-side matrix (Dx, Dy, Dz, Laplacian: hf/rbf.h:79-270, 274-380), and
// for restriction/prolongation targets the first m entries of
// [phi(|x_t - x_j|); p(x_t - x_seed)]^T A^{-1} (hf/mesh_functions.h:533-670).
// A is symmetric, so each of these is the solution w of A w = b.  Here every
// system is factorised once (LU, partial pivoting) and all its right-hand
// sides -- the four operators and every interpolation target that maps to
// this node -- are solved with that one factorisation, instead of the
// reference's one Gauss-Jordan inverse per operator and per transfer row.
//
// Layout: one lane per system; the batch workspace is interleaved
// (element (i,j) of system b at ws[(i*mpn+j)*batch + b]) so the 64 lanes of a
// wave touch 64 consecutive doubles on every access.
#pragma clang fp contract(off)

#include "device.hpp"

namespace mgm {

namespace {

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

__device__ inline double dist_sum(double ax, double ay, double az, double bx, double by, double bz,
                                  int dim)
{
    double s = 0, d;
    d = ax - bx; s += d * d;
    d = ay - by; s += d * d;
    if (dim == 3) { d = az - bz; s += d * d; }
    return s;
}

__global__ void __launch_bounds__(64)
weights_kernel(WeightsArgs a, int base, int count)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= count) return;
    const int s = a.sys[base + b];
    const int m = a.m, n = a.n, mpn = m + n, B = a.batch, dim = a.dim, phs = a.phs;
    double *__restrict__ ws = a.ws;
    double *__restrict__ rhs = a.rhs;
#define WS(i, j) ws[((size_t)(i) * mpn + (j)) * B + b]
#define RHS(r, i) rhs[(((size_t)(r)) * mpn + (i)) * B + b]
    const int *cl = a.cloud + (size_t)s * m;
    const double sx = a.x[cl[0]], sy = a.y[cl[0]], sz = a.z[cl[0]];

    // ---- assemble A (hf/rbf.h:42-77) ----
    for (int i = 0; i < m; i++) {
        const double xi = a.x[cl[i]], yi = a.y[cl[i]], zi = a.z[cl[i]];
        WS(i, i) = 0;
        for (int j = i + 1; j < m; j++) {
            double v = rbf_phs(dist_sum(xi, yi, zi, a.x[cl[j]], a.y[cl[j]], a.z[cl[j]], dim), phs);
            WS(i, j) = v;
            WS(j, i) = v;
        }
        const double dx = xi - sx, dy = yi - sy, dz = zi - sz;
        for (int k = 0; k < n; k++) {
            double v = 1;
            if (a.px[k]) v *= ipow(dx, a.px[k]);
            if (a.py[k]) v *= ipow(dy, a.py[k]);
            if (a.pz[k]) v *= ipow(dz, a.pz[k]);
            WS(i, m + k) = v;
            WS(m + k, i) = v;
        }
    }
    for (int k = 0; k < n; k++)
        for (int l = 0; l < n; l++) WS(m + k, m + l) = 0;

    // ---- LU with partial pivoting, in place ----
    int bad = 0;
    for (int k = 0; k < mpn; k++) {
        int p = k;
        double best = fabs(WS(k, k));
        for (int i = k + 1; i < mpn; i++) {
            double v = fabs(WS(i, k));
            if (v > best) { best = v; p = i; }
        }
        a.piv[(size_t)k * B + b] = p;
        if (best == 0) { bad = 1; break; }
        if (p != k)
            for (int j = 0; j < mpn; j++) {
                double t = WS(k, j);
                WS(k, j) = WS(p, j);
                WS(p, j) = t;
            }
        const double piv = WS(k, k);
        // rows in groups of 4 so each pivot-row element is loaded once per group
        int i = k + 1;
        for (; i + 3 < mpn; i += 4) {
            double l0 = WS(i, k) / piv, l1 = WS(i + 1, k) / piv, l2 = WS(i + 2, k) / piv,
                   l3 = WS(i + 3, k) / piv;
            WS(i, k) = l0; WS(i + 1, k) = l1; WS(i + 2, k) = l2; WS(i + 3, k) = l3;
            for (int j = k + 1; j < mpn; j++) {
                double u = WS(k, j);
                WS(i, j) -= l0 * u;
                WS(i + 1, j) -= l1 * u;
                WS(i + 2, j) -= l2 * u;
                WS(i + 3, j) -= l3 * u;
            }
        }
        for (; i < mpn; i++) {
            double l0 = WS(i, k) / piv;
            WS(i, k) = l0;
            for (int j = k + 1; j < mpn; j++) WS(i, j) -= l0 * WS(k, j);
        }
    }
    if (bad) { atomicOr(a.status, 1); return; }

    // ---- right-hand sides: operators, then interpolation targets ----
    int ops[4], nops = 0;
    if (a.Dx) ops[nops++] = 0;
    if (a.Dy) ops[nops++] = 1;
    if (a.Dz && dim == 3) ops[nops++] = 2;
    if (a.lap) ops[nops++] = 3;
    const int nr = a.rt_off ? a.rt_off[s + 1] - a.rt_off[s] : 0;
    const int np = a.pt_off ? a.pt_off[s + 1] - a.pt_off[s] : 0;
    const int total = nops + nr + np;
    for (int g0 = 0; g0 < total; g0 += 4) {
        const int gn = (total - g0 < 4) ? total - g0 : 4;
        for (int r = 0; r < 4; r++) {
            const int q = g0 + r;
            if (r >= gn) {
                for (int i = 0; i < mpn; i++) RHS(r, i) = 0;
            } else if (q < nops) {
                // row 0 of the operator's RHS matrix (x = seed)
                const int op = ops[q];
                for (int j = 0; j < m; j++) {
                    double s2 = dist_sum(sx, sy, sz, a.x[cl[j]], a.y[cl[j]], a.z[cl[j]], dim);
                    double rb = rbf_phs(s2, phs - 2);
                    if (op < 3) {
                        double d = (op == 0) ? sx - a.x[cl[j]] : (op == 1 ? sy - a.y[cl[j]] : sz - a.z[cl[j]]);
                        RHS(r, j) = phs * d * rb;
                    } else {
                        RHS(r, j) = (phs * phs) * rb;
                    }
                }
                for (int k = 0; k < n; k++) {
                    const int ex = a.px[k], ey = a.py[k], ez = a.pz[k];
                    double v = 0;
                    if (op == 0 && ex == 1 && ey == 0 && ez == 0) v = 1;
                    if (op == 1 && ex == 0 && ey == 1 && ez == 0) v = 1;
                    if (op == 2 && ex == 0 && ey == 0 && ez == 1) v = 1;
                    if (op == 3 && ((ex == 2 && ey == 0 && ez == 0) || (ex == 0 && ey == 2 && ez == 0) ||
                                    (ex == 0 && ey == 0 && ez == 2)))
                        v = 2;
                    RHS(r, m + k) = v;
                }
            } else {
                const double *t = (q < nops + nr) ? a.rt_xyz + 3 * (size_t)(a.rt_off[s] + q - nops)
                                                  : a.pt_xyz + 3 * (size_t)(a.pt_off[s] + q - nops - nr);
                for (int j = 0; j < m; j++)
                    RHS(r, j) = rbf_phs(dist_sum(t[0], t[1], t[2], a.x[cl[j]], a.y[cl[j]], a.z[cl[j]], dim), phs);
                const double dx = t[0] - sx, dy = t[1] - sy, dz = t[2] - sz;
                for (int k = 0; k < n; k++)
                    RHS(r, m + k) = ipow(dx, a.px[k]) * ipow(dy, a.py[k]) * ipow(dz, a.pz[k]);
            }
        }
        // P b
        for (int k = 0; k < mpn; k++) {
            int p = a.piv[(size_t)k * B + b];
            if (p != k)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    double t = RHS(r, k);
                    RHS(r, k) = RHS(r, p);
                    RHS(r, p) = t;
                }
        }
        // forward (unit lower)
        for (int k = 0; k < mpn; k++) {
            double bk[4];
#pragma unroll
            for (int r = 0; r < 4; r++) bk[r] = RHS(r, k);
            for (int i = k + 1; i < mpn; i++) {
                double l = WS(i, k);
#pragma unroll
                for (int r = 0; r < 4; r++) RHS(r, i) -= l * bk[r];
            }
        }
        // backward
        for (int k = mpn - 1; k >= 0; k--) {
            double ukk = WS(k, k);
            double bk[4];
#pragma unroll
            for (int r = 0; r < 4; r++) { bk[r] = RHS(r, k) / ukk; RHS(r, k) = bk[r]; }
            for (int i = 0; i < k; i++) {
                double u = WS(i, k);
#pragma unroll
                for (int r = 0; r < 4; r++) RHS(r, i) -= u * bk[r];
            }
        }
        for (int r = 0; r < gn; r++) {
            const int q = g0 + r;
            double *o;
            if (q < nops) {
                const int op = ops[q];
                o = (op == 0 ? a.Dx : op == 1 ? a.Dy : op == 2 ? a.Dz : a.lap) + (size_t)s * m;
            } else if (q < nops + nr) {
                o = a.R_out + (size_t)a.rt_row[a.rt_off[s] + q - nops] * m;
            } else {
                o = a.P_out + (size_t)a.pt_row[a.pt_off[s] + q - nops - nr] * m;
            }
            for (int j = 0; j < m; j++) o[j] = RHS(r, j);
        }
    }
#undef WS
#undef RHS
}

}  // namespace

int launch_weights(const WeightsArgs &a, hipStream_t s)
{
    for (int base = 0; base < a.nsys; base += a.batch) {
        int count = (a.nsys - base < a.batch) ? a.nsys - base : a.batch;
        hipLaunchKernelGGL(weights_kernel, dim3((count + 63) / 64), dim3(64), 0, s, a, base, count);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

}  // namespace mgm
