// ns.hip -- pointwise kernels of the two time steps that call the V-cycle:
// the explicit fractional step (hf/fractional_step_solver.h:47-303) and the
// TIMPLE step (hf/timple_solver.h:40-410).  FP64, no FMA contraction, the
// reference's expression order; stencil products use the ELL slabs.
#pragma clang fp contract(off)

#include "ns.hpp"

namespace mgm {

namespace {

constexpr int BLK = 256;
inline int nblk(long n) { return (int)((n + BLK - 1) / BLK); }

__global__ void __launch_bounds__(BLK) copy_kernel(int N, const double *__restrict__ a, double *__restrict__ b)
{
    int i = blockIdx.x * BLK + threadIdx.x;
    if (i < N) b[i] = a[i];
}

// out = c - dt*(u*dx + v*dy + w*dz - nu*lap)  (hf/fractional_step_solver.h:88, 99, 110)
__global__ void __launch_bounds__(BLK)
fs_predict_kernel(int N, const double *__restrict__ c, const double *__restrict__ u,
                  const double *__restrict__ v, const double *__restrict__ w,
                  const double *__restrict__ dx, const double *__restrict__ dy,
                  const double *__restrict__ dz, const double *__restrict__ dl, double dt, double nu,
                  double *__restrict__ out)
{
    int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= N) return;
    out[i] = c[i] - dt * (u[i] * dx[i] + v[i] * dy[i] + w[i] * dz[i] - nu * dl[i]);
}

// source = rho*(dpdx+dpdy+dpdz)/dt on interior rows; boundary rows set to 0
// (TIMPLE, hf/timple_solver.h:178-186) or left alone (FS, :128-130).
__global__ void __launch_bounds__(BLK)
mass_source_kernel(int N, const uint8_t *__restrict__ btag, const double *__restrict__ dx,
                   const double *__restrict__ dy, const double *__restrict__ dz, double rho, double dt,
                   int zero_boundary, double *__restrict__ src)
{
    int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= N) return;
    if (btag[i] == ROW_INTERIOR) src[i] = rho * (dx[i] + dy[i] + dz[i]) / dt;
    else if (zero_boundary) src[i] = 0;
}

// calculate_boundary_dpdn (hf/fractional_step_solver.h:133-148)
__global__ void __launch_bounds__(BLK)
fs_boundary_dpdn_kernel(int N, int dim, const uint8_t *__restrict__ btag, const double *__restrict__ u,
                        const double *__restrict__ v, const double *__restrict__ w,
                        const double *__restrict__ un, const double *__restrict__ vn,
                        const double *__restrict__ wn, const double *__restrict__ nx,
                        const double *__restrict__ ny, const double *__restrict__ nz, double rho,
                        double dt, double *__restrict__ dpdn)
{
    int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= N || btag[i] != ROW_BOUNDARY) return;
    double dpdx = (un[i] - u[i]) * rho / dt;
    double dpdy = (vn[i] - v[i]) * rho / dt;
    double r = dpdx * nx[i] + dpdy * ny[i];
    if (dim == 3) {
        double dpdz = (wn[i] - w[i]) * rho / dt;
        r += dpdz * nz[i];
    }
    dpdn[i] = r;
}

// update_velocity interior part (hf/fractional_step_solver.h:257-265)
__global__ void __launch_bounds__(BLK)
fs_correct_kernel(int N, int dim, const uint8_t *__restrict__ btag, const double *__restrict__ un,
                  const double *__restrict__ vn, const double *__restrict__ wn,
                  const double *__restrict__ dx, const double *__restrict__ dy,
                  const double *__restrict__ dz, double dt, double rho, double *__restrict__ u,
                  double *__restrict__ v, double *__restrict__ w)
{
    int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= N || btag[i]) return;
    u[i] = un[i] - dt * dx[i] / rho;
    v[i] = vn[i] - dt * dy[i] / rho;
    if (dim == 3) w[i] = wn[i] - dt * dz[i] / rho;
}

// update_velocity_implicit interior part (hf/timple_solver.h:342-351)
__global__ void __launch_bounds__(BLK)
timple_correct_kernel(int N, int dim, const uint8_t *__restrict__ btag, const double *__restrict__ dx,
                      const double *__restrict__ dy, const double *__restrict__ dz,
                      const double *__restrict__ pp, double dt, double rho, double *__restrict__ un,
                      double *__restrict__ vn, double *__restrict__ wn, double *__restrict__ pold)
{
    int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= N || btag[i]) return;
    un[i] = un[i] - dt * dx[i] / rho;
    vn[i] = vn[i] - dt * dy[i] / rho;
    if (dim == 3) wn[i] = wn[i] - dt * dz[i] / rho;
    pold[i] = pold[i] + 1.0 * pp[i];
}

// sum of squared velocity changes over all rows (steady-state error)
__global__ void __launch_bounds__(BLK)
vel_change_partial_kernel(int N, int dim, const double *__restrict__ a0, const double *__restrict__ a1,
                          const double *__restrict__ b0, const double *__restrict__ b1,
                          const double *__restrict__ c0, const double *__restrict__ c1,
                          double *__restrict__ partial)
{
    __shared__ double sh[BLK];
    double acc = 0;
    for (long i = (long)blockIdx.x * BLK + threadIdx.x; i < N; i += (long)gridDim.x * BLK) {
        double d1 = a1[i] - a0[i], d2 = b1[i] - b0[i];
        double t = d1 * d1 + d2 * d2;
        if (dim == 3) { double d3 = c1[i] - c0[i]; t += d3 * d3; }
        acc += t;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = BLK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__device__ inline void momentum_node(const NsDev &d, int r)
{
    // calculate_intermediate_velocity_implicit, one node (hf/timple_solver.h:92-160)
    const int m = d.m;
    const int S = d.S;
    const double rho = d.rho, mu = d.mu, dt = d.dt, facRe = d.facRe, facdt = d.facdt;
    const size_t q0 = ell_at(m, S, r, 0);
    double t1 = 0, t2 = 0, t3 = 0, t4 = 0;
    // u
    double adx = d.u[r] * d.Dx[q0], ady = d.v[r] * d.Dy[q0];
    double diff = -facRe * mu * d.lap[q0];
    double uns = 1. / (dt * facdt);
    double den = rho * (adx + ady + uns) + diff;
    for (int j = 1; j < m; j++) {
        size_t q = ell_at(m, S, r, j);
        double x = d.un[d.col[q]];
        t1 += d.Dx[q] * x;
        t2 += d.Dy[q] * x;
        t4 += d.lap[q] * x;
    }
    if (d.dim == 3) {
        den += rho * d.w[r] * d.Dz[q0];
        for (int j = 1; j < m; j++) {
            size_t q = ell_at(m, S, r, j);
            t3 += d.Dz[q] * d.un[d.col[q]];
        }
    }
    double adv = d.u[r] * t1 + d.v[r] * t2 + d.w[r] * t3;
    uns = d.uo[r] / (dt * facdt);
    d.un[r] = (rho * (uns - adv) - (facRe - 1) * mu * d.lap[q0] * d.un[r] + mu * t4 - d.dpdx[r]) / den;
    // v
    t1 = 0; t2 = 0; t3 = 0; t4 = 0;
    adx = d.u[r] * d.Dx[q0]; ady = d.v[r] * d.Dy[q0];
    diff = -facRe * mu * d.lap[q0];
    uns = 1. / (dt * facdt);
    den = rho * (adx + ady + uns) + diff;
    for (int j = 1; j < m; j++) {
        size_t q = ell_at(m, S, r, j);
        double x = d.vn[d.col[q]];
        t1 += d.Dx[q] * x;
        t2 += d.Dy[q] * x;
        t4 += d.lap[q] * x;
    }
    if (d.dim == 3) {
        den += rho * d.w[r] * d.Dz[q0];
        for (int j = 1; j < m; j++) {
            size_t q = ell_at(m, S, r, j);
            t3 += d.Dz[q] * d.vn[d.col[q]];
        }
    }
    adv = d.u[r] * t1 + d.v[r] * t2 + d.w[r] * t3;
    uns = d.vo[r] / (dt * facdt);
    d.vn[r] = (rho * (uns - adv) - (facRe - 1) * mu * d.lap[q0] * d.vn[r] + mu * t4 - d.dpdy[r]) / den;
    if (d.dim == 3) {
        // w
        t1 = 0; t2 = 0; t3 = 0; t4 = 0;
        adx = d.u[r] * d.Dx[q0]; ady = d.v[r] * d.Dy[q0];
        double adz = d.w[r] * d.Dz[q0];
        diff = -facRe * mu * d.lap[q0];
        uns = 1. / (dt * facdt);
        den = rho * (adx + ady + adz + uns) + diff;
        for (int j = 1; j < m; j++) {
            size_t q = ell_at(m, S, r, j);
            double x = d.wn[d.col[q]];
            t1 += d.Dx[q] * x;
            t2 += d.Dy[q] * x;
            t3 += d.Dz[q] * x;
            t4 += d.lap[q] * x;
        }
        adv = d.u[r] * t1 + d.v[r] * t2 + d.w[r] * t3;
        uns = d.wo[r] / (dt * facdt);
        d.wn[r] = (rho * (uns - adv) - (facRe - 1) * mu * d.lap[q0] * d.wn[r] + mu * t4 - d.dpdz[r]) / den;
    }
}

// Interior rows of the listed slices (one colour class), one thread per row.
__global__ void __launch_bounds__(BLK) momentum_slices_kernel(NsDev d, const int *__restrict__ slices,
                                                              int nslices)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    long k = t / d.S;
    if (k >= nslices) return;
    int r = slices[k] * d.S + (int)(t % d.S);
    if (d.btag[r] != ROW_INTERIOR) return;
    momentum_node(d, r);
}

__global__ void momentum_natural_kernel(NsDev d, int N, const int *__restrict__ iperm)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int i = 0; i < N; i++) {
        int r = iperm[i];
        if (d.btag[r] == ROW_INTERIOR) momentum_node(d, r);
    }
}

// update_boundary_pressure (hf/timple_solver.h:358-410).  In the reference
// temp1..3 are reassigned inside the stencil loop, so only the last stencil
// entry survives; boundary stencils hold only interior nodes after slot 0, so
// boundary rows never read each other and the loop is parallel.
__global__ void __launch_bounds__(BLK) timple_boundary_pressure_kernel(NsDev d, int N)
{
    int r = blockIdx.x * BLK + threadIdx.x;
    if (r >= N || d.btag[r] != ROW_BOUNDARY) return;
    const int m = d.m;
    const int S = d.S;
    const double nu = d.mu / d.rho;
    const size_t q0 = ell_at(m, S, r, 0);
    double t1 = 0, t2 = 0, t3 = 0;
    {
        size_t q = ell_at(m, S, r, m - 1);
        int c = d.col[q];
        t1 = nu * d.lap[q] * d.un[c];
        t1 -= d.u[r] * d.un[c] * d.Dx[q];
        t1 -= d.v[r] * d.un[c] * d.Dy[q];
        if (d.dim == 3) t1 -= d.w[r] * d.un[c] * d.Dz[q];
        t2 = nu * d.lap[q] * d.vn[c];
        t2 -= d.u[r] * d.vn[c] * d.Dx[q];
        t2 -= d.v[r] * d.vn[c] * d.Dy[q];
        if (d.dim == 3) t2 -= d.w[r] * d.vn[c] * d.Dz[q];
        if (d.dim == 3) {
            t3 = nu * d.lap[q] * d.wn[c];
            t3 -= d.u[r] * d.wn[c] * d.Dx[q];
            t3 -= d.v[r] * d.wn[c] * d.Dy[q];
            t3 -= d.w[r] * d.wn[c] * d.Dz[q];
        }
    }
    double dn = d.field_rho * (t1 * d.nx[r] + t2 * d.ny[r] + t3 * d.nz[r]);
    d.dpdn[r] = dn;
    double sx = 0, sy = 0, sz = 0, Ap = 0.0;
    for (int j = 1; j < m; j++) {
        size_t q = ell_at(m, S, r, j);
        double pj = d.pold[d.col[q]];
        sx += d.Dx[q] * pj;
        sy += d.Dy[q] * pj;
        if (d.dim == 3) sz += d.Dz[q] * pj;
    }
    Ap += d.Dx[q0] * d.nx[r];
    Ap += d.Dy[q0] * d.ny[r];
    if (d.dim == 3) Ap += d.Dz[q0] * d.nz[r];
    d.pold[r] = (dn - sx * d.nx[r] - sy * d.ny[r] - sz * d.nz[r]) / Ap;
}

}  // namespace

#define LAUNCH(kern, n, ...)                                                        \
    do {                                                                            \
        if ((n) > 0) {                                                              \
            hipLaunchKernelGGL(kern, dim3(nblk(n)), dim3(BLK), 0, s, __VA_ARGS__);  \
            HIP_TRY(hipGetLastError());                                             \
        }                                                                           \
    } while (0)

int launch_copy(int N, const double *a, double *b, hipStream_t s)
{
    LAUNCH(copy_kernel, N, N, a, b);
    return 0;
}
int launch_fs_predict(int N, const double *c, const double *u, const double *v, const double *w,
                      const double *dx, const double *dy, const double *dz, const double *dl,
                      double dt, double nu, double *out, hipStream_t s)
{
    LAUNCH(fs_predict_kernel, N, N, c, u, v, w, dx, dy, dz, dl, dt, nu, out);
    return 0;
}
int launch_mass_source(int N, const uint8_t *btag, const double *dx, const double *dy,
                       const double *dz, double rho, double dt, int zero_boundary, double *src,
                       hipStream_t s)
{
    LAUNCH(mass_source_kernel, N, N, btag, dx, dy, dz, rho, dt, zero_boundary, src);
    return 0;
}
int launch_fs_boundary_dpdn(int N, int dim, const uint8_t *btag, const double *u, const double *v,
                            const double *w, const double *un, const double *vn, const double *wn,
                            const double *nx, const double *ny, const double *nz, double rho,
                            double dt, double *dpdn, hipStream_t s)
{
    LAUNCH(fs_boundary_dpdn_kernel, N, N, dim, btag, u, v, w, un, vn, wn, nx, ny, nz, rho, dt, dpdn);
    return 0;
}
int launch_fs_correct(int N, int dim, const uint8_t *btag, const double *un, const double *vn,
                      const double *wn, const double *dx, const double *dy, const double *dz,
                      double dt, double rho, double *u, double *v, double *w, hipStream_t s)
{
    LAUNCH(fs_correct_kernel, N, N, dim, btag, un, vn, wn, dx, dy, dz, dt, rho, u, v, w);
    return 0;
}
int launch_timple_correct(int N, int dim, const uint8_t *btag, const double *dx, const double *dy,
                          const double *dz, const double *pp, double dt, double rho, double *un,
                          double *vn, double *wn, double *pold, hipStream_t s)
{
    LAUNCH(timple_correct_kernel, N, N, dim, btag, dx, dy, dz, pp, dt, rho, un, vn, wn, pold);
    return 0;
}
int launch_vel_change(int N, int dim, const double *a0, const double *a1, const double *b0,
                      const double *b1, const double *c0, const double *c1, double *partial,
                      double *out, hipStream_t s)
{
    int nb = nblk(N);
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(vel_change_partial_kernel, dim3(nb), dim3(BLK), 0, s, N, dim, a0, a1, b0, b1,
                       c0, c1, partial);
    HIP_TRY(hipGetLastError());
    return launch_sum_partials(nb, partial, out, s);
}
int launch_momentum_slices(const NsDev &d, const int *slices, int nslices, hipStream_t s)
{
    LAUNCH(momentum_slices_kernel, (long)nslices * d.S, d, slices, nslices);
    return 0;
}
int launch_momentum_natural(const NsDev &d, int N, const int *iperm, hipStream_t s)
{
    hipLaunchKernelGGL(momentum_natural_kernel, dim3(1), dim3(64), 0, s, d, N, iperm);
    HIP_TRY(hipGetLastError());
    return 0;
}
int launch_timple_boundary_pressure(const NsDev &d, int N, hipStream_t s)
{
    LAUNCH(timple_boundary_pressure_kernel, N, d, N);
    return 0;
}

}  // namespace mgm
