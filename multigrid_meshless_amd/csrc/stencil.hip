// stencil.hip -- FP64 fixed-width (sliced-ELL) stencil kernels of the V-cycle.
//
// Layout: rows in slices of S consecutive device rows; within a slice the m
// columns follow each other and each column holds the slice's S entries
// contiguously (ell_at in device.hpp).  Device order groups nodes by colour
// and, within a colour, by Morton code, so the rows of one SOR colour are
// contiguous and neighbour gathers stay local.
//
// Work split: one wave per slice, 64/S lanes per row.  At step t the lanes of
// a wave load entries j = t*(64/S) + lane/S of the S rows -- 64 consecutive
// weights and 64 consecutive column indices -- and gather the corresponding
// x values, so a row's m loads are in flight together instead of forming one
// lane's dependent chain.  Each product w*x is rounded exactly as the
// reference rounds it and staged in LDS; one lane per row then adds the
// products in stencil order j = 0..m-1 starting from 0.0 (or subtracts them
// in order, for the residual forms that do).  No FMA contraction, IEEE
// division: with the same operators and sweep order the results are bitwise
// those of the reference loops cited at each mode.
#pragma clang fp contract(off)

#include "device.hpp"

namespace mgm {

namespace {

constexpr int BLK = 256;
constexpr int WAVES = BLK / 64;

inline int nblk(long n) { return (int)((n + BLK - 1) / BLK); }

enum Mode {
    M_SPMV = 0,      // y = sum_j w x                 hf/mat_lib.h:70-89
    M_SOR = 1,       // one SOR update                 hf/heat_conduction.h:29-34
    M_RES0 = 2,      // interior s - sum, bnd 0        hf/heat_conduction.h:101-112, timple_solver.h:274-290
    M_RES1 = 3,      // interior s -= in order, bnd 0  hf/timple_solver.h:254-266
    M_RES2 = 4,      // all rows s -= in order         hf/fractional_step_solver.h:185-191
    M_RES3 = 5,      // interior s - sum, bnd kept     hf/fractional_step_solver.h:236-246
    M_RESTRICT = 6,  // interior sum, bnd 0 or kept    hf/heat_conduction.h:58-77 (+T, F)
    M_PROLONG = 7,   // interior p += sum              hf/heat_conduction.h:82-92
};

__device__ inline bool skips_boundary(int mode)
{
    return mode == M_RES0 || mode == M_RES1 || mode == M_RES3 || mode == M_RESTRICT || mode == M_PROLONG;
}

template <int S>
__global__ void __launch_bounds__(BLK)
rows_kernel(Ell A, int r0, int r1, int mode, const uint8_t *__restrict__ btag,
            const double *__restrict__ src, const double *x, double *y, double w, double c,
            int keep_boundary)
{
    extern __shared__ double prod[];          // [WAVES][m][S]
    constexpr int LPR = 64 / S;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ri = lane % S, jo = lane / S;
    const int m = A.m;
    double *P = prod + (size_t)wave * m * S;
    const long slice = (long)(r0 / S) + (long)blockIdx.x * WAVES + wave;
    const long r = slice * S + ri;
    const bool act = (r >= r0 && r < r1);
    const bool bnd = act && btag && btag[r];
    const bool need = act && !(bnd && skips_boundary(mode));
    const size_t base = (size_t)slice * S * m;
    if (need) {
#pragma unroll 4
        for (int j = jo; j < m; j += LPR) {
            size_t q = base + (size_t)j * S + ri;
            P[j * S + ri] = A.w[q] * x[A.col[q]];
        }
    }
    __syncthreads();
    if (jo != 0 || !act) return;
    switch (mode) {
    case M_SPMV: {
        double acc = 0.0;
        for (int j = 0; j < m; j++) acc += P[j * S + ri];
        y[r] = acc;
        break;
    }
    case M_SOR: {
        double sum = 0.0;
        for (int j = 1; j < m; j++) sum += P[j * S + ri];
        y[r] = w * ((src[r] - sum) / A.w[base + ri]) + c * y[r];
        break;
    }
    case M_RES0:
    case M_RES3: {
        if (bnd) {
            if (mode == M_RES0) y[r] = 0.0;
            break;
        }
        double sum = 0;
        for (int j = 0; j < m; j++) sum += P[j * S + ri];
        y[r] = src[r] - sum;
        break;
    }
    case M_RES1:
    case M_RES2: {
        if (mode == M_RES1 && bnd) { y[r] = 0; break; }
        double v = src[r];
        for (int j = 0; j < m; j++) v = v - P[j * S + ri];
        y[r] = v;
        break;
    }
    case M_RESTRICT: {
        if (bnd) {
            if (!keep_boundary) y[r] = 0.0;
            break;
        }
        double acc = 0.0;
        for (int j = 0; j < m; j++) acc += P[j * S + ri];
        y[r] = acc;
        break;
    }
    case M_PROLONG: {
        if (bnd) break;
        double acc = 0.0;
        for (int j = 0; j < m; j++) acc += P[j * S + ri];
        y[r] = y[r] + acc;
        break;
    }
    }
}

// Sequential sweep in reference node order (the reference's host loop); one
// thread, for parity runs on small levels.
__global__ void sor_natural_kernel(Ell A, int N, const int *__restrict__ iperm,
                                   const double *__restrict__ src, double *p, double w, double c)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int i = 0; i < N; i++) {
        int r = iperm[i];
        double sum = 0.0;
        for (int j = 1; j < A.m; j++) {
            size_t q = ell_at(A.m, A.S, r, j);
            sum += A.w[q] * p[A.col[q]];
        }
        p[r] = w * ((src[r] - sum) / A.w[ell_at(A.m, A.S, r, 0)]) + c * p[r];
    }
}

__global__ void __launch_bounds__(BLK)
zero_kernel(int N, const uint8_t *__restrict__ btag, int interior_only, double *p)
{
    int r = blockIdx.x * BLK + threadIdx.x;
    if (r >= N) return;
    if (!interior_only || !btag[r]) p[r] = 0.0;
}

__global__ void copy_scalar_kernel(const int *ref0, const double *p, double *dst)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = p[*ref0];
}

// p -= p[node 0] (hf/timple_solver.h:248-252, hf/fractional_step_solver.h:180-183)
__global__ void __launch_bounds__(BLK)
pin_kernel(int N, const double *__restrict__ pref, double *p)
{
    int r = blockIdx.x * BLK + threadIdx.x;
    if (r >= N) return;
    p[r] = p[r] - *pref;
}

// Sum over interior rows of (a-b)^2: block partial sums, then one block.
__global__ void __launch_bounds__(BLK)
sqdiff_partial_kernel(int N, const uint8_t *__restrict__ btag, const double *__restrict__ a,
                      const double *__restrict__ b, double *__restrict__ partial)
{
    __shared__ double sh[BLK];
    double acc = 0;
    for (long r = (long)blockIdx.x * BLK + threadIdx.x; r < N; r += (long)gridDim.x * BLK) {
        if (btag[r]) continue;
        double d = a[r] - (b ? b[r] : 0.0);
        acc += d * d;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = BLK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(BLK)
sum_partial_kernel(int n, const double *__restrict__ partial, double *__restrict__ out)
{
    __shared__ double sh[BLK];
    double acc = 0;
    for (int i = threadIdx.x; i < n; i += BLK) acc += partial[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = BLK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// Row-major reference-order rows -> sliced-ELL device-order slab.
__global__ void __launch_bounds__(BLK)
gather_rows_kernel(int N, int m, int S, const double *__restrict__ rows, const int *__restrict__ perm,
                   const int *__restrict__ iperm, const int *__restrict__ cloud,
                   double *__restrict__ w_out, int *__restrict__ col_out)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t >= (long)N * m) return;
    int dev = (int)(t / m), j = (int)(t % m);
    int ref = perm[dev];
    size_t o = ell_at(m, S, dev, j);
    if (w_out) w_out[o] = rows[(size_t)ref * m + j];
    if (col_out) col_out[o] = iperm[cloud[(size_t)ref * m + j]];
}

// Transfer rows (reference order, width m of the source level) -> slab over the
// target level's device order, columns pre-composed through the source cloud.
__global__ void __launch_bounds__(BLK)
transfer_slab_kernel(int Nt, int m, int S, const double *__restrict__ rows, const int *__restrict__ perm,
                     const int *__restrict__ pts, const int *__restrict__ cloud_src,
                     const int *__restrict__ iperm_src, double *__restrict__ w_out,
                     int *__restrict__ col_out)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t >= (long)Nt * m) return;
    int dev = (int)(t / m), j = (int)(t % m);
    int ref = perm[dev];
    size_t o = ell_at(m, S, dev, j);
    w_out[o] = rows[(size_t)ref * m + j];
    col_out[o] = iperm_src[cloud_src[(size_t)pts[ref] * m + j]];
}

// create_laplacian_for_Poisson (hf/rbf.h:382-399) on row-major rows.
__global__ void __launch_bounds__(BLK)
lap_poisson_kernel(int N, int m, int dim, const double *__restrict__ Dx, const double *__restrict__ Dy,
                   const double *__restrict__ Dz, const double *__restrict__ lap,
                   const double *__restrict__ nx, const double *__restrict__ ny,
                   const double *__restrict__ nz, const uint8_t *__restrict__ btag,
                   double *__restrict__ lapP)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t >= (long)N * m) return;
    int i = (int)(t / m);
    if (!btag[i]) { lapP[t] = lap[t]; return; }
    double v = Dx[t] * nx[i] + Dy[t] * ny[i];
    if (dim == 3) v += Dz[t] * nz[i];
    lapP[t] = v;
}

__global__ void __launch_bounds__(BLK)
dirichlet_rows_kernel(int N, int m, const uint8_t *__restrict__ btag, double *__restrict__ rows)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t >= (long)N * m) return;
    int i = (int)(t / m), j = (int)(t % m);
    if (btag[i]) rows[t] = (j == 0) ? 1.0 : 0.0;
}

__global__ void __launch_bounds__(BLK)
permute_kernel(int N, const int *__restrict__ idx, const double *__restrict__ in, double *__restrict__ out)
{
    int t = blockIdx.x * BLK + threadIdx.x;
    if (t < N) out[t] = in[idx[t]];
}

__global__ void __launch_bounds__(BLK)
scatter_kernel(int N, const int *__restrict__ idx, const double *__restrict__ in, double *__restrict__ out)
{
    int t = blockIdx.x * BLK + threadIdx.x;
    if (t < N) out[idx[t]] = in[t];
}

int launch_rows(Ell A, int r0, int r1, int mode, const uint8_t *btag, const double *src,
                const double *x, double *y, double w, double c, int keep, hipStream_t s)
{
    if (r1 <= r0) return 0;
    long s0 = r0 / A.S, s1 = (r1 - 1) / A.S;
    int blocks = (int)((s1 - s0 + 1 + WAVES - 1) / WAVES);
    size_t shm = (size_t)WAVES * A.m * A.S * sizeof(double);
    if (A.S == 8)
        hipLaunchKernelGGL(rows_kernel<8>, dim3(blocks), dim3(BLK), shm, s, A, r0, r1, mode, btag, src,
                           x, y, w, c, keep);
    else if (A.S == 16)
        hipLaunchKernelGGL(rows_kernel<16>, dim3(blocks), dim3(BLK), shm, s, A, r0, r1, mode, btag, src,
                           x, y, w, c, keep);
    else
        MGM_FAIL(-1, "unsupported slice height");
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace

#define LAUNCH(kern, n, ...)                                                         \
    do {                                                                             \
        if ((n) > 0) {                                                               \
            hipLaunchKernelGGL(kern, dim3(nblk(n)), dim3(BLK), 0, s, __VA_ARGS__);   \
            HIP_TRY(hipGetLastError());                                              \
        }                                                                            \
    } while (0)

int launch_spmv(Ell A, int r0, int r1, const double *x, double *y, hipStream_t s)
{
    return launch_rows(A, r0, r1, M_SPMV, nullptr, nullptr, x, y, 0, 0, 0, s);
}
int launch_sor_rows(Ell A, int r0, int r1, const double *src, double *p, double w, double c,
                    hipStream_t s)
{
    return launch_rows(A, r0, r1, M_SOR, nullptr, src, p, p, w, c, 0, s);
}
int launch_sor_natural(Ell A, int N, const int *iperm, const double *src, double *p, double w,
                       double c, hipStream_t s)
{
    hipLaunchKernelGGL(sor_natural_kernel, dim3(1), dim3(64), 0, s, A, N, iperm, src, p, w, c);
    HIP_TRY(hipGetLastError());
    return 0;
}
int launch_residual(Ell A, int N, int form, const uint8_t *btag, const double *src,
                    const double *p, double *res, hipStream_t s)
{
    static const int modes[4] = {M_RES0, M_RES1, M_RES2, M_RES3};
    if (form < 0 || form > 3) MGM_FAIL(-1, "bad residual form");
    return launch_rows(A, 0, N, modes[form], btag, src, p, res, 0, 0, 0, s);
}
int launch_restrict(Ell R, int N, const uint8_t *btag, int keep_boundary, const double *in,
                    double *out, hipStream_t s)
{
    return launch_rows(R, 0, N, M_RESTRICT, btag, nullptr, in, out, 0, 0, keep_boundary, s);
}
int launch_prolong(Ell P, int N, const uint8_t *btag, const double *pc, double *p, hipStream_t s)
{
    return launch_rows(P, 0, N, M_PROLONG, btag, nullptr, pc, p, 0, 0, 0, s);
}
int launch_zero(int N, const uint8_t *btag, int interior_only, double *p, hipStream_t s)
{
    LAUNCH(zero_kernel, N, N, btag, interior_only, p);
    return 0;
}
int launch_pin(int N, const int *ref0_dev, double *p, double *scratch, hipStream_t s)
{
    hipLaunchKernelGGL(copy_scalar_kernel, dim3(1), dim3(64), 0, s, ref0_dev, p, scratch);
    HIP_TRY(hipGetLastError());
    LAUNCH(pin_kernel, N, N, scratch, p);
    return 0;
}
int launch_sqdiff_sum(int N, const uint8_t *btag, const double *a, const double *b,
                      double *partial, double *out, hipStream_t s)
{
    int nb = nblk(N);
    if (nb > 1024) nb = 1024;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(sqdiff_partial_kernel, dim3(nb), dim3(BLK), 0, s, N, btag, a, b, partial);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(sum_partial_kernel, dim3(1), dim3(BLK), 0, s, nb, partial, out);
    HIP_TRY(hipGetLastError());
    return 0;
}
int launch_sum_partials(int n, const double *partial, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(sum_partial_kernel, dim3(1), dim3(BLK), 0, s, n, partial, out);
    HIP_TRY(hipGetLastError());
    return 0;
}
int launch_gather_rows(int N, int m, int S, const double *rows, const int *perm, const int *iperm,
                       const int *cloud, double *w_out, int *col_out, hipStream_t s)
{
    LAUNCH(gather_rows_kernel, (long)N * m, N, m, S, rows, perm, iperm, cloud, w_out, col_out);
    return 0;
}
int launch_transfer_slab(int Nt, int m, int S, const double *rows, const int *perm, const int *pts,
                         const int *cloud_src, const int *iperm_src, double *w_out, int *col_out,
                         hipStream_t s)
{
    LAUNCH(transfer_slab_kernel, (long)Nt * m, Nt, m, S, rows, perm, pts, cloud_src, iperm_src, w_out,
           col_out);
    return 0;
}
int launch_lap_poisson(int N, int m, int dim, const double *Dx, const double *Dy, const double *Dz,
                       const double *lap, const double *nx, const double *ny, const double *nz,
                       const uint8_t *btag, double *lapP, hipStream_t s)
{
    LAUNCH(lap_poisson_kernel, (long)N * m, N, m, dim, Dx, Dy, Dz, lap, nx, ny, nz, btag, lapP);
    return 0;
}
int launch_dirichlet_rows(int N, int m, const uint8_t *btag, double *rows, hipStream_t s)
{
    LAUNCH(dirichlet_rows_kernel, (long)N * m, N, m, btag, rows);
    return 0;
}
int launch_permute(int N, const int *idx, const double *in, double *out, hipStream_t s)
{
    LAUNCH(permute_kernel, N, N, idx, in, out);
    return 0;
}
int launch_scatter(int N, const int *idx, const double *in, double *out, hipStream_t s)
{
    LAUNCH(scatter_kernel, N, N, idx, in, out);
    return 0;
}

}  // namespace mgm
