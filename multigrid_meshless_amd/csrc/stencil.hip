// stencil.hip -- FP64 fixed-width (ELL) stencil kernels of the V-cycle.
//
// Every operator row has exactly m entries stored column-major over the
// level's device node order (entry j of row r at [j*ld + r]), so the 64 lanes
// of a wave read 64 consecutive weights and indices per entry.  Device order
// groups nodes by colour and, within a colour, by Morton code, so the rows of
// one SOR colour are contiguous and neighbour gathers stay local.
//
// Arithmetic is the reference's, operation for operation: sums start from 0.0
// and add entries in stencil order j = 0..m-1, no FMA contraction, IEEE
// division.  With the same operators and the same sweep order the results are
// bitwise those of the reference loops cited at each kernel.
#pragma clang fp contract(off)

#include "device.hpp"

namespace mgm {

namespace {

constexpr int BLK = 256;

inline int nblk(long n) { return (int)((n + BLK - 1) / BLK); }

// multiply_sparse_matrix_vector(_gpu) (hf/mat_lib.h:70-89)
__global__ void __launch_bounds__(BLK)
spmv_kernel(Ell A, int r0, int r1, const double *__restrict__ x, double *__restrict__ y)
{
    int r = r0 + blockIdx.x * BLK + threadIdx.x;
    if (r >= r1) return;
    double acc = 0.0;
    for (int j = 0; j < A.m; j++) {
        size_t q = (size_t)j * A.ld + r;
        acc += A.w[q] * x[A.col[q]];
    }
    y[r] = acc;
}

// One SOR update of rows [r0, r1) (one colour class): the loop body of
// relaxation (hf/heat_conduction.h:29-34, hf/timple_solver.h:240-247,
// hf/fractional_step_solver.h:172-179).  Rows of a class never read each other.
__global__ void __launch_bounds__(BLK)
sor_rows_kernel(Ell A, int r0, int r1, const double *__restrict__ src, double *p, double w, double c)
{
    int r = r0 + blockIdx.x * BLK + threadIdx.x;
    if (r >= r1) return;
    double sum = 0.0;
    for (int j = 1; j < A.m; j++) {
        size_t q = (size_t)j * A.ld + r;
        sum += A.w[q] * p[A.col[q]];
    }
    p[r] = w * ((src[r] - sum) / A.w[r]) + c * p[r];
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
            size_t q = (size_t)j * A.ld + r;
            sum += A.w[q] * p[A.col[q]];
        }
        p[r] = w * ((src[r] - sum) / A.w[r]) + c * p[r];
    }
}

// calculate_residuals and the residual inside relaxation, four forms.
__global__ void __launch_bounds__(BLK)
residual_kernel(Ell A, int N, int form, const uint8_t *__restrict__ btag,
                const double *__restrict__ src, const double *__restrict__ p, double *__restrict__ res)
{
    int r = blockIdx.x * BLK + threadIdx.x;
    if (r >= N) return;
    bool bnd = btag[r] != 0;
    if (form == 0) {            // hf/heat_conduction.h:101-112, hf/timple_solver.h:274-290
        if (bnd) { res[r] = 0.0; return; }
        double sum = 0;
        for (int j = 0; j < A.m; j++) {
            size_t q = (size_t)j * A.ld + r;
            sum += A.w[q] * p[A.col[q]];
        }
        res[r] = src[r] - sum;
    } else if (form == 3) {     // hf/fractional_step_solver.h:236-246
        if (bnd) return;
        double sum = 0;
        for (int j = 0; j < A.m; j++) {
            size_t q = (size_t)j * A.ld + r;
            sum += A.w[q] * p[A.col[q]];
        }
        res[r] = src[r] - sum;
    } else {                    // form 1: hf/timple_solver.h:254-266; form 2: hf/fractional_step_solver.h:185-191
        if (form == 1 && bnd) { res[r] = 0; return; }
        double v = src[r];
        for (int j = 0; j < A.m; j++) {
            size_t q = (size_t)j * A.ld + r;
            v = v - A.w[q] * p[A.col[q]];
        }
        res[r] = v;
    }
}

// restrict_residuals (hf/heat_conduction.h:58-77; T :292-306; F :198-212);
// the restriction point's fine cloud is pre-composed into the column indices.
__global__ void __launch_bounds__(BLK)
restrict_kernel(Ell R, int N, const uint8_t *__restrict__ btag, int keep_boundary,
                const double *__restrict__ in, double *__restrict__ out)
{
    int r = blockIdx.x * BLK + threadIdx.x;
    if (r >= N) return;
    if (btag[r]) {
        if (!keep_boundary) out[r] = 0.0;
        return;
    }
    double acc = 0.0;
    for (int j = 0; j < R.m; j++) {
        size_t q = (size_t)j * R.ld + r;
        acc += R.w[q] * in[R.col[q]];
    }
    out[r] = acc;
}

// prolongate_corrections, first loop (hf/heat_conduction.h:82-92).
__global__ void __launch_bounds__(BLK)
prolong_kernel(Ell P, int N, const uint8_t *__restrict__ btag, const double *__restrict__ pc,
               double *__restrict__ p)
{
    int r = blockIdx.x * BLK + threadIdx.x;
    if (r >= N || btag[r]) return;
    double acc = 0.0;
    for (int j = 0; j < P.m; j++) {
        size_t q = (size_t)j * P.ld + r;
        acc += P.w[q] * pc[P.col[q]];
    }
    p[r] = p[r] + acc;
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

// Row-major reference-order rows -> column-major device-order slab.
__global__ void __launch_bounds__(BLK)
gather_rows_kernel(int N, int m, const double *__restrict__ rows, const int *__restrict__ perm,
                   const int *__restrict__ iperm, const int *__restrict__ cloud,
                   double *__restrict__ w_out, int *__restrict__ col_out)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t >= (long)N * m) return;
    int dev = (int)(t % N), j = (int)(t / N);
    int ref = perm[dev];
    if (w_out) w_out[t] = rows[(size_t)ref * m + j];
    if (col_out) col_out[t] = iperm[cloud[(size_t)ref * m + j]];
}

// Transfer rows (reference order, width m of the source level) -> slab over the
// target level's device order, columns pre-composed through the source cloud.
__global__ void __launch_bounds__(BLK)
transfer_slab_kernel(int Nt, int m, const double *__restrict__ rows, const int *__restrict__ perm,
                     const int *__restrict__ pts, const int *__restrict__ cloud_src,
                     const int *__restrict__ iperm_src, double *__restrict__ w_out,
                     int *__restrict__ col_out)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t >= (long)Nt * m) return;
    int dev = (int)(t % Nt), j = (int)(t / Nt);
    int ref = perm[dev];
    w_out[t] = rows[(size_t)ref * m + j];
    col_out[t] = iperm_src[cloud_src[(size_t)pts[ref] * m + j]];
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
    LAUNCH(spmv_kernel, r1 - r0, A, r0, r1, x, y);
    return 0;
}
int launch_sor_rows(Ell A, int r0, int r1, const double *src, double *p, double w, double c,
                    hipStream_t s)
{
    LAUNCH(sor_rows_kernel, r1 - r0, A, r0, r1, src, p, w, c);
    return 0;
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
    LAUNCH(residual_kernel, N, A, N, form, btag, src, p, res);
    return 0;
}
int launch_restrict(Ell R, int N, const uint8_t *btag, int keep_boundary, const double *in,
                    double *out, hipStream_t s)
{
    LAUNCH(restrict_kernel, N, R, N, btag, keep_boundary, in, out);
    return 0;
}
int launch_prolong(Ell P, int N, const uint8_t *btag, const double *pc, double *p, hipStream_t s)
{
    LAUNCH(prolong_kernel, N, P, N, btag, pc, p);
    return 0;
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
int launch_gather_rows(int N, int m, const double *rows, const int *perm, const int *iperm,
                       const int *cloud, double *w_out, int *col_out, hipStream_t s)
{
    LAUNCH(gather_rows_kernel, (long)N * m, N, m, rows, perm, iperm, cloud, w_out, col_out);
    return 0;
}
int launch_transfer_slab(int Nt, int m, const double *rows, const int *perm, const int *pts,
                         const int *cloud_src, const int *iperm_src, double *w_out, int *col_out,
                         hipStream_t s)
{
    LAUNCH(transfer_slab_kernel, (long)Nt * m, Nt, m, rows, perm, pts, cloud_src, iperm_src, w_out,
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
