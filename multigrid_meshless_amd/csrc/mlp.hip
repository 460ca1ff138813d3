// mlp.hip -- the learned smoother's per-node network: training kernels
// (forward, backward, Adam, FP32) and the inference correction (BF16 weights,
// FP32 accumulation).  No reference counterpart: the reference smooths with
// SOR only.
//
// Network: in -> H -> H -> 1, ReLU after the two hidden layers, H = 64.
// Weight vector layout (FP32 masters): W1[H][in], b1[H], W2[H][H], b2[H],
// w3[H], b3.  Input of node i on a level with stencil width m:
//   [ (x_cj - x_i)/h for j < m, per axis d < dim ]  (m*dim values, h = level spacing)
//   [ r_cj / sigma_i for j < m ]                     (m values, sigma_i = RMS of r over the stencil)
// Output y; the correction is delta_i = y * sigma_i / a_ii.
#include <cstdlib>

#include "mlp.hpp"

namespace mgm {

namespace {

constexpr int BLK = 256;
inline int nblk(long n) { return (int)((n + BLK - 1) / BLK); }

__device__ inline float relu(float v) { return v > 0.f ? v : 0.f; }
__device__ inline float bf(bf16 v) { return __uint_as_float((unsigned)v.bits << 16); }

// ---- training ---------------------------------------------------------------

// One sample per thread: H1, H2 (stored for the backward pass) and y.
__global__ void __launch_bounds__(BLK)
fwd_kernel(int n, int in, const float *__restrict__ X, const float *__restrict__ W, float *__restrict__ H1,
           float *__restrict__ H2, float *__restrict__ y)
{
    const int b = blockIdx.x * BLK + threadIdx.x;
    if (b >= n) return;
    const float *W1 = W, *b1 = W1 + MLP_H * in, *W2 = b1 + MLP_H, *b2 = W2 + MLP_H * MLP_H;
    const float *w3 = b2 + MLP_H, *b3 = w3 + MLP_H;
    const float *x = X + (size_t)b * in;
    float h1[MLP_H];
    for (int k = 0; k < MLP_H; k++) {
        float a = b1[k];
        for (int f = 0; f < in; f++) a += W1[k * in + f] * x[f];
        h1[k] = relu(a);
        H1[(size_t)b * MLP_H + k] = h1[k];
    }
    float out = b3[0];
    for (int k = 0; k < MLP_H; k++) {
        float a = b2[k];
        for (int q = 0; q < MLP_H; q++) a += W2[k * MLP_H + q] * h1[q];
        const float h = relu(a);
        H2[(size_t)b * MLP_H + k] = h;
        out += w3[k] * h;
    }
    y[b] = out;
}

// Per-sample deltas of the loss L = mean_b (y_b - t_b)^2:
//   dy = 2 (y - t) / n, dZ2 = dy w3 [H2 > 0], dZ1 = W2^T dZ2 [H1 > 0].
__global__ void __launch_bounds__(BLK)
delta_kernel(int n, const float *__restrict__ W, int in, const float *__restrict__ H1,
             const float *__restrict__ H2, const float *__restrict__ y, const float *__restrict__ t,
             float *__restrict__ dy, float *__restrict__ dZ2, float *__restrict__ dZ1)
{
    const int b = blockIdx.x * BLK + threadIdx.x;
    if (b >= n) return;
    const float *W2 = W + MLP_H * in + MLP_H, *w3 = W2 + MLP_H * MLP_H + MLP_H;
    const float d = 2.f * (y[b] - t[b]) / (float)n;
    dy[b] = d;
    float z2[MLP_H];
    for (int k = 0; k < MLP_H; k++) {
        z2[k] = H2[(size_t)b * MLP_H + k] > 0.f ? d * w3[k] : 0.f;
        dZ2[(size_t)b * MLP_H + k] = z2[k];
    }
    for (int q = 0; q < MLP_H; q++) {
        float a = 0.f;
        for (int k = 0; k < MLP_H; k++) a += W2[k * MLP_H + q] * z2[k];
        dZ1[(size_t)b * MLP_H + q] = H1[(size_t)b * MLP_H + q] > 0.f ? a : 0.f;
    }
}

// Weight gradients, one weight per thread, summed over the samples:
//   dW1[k][f] = sum_b dZ1[b][k] X[b][f]   db1[k] = sum_b dZ1[b][k]
//   dW2[k][q] = sum_b dZ2[b][k] H1[b][q]  db2[k] = sum_b dZ2[b][k]
//   dw3[k]    = sum_b dy[b] H2[b][k]      db3    = sum_b dy[b]
__global__ void __launch_bounds__(BLK)
wgrad_kernel(int n, int in, const float *__restrict__ X, const float *__restrict__ H1,
             const float *__restrict__ H2, const float *__restrict__ dy, const float *__restrict__ dZ2,
             const float *__restrict__ dZ1, float *__restrict__ G)
{
    const int nw1 = MLP_H * in, o_b1 = nw1, o_w2 = o_b1 + MLP_H, o_b2 = o_w2 + MLP_H * MLP_H;
    const int o_w3 = o_b2 + MLP_H, o_b3 = o_w3 + MLP_H, total = o_b3 + 1;
    const int g = blockIdx.x * BLK + threadIdx.x;
    if (g >= total) return;
    float a = 0.f;
    if (g < o_b1) {
        const int k = g / in, f = g % in;
        for (int b = 0; b < n; b++) a += dZ1[(size_t)b * MLP_H + k] * X[(size_t)b * in + f];
    } else if (g < o_w2) {
        const int k = g - o_b1;
        for (int b = 0; b < n; b++) a += dZ1[(size_t)b * MLP_H + k];
    } else if (g < o_b2) {
        const int k = (g - o_w2) / MLP_H, q = (g - o_w2) % MLP_H;
        for (int b = 0; b < n; b++) a += dZ2[(size_t)b * MLP_H + k] * H1[(size_t)b * MLP_H + q];
    } else if (g < o_w3) {
        const int k = g - o_b2;
        for (int b = 0; b < n; b++) a += dZ2[(size_t)b * MLP_H + k];
    } else if (g < o_b3) {
        const int k = g - o_w3;
        for (int b = 0; b < n; b++) a += dy[b] * H2[(size_t)b * MLP_H + k];
    } else {
        for (int b = 0; b < n; b++) a += dy[b];
    }
    G[g] = a;
}

__global__ void __launch_bounds__(BLK)
sqerr_partial_kernel(int n, const float *__restrict__ y, const float *__restrict__ t, double *__restrict__ partial)
{
    __shared__ double sh[BLK];
    double a = 0;
    for (int b = blockIdx.x * BLK + threadIdx.x; b < n; b += gridDim.x * BLK) {
        const double d = (double)y[b] - (double)t[b];
        a += d * d;
    }
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int s = BLK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// Adam (Kingma & Ba 2015), bias-corrected step sizes passed in.
__global__ void __launch_bounds__(BLK)
adam_kernel(int nw, float *__restrict__ W, const float *__restrict__ G, float *__restrict__ M,
            float *__restrict__ V, float lr_t, float b1, float b2, float eps)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= nw) return;
    const float g = G[i];
    const float m = b1 * M[i] + (1.f - b1) * g;
    const float v = b2 * V[i] + (1.f - b2) * g * g;
    M[i] = m;
    V[i] = v;
    W[i] -= lr_t * m / (sqrtf(v) + eps);
}

__global__ void __launch_bounds__(BLK) diff_kernel(int n, const double *__restrict__ a, const double *__restrict__ b,
                                                    double *__restrict__ out)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i < n) out[i] = a[i] - b[i];
}

// Row gather of a training batch: Xb[i] = X[idx[i]], tb[i] = t[idx[i]].
__global__ void __launch_bounds__(BLK) gather_batch_kernel(int n, int in, const int *__restrict__ idx,
                                                           const float *__restrict__ X, const float *__restrict__ t,
                                                           float *__restrict__ Xb, float *__restrict__ tb)
{
    const long g = (long)blockIdx.x * BLK + threadIdx.x;
    if (g >= (long)n * in) return;
    const int i = (int)(g / in), f = (int)(g % in);
    Xb[g] = X[(size_t)idx[i] * in + f];
    if (f == 0) tb[i] = t[idx[i]];
}

__global__ void __launch_bounds__(BLK) to_bf16_kernel(int n, const float *__restrict__ in, bf16 *__restrict__ out)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    unsigned u = __float_as_uint(in[i]);
    u += 0x7fffu + ((u >> 16) & 1u);
    out[i].bits = (uint16_t)(u >> 16);
}

// ---- features ---------------------------------------------------------------

// Stencil residuals of device row r scaled by their RMS; returns sigma (0 if
// the stencil residual vanishes, then u is left 0).
__device__ inline float stencil_residual(const Ell &A, int r, const double *__restrict__ res, float *u)
{
    double s2 = 0;
    for (int j = 0; j < A.m; j++) {
        const double v = res[A.col[ell_at(A.m, A.S, r, j)]];
        u[j] = (float)v;
        s2 += v * v;
    }
    const double sigma = sqrt(s2 / A.m);
    const float inv = sigma > 0 ? (float)(1.0 / sigma) : 0.f;
    for (int j = 0; j < A.m; j++) u[j] *= inv;
    return (float)sigma;
}

// Training rows: X[b] = [offsets, scaled residuals] of device row rows[b] in
// snapshot snap[b]; target t[b] = e * a_ii / sigma (0 where sigma = 0).
__global__ void __launch_bounds__(BLK)
features_kernel(int n, Ell A, int dim, double h, const double *__restrict__ dx, const double *__restrict__ dy,
                const double *__restrict__ dz, const int *__restrict__ rows, const int *__restrict__ snap,
                int Nd, const double *__restrict__ res_snaps, const double *__restrict__ err_snaps,
                float *__restrict__ X, float *__restrict__ t)
{
    const int b = blockIdx.x * BLK + threadIdx.x;
    if (b >= n) return;
    const int m = A.m, in = m * dim + m;
    const int r = rows[b];
    const double *res = res_snaps + (size_t)snap[b] * Nd;
    float *x = X + (size_t)b * in;
    const double *crd[3] = {dx, dy, dz};
    for (int j = 0; j < m; j++) {
        const int c = A.col[ell_at(m, A.S, r, j)];
        for (int d = 0; d < dim; d++) x[j * dim + d] = (float)((crd[d][c] - crd[d][r]) / h);
    }
    const float sigma = stencil_residual(A, r, res, x + m * dim);
    const double diag = A.w[ell_at(m, A.S, r, 0)];
    t[b] = sigma > 0.f ? (float)(err_snaps[(size_t)snap[b] * Nd + r] * diag / (double)sigma) : 0.f;
}

// Per-row constant part of layer 1: G[r][k] = b1[k] + sum_{j,d} W1[k][j*dim+d] off_jd.
__global__ void __launch_bounds__(BLK)
geom_kernel(int Nd, Ell A, int dim, double h, const double *__restrict__ dx, const double *__restrict__ dy,
            const double *__restrict__ dz, const uint8_t *__restrict__ btag, const float *__restrict__ W,
            float *__restrict__ G)
{
    const int r = blockIdx.x * BLK + threadIdx.x;
    if (r >= Nd || btag[r] != ROW_INTERIOR) return;
    const int m = A.m, in = m * dim + m;
    const float *b1 = W + MLP_H * in;
    const double *crd[3] = {dx, dy, dz};
    float acc[MLP_H];
#pragma unroll
    for (int k = 0; k < MLP_H; k++) acc[k] = b1[k];
    for (int j = 0; j < m; j++) {
        const int c = A.col[ell_at(m, A.S, r, j)];
        for (int d = 0; d < dim; d++) {
            const float o = (float)((crd[d][c] - crd[d][r]) / h);
            const int f = j * dim + d;
#pragma unroll
            for (int k = 0; k < MLP_H; k++) acc[k] += W[k * in + f] * o;
        }
    }
#pragma unroll
    for (int k = 0; k < MLP_H; k++) G[(size_t)r * MLP_H + k] = acc[k];
}

// ---- inference ----------------------------------------------------------------

// p[r] += y * sigma / a_rr on interior rows, y = network(offsets, scaled
// stencil residuals) with the geometric part of layer 1 taken from G.
// BF16 weights W1rT[j][k] (residual columns of W1, transposed), W2[k][q],
// w3[k] are widened into LDS once per block (rows padded to 65 floats so the
// threads of one node read different banks); FP32 biases and accumulation.
// T threads share a node: T = 1 for large levels (the node's work in one
// thread's registers), T = 16 for small ones (more threads in flight): the
// node's stencil residuals go to LDS, each thread computes 64/T hidden units
// of each layer, and the output is summed across the T lanes.
constexpr int MLP_LD = MLP_H + 1;
template <typename TV, int T>
__global__ void __launch_bounds__(BLK)
correct_kernel(int Nd, Ell A, const uint8_t *__restrict__ btag, const float *__restrict__ G,
               const bf16 *__restrict__ W1rT, const bf16 *__restrict__ W2, const float *__restrict__ b2,
               const bf16 *__restrict__ w3, float b3, const TV *__restrict__ res, TV *__restrict__ p)
{
    constexpr int NPB = BLK / T;            // nodes per block
    constexpr int KPT = MLP_H / T;          // hidden units per thread
    __shared__ float sW1[MLP_MAX_M * MLP_LD], sW2[MLP_H * MLP_LD], sw3[MLP_H], sb2[MLP_H];
    __shared__ double sR[T > 1 ? NPB : 1][MLP_MAX_M + 1];
    __shared__ float sH[T > 1 ? NPB : 1][MLP_LD];
    const int m = A.m;
    for (int i = threadIdx.x; i < m * MLP_H; i += BLK) sW1[(i / MLP_H) * MLP_LD + i % MLP_H] = bf(W1rT[i]);
    for (int i = threadIdx.x; i < MLP_H * MLP_H; i += BLK) sW2[(i / MLP_H) * MLP_LD + i % MLP_H] = bf(W2[i]);
    if (threadIdx.x < MLP_H) {
        sw3[threadIdx.x] = bf(w3[threadIdx.x]);
        sb2[threadIdx.x] = b2[threadIdx.x];
    }
    __syncthreads();
    const int nl = threadIdx.x / T, sub = threadIdx.x % T;
    const int r = blockIdx.x * NPB + nl;
    const bool valid = r < Nd && btag[r] == ROW_INTERIOR;
    if constexpr (T == 1) {
        if (!valid) return;
        double s2 = 0;
        for (int j = 0; j < m; j++) {
            const double v = res[A.col[ell_at(m, A.S, r, j)]];
            s2 += v * v;
        }
        const double sigma = sqrt(s2 / m);
        if (!(sigma > 0)) return;
        const double inv = 1.0 / sigma;
        float h1[MLP_H];
#pragma unroll
        for (int k = 0; k < MLP_H; k++) h1[k] = G[(size_t)r * MLP_H + k];
        for (int j = 0; j < m; j++) {
            const float u = (float)(res[A.col[ell_at(m, A.S, r, j)]] * inv);
#pragma unroll
            for (int k = 0; k < MLP_H; k++) h1[k] += sW1[j * MLP_LD + k] * u;
        }
#pragma unroll
        for (int k = 0; k < MLP_H; k++) h1[k] = relu(h1[k]);
        float y = b3;
        for (int k = 0; k < MLP_H; k++) {
            float a = sb2[k];
#pragma unroll
            for (int q = 0; q < MLP_H; q++) a += sW2[k * MLP_LD + q] * h1[q];
            y += sw3[k] * relu(a);
        }
        const double diag = A.w[ell_at(m, A.S, r, 0)];
        p[r] = p[r] + (TV)((double)y * sigma / diag);
    } else {
        // stencil residuals of the node, split over its T lanes
        double s2 = 0;
        if (valid)
            for (int j = sub; j < m; j += T) {
                const double v = res[A.col[ell_at(m, A.S, r, j)]];
                sR[nl][j] = v;
                s2 += v * v;
            }
#pragma unroll
        for (int o = T / 2; o > 0; o /= 2) s2 += __shfl_xor(s2, o, T);
        const double sigma = sqrt(s2 / m);
        const bool go = valid && sigma > 0;
        const double inv = go ? 1.0 / sigma : 0.0;
        __syncthreads();
        if (go) {
            float h[KPT];
#pragma unroll
            for (int i = 0; i < KPT; i++) h[i] = G[(size_t)r * MLP_H + sub * KPT + i];
            for (int j = 0; j < m; j++) {
                const float u = (float)(sR[nl][j] * inv);
#pragma unroll
                for (int i = 0; i < KPT; i++) h[i] += sW1[j * MLP_LD + sub * KPT + i] * u;
            }
#pragma unroll
            for (int i = 0; i < KPT; i++) sH[nl][sub * KPT + i] = relu(h[i]);
        }
        __syncthreads();
        float y = 0.f;
        if (go) {
#pragma unroll
            for (int i = 0; i < KPT; i++) {
                const int k = sub * KPT + i;
                float a = sb2[k];
                for (int q = 0; q < MLP_H; q++) a += sW2[k * MLP_LD + q] * sH[nl][q];
                y += sw3[k] * relu(a);
            }
        }
#pragma unroll
        for (int o = T / 2; o > 0; o /= 2) y += __shfl_xor(y, o, T);
        if (go && sub == 0) {
            const double diag = A.w[ell_at(m, A.S, r, 0)];
            p[r] = p[r] + (TV)((double)(y + b3) * sigma / diag);
        }
    }
}

}  // namespace

#define LAUNCH(kern, n, ...)                                                        \
    do {                                                                            \
        if ((n) > 0) {                                                              \
            hipLaunchKernelGGL(kern, dim3(nblk(n)), dim3(BLK), 0, s, __VA_ARGS__);  \
            HIP_TRY(hipGetLastError());                                             \
        }                                                                           \
    } while (0)

size_t mlp_weight_count(int in) { return (size_t)MLP_H * in + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1; }

int mlp_forward(int n, int in, const float *X, const float *W, float *H1, float *H2, float *y, hipStream_t s)
{
    LAUNCH(fwd_kernel, n, n, in, X, W, H1, H2, y);
    return 0;
}

int mlp_gradient(int n, int in, const float *X, const float *t, const float *W, const float *H1, const float *H2,
                 const float *y, float *dy, float *dZ2, float *dZ1, float *G, hipStream_t s)
{
    LAUNCH(delta_kernel, n, n, W, in, H1, H2, y, t, dy, dZ2, dZ1);
    LAUNCH(wgrad_kernel, (long)mlp_weight_count(in), n, in, X, H1, H2, dy, dZ2, dZ1, G);
    return 0;
}

int mlp_sqerr(int n, const float *y, const float *t, double *partial, double *out, hipStream_t s)
{
    int nb = nblk(n);
    if (nb > 1024) nb = 1024;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(sqerr_partial_kernel, dim3(nb), dim3(BLK), 0, s, n, y, t, partial);
    HIP_TRY(hipGetLastError());
    return launch_sum_partials(nb, partial, out, s);
}

int mlp_adam(int nw, float *W, const float *G, float *M, float *V, float lr_t, float b1, float b2, float eps,
             hipStream_t s)
{
    LAUNCH(adam_kernel, nw, nw, W, G, M, V, lr_t, b1, b2, eps);
    return 0;
}

int mlp_diff(int n, const double *a, const double *b, double *out, hipStream_t s)
{
    LAUNCH(diff_kernel, n, n, a, b, out);
    return 0;
}

int mlp_gather_batch(int n, int in, const int *idx, const float *X, const float *t, float *Xb, float *tb,
                     hipStream_t s)
{
    LAUNCH(gather_batch_kernel, (long)n * in, n, in, idx, X, t, Xb, tb);
    return 0;
}

int mlp_to_bf16(int n, const float *in, bf16 *out, hipStream_t s)
{
    LAUNCH(to_bf16_kernel, n, n, in, out);
    return 0;
}

int mlp_features(int n, Ell A, int dim, double h, const double *dx, const double *dy, const double *dz,
                 const int *rows, const int *snap, int Nd, const double *res_snaps, const double *err_snaps,
                 float *X, float *t, hipStream_t s)
{
    if (A.m > MLP_MAX_M) MGM_FAIL(-1, "learned smoother: stencil wider than supported");
    LAUNCH(features_kernel, n, n, A, dim, h, dx, dy, dz, rows, snap, Nd, res_snaps, err_snaps, X, t);
    return 0;
}

int mlp_geometry(int Nd, Ell A, int dim, double h, const double *dx, const double *dy, const double *dz,
                 const uint8_t *btag, const float *W, float *G, hipStream_t s)
{
    if (A.m > MLP_MAX_M) MGM_FAIL(-1, "learned smoother: stencil wider than supported");
    LAUNCH(geom_kernel, Nd, Nd, A, dim, h, dx, dy, dz, btag, W, G);
    return 0;
}

template <typename TV>
int mlp_correct(int Nd, Ell A, const uint8_t *btag, const float *G, const bf16 *W1rT, const bf16 *W2,
                const float *b2, const bf16 *w3, float b3, const TV *res, TV *p, hipStream_t s)
{
    if (A.m > MLP_MAX_M) MGM_FAIL(-1, "learned smoother: stencil wider than supported");
    const char *e = getenv("MGM_MLP_WIDE_ROWS");
    const int wide = e ? atoi(e) : 200000;   // levels with more rows: one thread per node
    if (Nd > wide) {
        LAUNCH((correct_kernel<TV, 1>), Nd, Nd, A, btag, G, W1rT, W2, b2, w3, b3, res, p);
    } else {
        LAUNCH((correct_kernel<TV, 16>), (long)Nd * 16, Nd, A, btag, G, W1rT, W2, b2, w3, b3, res, p);
    }
    return 0;
}
template int mlp_correct<double>(int, Ell, const uint8_t *, const float *, const bf16 *, const bf16 *,
                                 const float *, const bf16 *, float, const double *, double *, hipStream_t);
template int mlp_correct<float>(int, Ell, const uint8_t *, const float *, const bf16 *, const bf16 *,
                                const float *, const bf16 *, float, const float *, float *, hipStream_t);

}  // namespace mgm
