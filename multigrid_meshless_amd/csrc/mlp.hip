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
#include <algorithm>
#include <cstdlib>

#include "mlp.hpp"

namespace mgm {

namespace {

constexpr int BLK = 256;
inline int nblk(long n) { return (int)((n + BLK - 1) / BLK); }

__device__ inline float relu(float v) { return v > 0.f ? v : 0.f; }
__device__ inline float bf(bf16 v) { return __uint_as_float((unsigned)v.bits << 16); }

// ---- training ---------------------------------------------------------------

// Forward pass storing H1, H2 (for the backward pass) and y.  A block takes
// FWD_BS samples, their inputs staged in LDS transposed (xs[f][sample]); the
// thread for hidden unit k and sample group g computes k for four samples,
// streaming row k of W1 / W2 through L1.  Each unit's sum runs in input order
// from its bias.
constexpr int FWD_BS = 16;
constexpr int FWD_MAX_IN = MLP_MAX_M * 3 + MLP_MAX_M;
__global__ void __launch_bounds__(BLK)
fwd_kernel(int n, int in, const float *__restrict__ X, const float *__restrict__ W, float *__restrict__ H1,
           float *__restrict__ H2, float *__restrict__ y)
{
    __shared__ __attribute__((aligned(16))) float xs[FWD_MAX_IN][FWD_BS];
    __shared__ __attribute__((aligned(16))) float hs[FWD_BS][MLP_H];
    const float *W1 = W, *b1 = W1 + MLP_H * in, *W2 = b1 + MLP_H, *b2 = W2 + MLP_H * MLP_H;
    const float *w3 = b2 + MLP_H, *b3 = w3 + MLP_H;
    const long b0 = (long)blockIdx.x * FWD_BS;
    const int nb = (int)((n - b0 < FWD_BS) ? n - b0 : FWD_BS);
    for (int q = threadIdx.x; q < FWD_BS * in; q += BLK) {
        const int sm = q / in, f = q % in;
        xs[f][sm] = sm < nb ? X[(b0 + sm) * in + f] : 0.f;
    }
    __syncthreads();
    const int k = threadIdx.x & 63, g = threadIdx.x >> 6;   // wave g: samples 4g .. 4g+3
    float a[4];
#pragma unroll
    for (int v = 0; v < 4; v++) a[v] = b1[k];
    const float *w1 = W1 + (size_t)k * in;
    for (int f = 0; f < in; f++) {
        const float w = w1[f];
        const float4 xv = *reinterpret_cast<const float4 *>(&xs[f][4 * g]);
        a[0] += w * xv.x;
        a[1] += w * xv.y;
        a[2] += w * xv.z;
        a[3] += w * xv.w;
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int sm = 4 * g + v;
        const float h = relu(a[v]);
        hs[sm][k] = h;
        if (sm < nb) H1[(b0 + sm) * MLP_H + k] = h;
        a[v] = b2[k];
    }
    __syncthreads();
    const float *w2 = W2 + (size_t)k * MLP_H;
    for (int q = 0; q < MLP_H; q++) {
        const float w = w2[q];
#pragma unroll
        for (int v = 0; v < 4; v++) a[v] += w * hs[4 * g + v][q];
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int sm = 4 * g + v;
        const float h = relu(a[v]);
        if (sm < nb) H2[(b0 + sm) * MLP_H + k] = h;
        float o = w3[k] * h;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) o += __shfl_xor(o, off);
        if (k == 0 && sm < nb) y[b0 + sm] = b3[0] + o;
    }
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

// Weight gradients summed over the samples, in WG_SPLIT chunks of samples
// (partial sums part[c][g], one weight g per thread) and then over the chunks
// in order:
//   dW1[k][f] = sum_b dZ1[b][k] X[b][f]   db1[k] = sum_b dZ1[b][k]
//   dW2[k][q] = sum_b dZ2[b][k] H1[b][q]  db2[k] = sum_b dZ2[b][k]
//   dw3[k]    = sum_b dy[b] H2[b][k]      db3    = sum_b dy[b]
// Consecutive threads take consecutive inputs f (or q) of one unit k, so the
// sample rows are read contiguously and the delta is one broadcast value.
constexpr int WG_SPLIT = 32;
__global__ void __launch_bounds__(BLK)
wgrad_part_kernel(int n, int in, const float *__restrict__ X, const float *__restrict__ H1,
                  const float *__restrict__ H2, const float *__restrict__ dy, const float *__restrict__ dZ2,
                  const float *__restrict__ dZ1, float *__restrict__ part)
{
    const int nw1 = MLP_H * in, o_b1 = nw1, o_w2 = o_b1 + MLP_H, o_b2 = o_w2 + MLP_H * MLP_H;
    const int o_w3 = o_b2 + MLP_H, o_b3 = o_w3 + MLP_H, total = o_b3 + 1;
    const int g = blockIdx.x * BLK + threadIdx.x;
    const int c = blockIdx.y;
    if (g >= total) return;
    const int chunk = (n + WG_SPLIT - 1) / WG_SPLIT;
    const int bs = c * chunk, be = (bs + chunk < n) ? bs + chunk : n;
    float a = 0.f;
    if (g < o_b1) {
        const int k = g / in, f = g % in;
        for (int b = bs; b < be; b++) a += dZ1[(size_t)b * MLP_H + k] * X[(size_t)b * in + f];
    } else if (g < o_w2) {
        const int k = g - o_b1;
        for (int b = bs; b < be; b++) a += dZ1[(size_t)b * MLP_H + k];
    } else if (g < o_b2) {
        const int k = (g - o_w2) / MLP_H, q = (g - o_w2) % MLP_H;
        for (int b = bs; b < be; b++) a += dZ2[(size_t)b * MLP_H + k] * H1[(size_t)b * MLP_H + q];
    } else if (g < o_w3) {
        const int k = g - o_b2;
        for (int b = bs; b < be; b++) a += dZ2[(size_t)b * MLP_H + k];
    } else if (g < o_b3) {
        const int k = g - o_w3;
        for (int b = bs; b < be; b++) a += dy[b] * H2[(size_t)b * MLP_H + k];
    } else {
        for (int b = bs; b < be; b++) a += dy[b];
    }
    part[(size_t)c * total + g] = a;
}

__global__ void __launch_bounds__(BLK)
wgrad_sum_kernel(int total, const float *__restrict__ part, float *__restrict__ G)
{
    const int g = blockIdx.x * BLK + threadIdx.x;
    if (g >= total) return;
    float a = 0.f;
    for (int c = 0; c < WG_SPLIT; c++) a += part[(size_t)c * total + g];
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
// snapshot snap[b] (snapshot 0 when snap is null); target t[b] = e * a_ii /
// sigma (0 where sigma = 0).
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
    const size_t so = snap ? (size_t)snap[b] * Nd : 0;
    const double *res = res_snaps + so;
    float *x = X + (size_t)b * in;
    const double *crd[3] = {dx, dy, dz};
    for (int j = 0; j < m; j++) {
        const int c = A.col[ell_at(m, A.S, r, j)];
        for (int d = 0; d < dim; d++) x[j * dim + d] = (float)((crd[d][c] - crd[d][r]) / h);
    }
    const float sigma = stencil_residual(A, r, res, x + m * dim);
    const double diag = A.w[ell_at(m, A.S, r, 0)];
    t[b] = sigma > 0.f ? (float)(err_snaps[so + r] * diag / (double)sigma) : 0.f;
}

// Per-row constant part of layer 1: b1[k] + sum_{j,d} W1[k][j*dim+d] off_jd,
// stored in 16-row tiles, hidden-unit-major inside a tile: row r, unit k at
// G[geo_at(r, k)].  The correction reads four consecutive rows of one hidden
// unit per lane (16 bytes), and a 16-row tile's 64 units are 4 KB in one run.
__host__ __device__ inline size_t geo_at(long r, int k)
{
    return (size_t)(r >> 4) * (16 * MLP_H) + (size_t)k * 16 + (size_t)(r & 15);
}
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
    for (int k = 0; k < MLP_H; k++) G[geo_at(r, k)] = acc[k];
}

// ---- inference ----------------------------------------------------------------

// p[r] += y * sigma / a_rr on interior rows, y = network(offsets, scaled
// stencil residuals) with the geometric part of layer 1 taken from G.  Both
// layers run on BF16 matrix cores (v_mfma_f32_16x16x32_bf16): BF16 weights
// and activations, FP32 accumulation.  A wave takes tiles of 16 nodes:
//   layer 1  C[16 x 64] = G tile + U[16 x K] W1r[K x 64], K = m residual
//            columns padded to 32-multiples, U = BF16(r_cj / sigma);
//   layer 2  C2 = b2 + BF16(relu(C)) W2^T (C moved through LDS into the
//            operand layout); y = b3 + w3 . relu(C2), summed across lanes.
// Operand layouts (lane l): A[i][k] at i = l%16, k = (l/16)*8 + e; B[k][n] at
// k = (l/16)*8 + e, n = l%16; C[i][n] at n = l%16, i = (l/16)*4 + v.  The
// weight fragments are laid out in LDS once per block in exactly the B
// layout, so each is one 16-byte read.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int MF_KS1 = (MLP_MAX_M + 31) / 32;   // K steps of layer 1
constexpr int MF_KS2 = MLP_H / 32;              // K steps of layer 2
constexpr int MF_NT = MLP_H / 16;               // 16-wide column tiles of a hidden layer
constexpr int MF_WAVES = BLK / 64;
constexpr int MF_TPW = 4;                       // most node tiles per wave
constexpr int MF_COLS = MF_KS1 * 8;             // residual columns a lane holds per node

__device__ inline __bf16 bits_bf(bf16 v) { return __builtin_bit_cast(__bf16, v.bits); }

// LDS written by some lanes of a wave is visible to the other lanes of that
// wave afterwards (each wave uses its own LDS rows, so no workgroup barrier).
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The network's weights as the correction reads them, B-operand layout (f1:
// residual columns of W1, f2: W2), with b2 and BF16(w3).  Built once per
// weight update in memory (pack_net_kernel); each correction block copies the
// image into LDS with 16-byte loads (load_net).
struct NetLds {
    bf16x8 f1[MF_KS1][MF_NT][64];
    bf16x8 f2[MF_KS2][MF_NT][64];
    float b2[MLP_H], w3[MLP_H];
};
static_assert(sizeof(NetLds) % 16 == 0, "the image copies in 16-byte pieces");
__global__ void __launch_bounds__(BLK)
pack_net_kernel(NetLds *__restrict__ out, int m, const bf16 *__restrict__ W1rT, const bf16 *__restrict__ W2,
                const float *__restrict__ b2, const bf16 *__restrict__ w3)
{
    NetLds &n = *out;
    const int tid = threadIdx.x, nthr = BLK;
    for (int q = tid; q < MF_KS1 * MF_NT * 64; q += nthr) {
        const int l = q % 64, t = (q / 64) % MF_NT, ks = q / (64 * MF_NT);
        bf16x8 v;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int k = ks * 32 + (l / 16) * 8 + e, n = t * 16 + l % 16;   // residual column k, hidden unit n
            v[e] = k < m ? bits_bf(W1rT[k * MLP_H + n]) : (__bf16)0.f;
        }
        n.f1[ks][t][l] = v;
    }
    for (int q = tid; q < MF_KS2 * MF_NT * 64; q += nthr) {
        const int l = q % 64, t = (q / 64) % MF_NT, ks = q / (64 * MF_NT);
        bf16x8 v;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int k = ks * 32 + (l / 16) * 8 + e, n = t * 16 + l % 16;   // layer-2 input k, output n
            v[e] = bits_bf(W2[n * MLP_H + k]);
        }
        n.f2[ks][t][l] = v;
    }
    if (tid < MLP_H) {
        n.b2[tid] = b2[tid];
        n.w3[tid] = bf(w3[tid]);
    }
}
// Threads tid = 0 .. nthr-1 of the block copy the packed image into LDS; the
// caller synchronises the block afterwards.
__device__ inline void load_net(NetLds &n, const NetLds *__restrict__ img, int tid, int nthr)
{
    const uint4 *s = reinterpret_cast<const uint4 *>(img);
    uint4 *d = reinterpret_cast<uint4 *>(&n);
    for (int q = tid; q < (int)(sizeof(NetLds) / 16); q += nthr) d[q] = s[q];
}

// The geometric part of layer 1 for the four rows n4 .. n4+3 (a multiple of
// 4, so the four sit in one 16-row tile of G) of the hidden units
// t*16 + li: one aligned 16-byte load per column tile, zero when !have.
__device__ inline void load_geo(f32x4 (&c)[MF_NT], const float *__restrict__ G, long n4, bool have, int li)
{
#pragma unroll
    for (int t = 0; t < MF_NT; t++)
        c[t] = have ? *reinterpret_cast<const f32x4 *>(G + geo_at(n4, t * 16 + li))
                    : (f32x4){0.f, 0.f, 0.f, 0.f};
}

// Both layers of one 16-node tile from the layer-1 accumulators c (geometric
// part) and this lane's scaled residual operands u; returns the network
// output before b3 of node l%16, in every lane.  hb / yb are the wave's own
// LDS rows; the caller syncs the wave before they are written again.
__device__ inline float net_tile(f32x4 (&c)[MF_NT], const bf16x8 (&u)[MF_KS1], const NetLds &n,
                                 __bf16 (*hb)[MLP_H], float *yb, int l)
{
    const int li = l % 16, g = l / 16;
#pragma unroll
    for (int ks = 0; ks < MF_KS1; ks++)
#pragma unroll
        for (int t = 0; t < MF_NT; t++)
            c[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(u[ks], n.f1[ks][t][l], c[t], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < MF_NT; t++)
#pragma unroll
        for (int v = 0; v < 4; v++) hb[g * 4 + v][t * 16 + li] = (__bf16)relu(c[t][v]);
    wave_sync();
    f32x4 c2[MF_NT];
#pragma unroll
    for (int t = 0; t < MF_NT; t++) {
        const float bb = n.b2[t * 16 + li];
        c2[t] = (f32x4){bb, bb, bb, bb};
    }
#pragma unroll
    for (int ks = 0; ks < MF_KS2; ks++) {
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&hb[li][ks * 32 + g * 8]);
#pragma unroll
        for (int t = 0; t < MF_NT; t++)
            c2[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, n.f2[ks][t][l], c2[t], 0, 0, 0);
    }
    float sv[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        float acc = 0.f;
#pragma unroll
        for (int t = 0; t < MF_NT; t++) acc += n.w3[t * 16 + li] * relu(c2[t][v]);
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) acc += __shfl_xor(acc, o);   // the 16 lanes of group g
        sv[v] = acc;
    }
    if (li == 0)
#pragma unroll
        for (int v = 0; v < 4; v++) yb[g * 4 + v] = sv[v];
    wave_sync();
    return yb[li];
}

// sigma = RMS of the m stencil residuals of node l%16 (this lane holds
// MF_COLS of them, zero beyond m or on rows without a correction), summed in
// column order per lane, then across the four lane groups.
__device__ inline double stencil_sigma(double s2, int m)
{
    s2 += __shfl_xor(s2, 16);
    s2 += __shfl_xor(s2, 32);
    return sqrt(s2 / m);
}

// Four waves per SIMD (114 registers with the 24 residual gathers a lane
// keeps in flight).  S: the slab's slice height (entry j of row r at the
// row's base + j*S, so a lane's columns sit at compile-time offsets).
template <int S, typename TV>
__global__ void __launch_bounds__(BLK, 4)
correct_kernel(int Nd, Ell A, const uint8_t *__restrict__ btag, const float *__restrict__ G,
               const NetLds *__restrict__ img, float b3, const TV *__restrict__ res, TV *__restrict__ p, int tpw)
{
    __shared__ NetLds net;
    __shared__ __attribute__((aligned(16))) __bf16 hb[MF_WAVES][16][MLP_H];
    __shared__ float yb[MF_WAVES][16];
    const int m = A.m;
    load_net(net, img, threadIdx.x, BLK);
    __syncthreads();
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, li = l % 16, g = l / 16;
    const long ntiles = ((long)Nd + 15) / 16;
    for (int it = 0; it < tpw; it++) {
        const long tile = ((long)blockIdx.x * MF_WAVES + wave) * tpw + it;
        if (tile >= ntiles) break;
        const long nb = tile * 16, r = nb + li;
        // No load waits on the row tag: rows that get no correction (boundary,
        // padding: valid columns, weight 0) gather their stencils too, and
        // since each row of the tile's products depends only on its own
        // node's inputs, their values reach no other node's correction.
        const bool inr = r < Nd;
        const uint8_t tag = inr ? btag[r] : (uint8_t)ROW_PAD;
        const size_t base = (size_t)(r / S) * S * m + (size_t)(r % S);
        const int *cp = A.col + base + (size_t)g * 8 * S;
        TV pold = 0;
        double diag = 1;
        if (g == 0 && inr) {
            pold = p[r];
            diag = A.w[base];
        }
        // this lane's residual columns of node li: j = ks*32 + g*8 + e
        double vq[MF_COLS];
        double s2 = 0;
#pragma unroll
        for (int ks = 0; ks < MF_KS1; ks++)
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int j = ks * 32 + g * 8 + e;
                double v = 0;
                if (inr && j < m) v = (double)res[cp[(ks * 32 + e) * S]];
                vq[ks * 8 + e] = v;
                s2 += v * v;
            }
        const bool valid = tag == ROW_INTERIOR;
        const double sigma = stencil_sigma(s2, m);
        const double inv = sigma > 0 ? 1.0 / sigma : 0.0;
        // Nd and nb are multiples of 4: the lane's four rows are all rows or all beyond Nd
        f32x4 c[MF_NT];
        load_geo(c, G, nb + g * 4, nb + g * 4 < Nd, li);
        bf16x8 u[MF_KS1];
#pragma unroll
        for (int ks = 0; ks < MF_KS1; ks++)
#pragma unroll
            for (int e = 0; e < 8; e++) u[ks][e] = (__bf16)(float)(vq[ks * 8 + e] * inv);
        const float y = net_tile(c, u, net, hb[wave], yb[wave], l);
        if (g == 0 && valid && sigma > 0) p[r] = pold + (TV)((double)(y + b3) * sigma / diag);
        wave_sync();   // hb and yb are rewritten by the next tile
    }
}

// The same correction over a level in Morton blocks with LDS windows, one
// workgroup per block: the block's residual window -- its own rows, then its
// halo rows (A.win_*), plus one zero -- is loaded into LDS once, and each
// stencil residual is read there through its 16-bit place instead of being
// gathered from memory (about 70 gathers per node).  A lane reads its places
// once (two per register) and the window twice (the RMS, then the scaled
// operands), so it holds no residual values across the tile.  Tiles of 16 rows start at the block's
// first row (a multiple of 4).  Same operations in the same order as
// correct_kernel, so the corrections are bitwise the same.
constexpr int CW_WAVES = 16;
template <int S, typename TV>
__global__ void __launch_bounds__(64 * CW_WAVES)
correct_window_kernel(Ell A, const uint8_t *__restrict__ btag, const float *__restrict__ G,
                      const NetLds *__restrict__ img, float b3, const TV *__restrict__ res, TV *__restrict__ p)
{
    __shared__ NetLds net;
    __shared__ __attribute__((aligned(16))) __bf16 hb[CW_WAVES][16][MLP_H];
    __shared__ float yb[CW_WAVES][16];
    extern __shared__ unsigned char lds_raw[];
    TV *win = reinterpret_cast<TV *>(lds_raw);
    const int m = A.m;
    const int b = (int)xcd_block(blockIdx.x, gridDim.x);
    const int r0 = A.win_row0[b], r1 = A.win_row0[b + 1];
    const int h0 = A.win_hoff[b], nh = A.win_hoff[b + 1] - h0, nown = r1 - r0, zero = nown + nh;
    for (int i = threadIdx.x; i < nown; i += 64 * CW_WAVES) win[i] = res[r0 + i];
    for (int i = threadIdx.x; i < nh; i += 64 * CW_WAVES) win[nown + i] = res[A.win_halo[h0 + i]];
    if (threadIdx.x == 0) win[zero] = TV(0);
    load_net(net, img, threadIdx.x, 64 * CW_WAVES);
    __syncthreads();
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, li = l % 16, g = l / 16;
    for (int nb = r0 + 16 * wave; nb < r1; nb += 16 * CW_WAVES) {
        const int r = nb + li;
        // no load waits on the row tag (as in correct_kernel)
        const bool inr = r < r1;
        const uint8_t tag = inr ? btag[r] : (uint8_t)ROW_PAD;
        // entry j of row r at base + j * S; this lane's columns j = ks*32 + g*8 + e
        // sit at compile-time offsets from cp
        const size_t base = (size_t)(r / S) * S * m + (size_t)(r % S);
        const uint16_t *cp = A.win_col + base + (size_t)g * 8 * S;
        TV pold = 0;
        double diag = 1;
        if (g == 0 && inr) {
            pold = p[r];
            diag = A.w[base];
        }
        unsigned pl[MF_COLS / 2];   // two 16-bit places per register, column q in half q % 2
        double s2 = 0;
#pragma unroll
        for (int ks = 0; ks < MF_KS1; ks++)
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int j = ks * 32 + g * 8 + e, q = ks * 8 + e;
                const unsigned v = (inr && j < m) ? (unsigned)cp[(ks * 32 + e) * S] : (unsigned)zero;
                pl[q / 2] = (q % 2) ? (pl[q / 2] | v << 16) : v;
            }
        const bool valid = tag == ROW_INTERIOR;
#pragma unroll
        for (int q = 0; q < MF_COLS; q++) {
            const double v = (double)win[(pl[q / 2] >> (16 * (q % 2))) & 0xffff];
            s2 += v * v;
        }
        const double sigma = stencil_sigma(s2, m);
        const double inv = sigma > 0 ? 1.0 / sigma : 0.0;
        f32x4 c[MF_NT];
        load_geo(c, G, nb + g * 4, nb + g * 4 < r1, li);   // r0, r1: multiples of 4
        bf16x8 u[MF_KS1];
#pragma unroll
        for (int ks = 0; ks < MF_KS1; ks++)
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int q = ks * 8 + e;
                u[ks][e] = (__bf16)(float)((double)win[(pl[q / 2] >> (16 * (q % 2))) & 0xffff] * inv);
            }
        const float y = net_tile(c, u, net, hb[wave], yb[wave], l);
        if (g == 0 && valid && sigma > 0) p[r] = pold + (TV)((double)(y + b3) * sigma / diag);
        wave_sync();   // hb and yb are rewritten by the next tile
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
    if (in > FWD_MAX_IN) MGM_FAIL(-1, "mlp_forward: too many inputs");
    const int blocks = (n + FWD_BS - 1) / FWD_BS;
    if (blocks > 0) {
        hipLaunchKernelGGL(fwd_kernel, dim3(blocks), dim3(BLK), 0, s, n, in, X, W, H1, H2, y);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}
size_t mlp_gradient_workspace(int in) { return (size_t)WG_SPLIT * mlp_weight_count(in); }
int mlp_gradient(int n, int in, const float *X, const float *t, const float *W, const float *H1, const float *H2,
                 const float *y, float *dy, float *dZ2, float *dZ1, float *part, float *G, hipStream_t s)
{
    LAUNCH(delta_kernel, n, n, W, in, H1, H2, y, t, dy, dZ2, dZ1);
    const int total = (int)mlp_weight_count(in);
    if (n > 0) {
        hipLaunchKernelGGL(wgrad_part_kernel, dim3(nblk(total), WG_SPLIT), dim3(BLK), 0, s, n, in, X, H1, H2, dy,
                           dZ2, dZ1, part);
        HIP_TRY(hipGetLastError());
    }
    LAUNCH(wgrad_sum_kernel, total, total, part, G);
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
    if (A.sorted) MGM_FAIL(-1, "learned smoother: needs the plain slab");
    LAUNCH(features_kernel, n, n, A, dim, h, dx, dy, dz, rows, snap, Nd, res_snaps, err_snaps, X, t);
    return 0;
}

int mlp_geometry(int Nd, Ell A, int dim, double h, const double *dx, const double *dy, const double *dz,
                 const uint8_t *btag, const float *W, float *G, hipStream_t s)
{
    if (A.m > MLP_MAX_M) MGM_FAIL(-1, "learned smoother: stencil wider than supported");
    if (A.sorted) MGM_FAIL(-1, "learned smoother: needs the plain slab");
    LAUNCH(geom_kernel, Nd, Nd, A, dim, h, dx, dy, dz, btag, W, G);
    return 0;
}

// LDS of the window form: the kernel's static arrays and the window.
constexpr size_t kLdsPerCu = 160 * 1024;
template <typename TV>
static size_t correct_window_static()
{
    return sizeof(NetLds) + sizeof(__bf16) * CW_WAVES * 16 * MLP_H + sizeof(float) * CW_WAVES * 16;
}
template <typename TV>
static bool correct_window_fits(const Ell &A)
{
    return A.win_blocks > 0 && (A.S == 4 || A.S == 8 || A.S == 16) &&
           correct_window_static<TV>() + ((size_t)A.win_max + 1) * sizeof(TV) <= kLdsPerCu;
}
// MGM_LEARNED_WINDOW=0 keeps the gathering form on every level.
static bool window_learned_on()
{
    const char *e = getenv("MGM_LEARNED_WINDOW");
    return !(e && e[0] == '0');
}

int mlp_device_init()
{
    const int dbl = (int)(kLdsPerCu - correct_window_static<double>());
    const int flt = (int)(kLdsPerCu - correct_window_static<float>());
    for (const void *f : {(const void *)correct_window_kernel<4, double>, (const void *)correct_window_kernel<8, double>,
                          (const void *)correct_window_kernel<16, double>})
        HIP_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, dbl));
    for (const void *f : {(const void *)correct_window_kernel<4, float>, (const void *)correct_window_kernel<8, float>,
                          (const void *)correct_window_kernel<16, float>})
        HIP_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, flt));
    return 0;
}

size_t mlp_geometry_size(int Nd) { return (size_t)((Nd + 15) / 16) * 16 * MLP_H; }

size_t mlp_net_bytes() { return sizeof(NetLds); }

int mlp_pack_net(int m, const bf16 *W1rT, const bf16 *W2, const float *b2, const bf16 *w3, void *out, hipStream_t s)
{
    if (m > MLP_MAX_M) MGM_FAIL(-1, "learned smoother: stencil wider than supported");
    hipLaunchKernelGGL(pack_net_kernel, dim3(1), dim3(BLK), 0, s, (NetLds *)out, m, W1rT, W2, b2, w3);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <typename TV>
int mlp_correct(int Nd, Ell A, const uint8_t *btag, const float *G, const void *net, float b3, const TV *res, TV *p,
                hipStream_t s)
{
    if (A.m > MLP_MAX_M) MGM_FAIL(-1, "learned smoother: stencil wider than supported");
    if (A.sorted) MGM_FAIL(-1, "learned smoother: needs the plain slab");
    if (Nd % 4) MGM_FAIL(-1, "learned smoother: device rows must be a multiple of 4");
    const NetLds *img = (const NetLds *)net;
    if (correct_window_fits<TV>(A) && window_learned_on()) {
        const size_t shm = ((size_t)A.win_max + 1) * sizeof(TV);
        const dim3 grid(A.win_blocks), block(64 * CW_WAVES);
        if (A.S == 4)
            hipLaunchKernelGGL((correct_window_kernel<4, TV>), grid, block, shm, s, A, btag, G, img, b3, res, p);
        else if (A.S == 8)
            hipLaunchKernelGGL((correct_window_kernel<8, TV>), grid, block, shm, s, A, btag, G, img, b3, res, p);
        else
            hipLaunchKernelGGL((correct_window_kernel<16, TV>), grid, block, shm, s, A, btag, G, img, b3, res, p);
        HIP_TRY(hipGetLastError());
        return 0;
    }
    // Tiles per wave: one on small levels (every tile's latency chain runs at
    // once), up to MF_TPW where the level has enough tiles to keep about
    // 12k waves busy (each block copies the network image once).
    const long tiles = ((long)Nd + 15) / 16;
    const int tpw = (int)std::max(1L, std::min((long)MF_TPW, tiles / 12288));
    const long waves = (tiles + tpw - 1) / tpw;
    const int blocks = (int)((waves + MF_WAVES - 1) / MF_WAVES);
    if (blocks > 0) {
        if (A.S == 4)
            hipLaunchKernelGGL((correct_kernel<4, TV>), dim3(blocks), dim3(BLK), 0, s, Nd, A, btag, G, img, b3, res, p,
                               tpw);
        else if (A.S == 8)
            hipLaunchKernelGGL((correct_kernel<8, TV>), dim3(blocks), dim3(BLK), 0, s, Nd, A, btag, G, img, b3, res, p,
                               tpw);
        else if (A.S == 16)
            hipLaunchKernelGGL((correct_kernel<16, TV>), dim3(blocks), dim3(BLK), 0, s, Nd, A, btag, G, img, b3, res,
                               p, tpw);
        else
            MGM_FAIL(-1, "learned smoother: unsupported slice height");
        HIP_TRY(hipGetLastError());
    }
    return 0;
}
template int mlp_correct<double>(int, Ell, const uint8_t *, const float *, const void *, float, const double *,
                                 double *, hipStream_t);
template int mlp_correct<float>(int, Ell, const uint8_t *, const float *, const void *, float, const float *, float *,
                                hipStream_t);

}  // namespace mgm
