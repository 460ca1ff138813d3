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

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

#include "device.hpp"
#include "rbf_entry.hpp"

namespace mgm {

namespace {

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

// ---- register-tiled LU: one workgroup per system ---------------------------
//
// The augmented matrix [A | b_1 .. b_R] (NC = 16*TILE rows and columns, the
// first mpn of each used, then up to NC - mpn right-hand sides) is spread
// cyclically over a 16 x 16 grid of the 256 threads: element (i, j) lives in
// thread (i % 16, j % 16), register slot (i / 16, j / 16).  Gaussian
// elimination with partial pivoting (largest |a_ik| among the rows not yet
// used as pivot, first such row on ties) permutes rows logically: each step
// broadcasts the pivot column and the pivot row through LDS (double-buffered
// by step parity) and updates every remaining row in registers, right-hand
// sides included.  Back substitution then walks the pivot rows in reverse
// pivot order.  Systems with more right-hand sides than free columns run in
// several passes.
constexpr int LU_T = 256;

// Entry (i, j) of the augmented matrix: A (PHS block, polynomial blocks,
// zero block), right-hand side j - mpn for j >= mpn, 0 in the padding.  Kept
// out of line: the kernel calls it TILE^2 times per pass.
// Exponent tables px/py/pz, dimension and PHS degree come as plain values:
// a reference to the kernel's argument block would copy it to scratch.
struct PolyTerms {
    const short *px, *py, *pz;
};

__device__ __attribute__((noinline)) double lu_entry(PolyTerms a, int dim, int phs, int i, int j, int m, int n,
                                                      int nrhs, int q0, int nops, const int *ops, const double *cx,
                                                      const double *cy, const double *cz, const double (*tgt)[3],
                                                      double sx, double sy, double sz)
{
    const int mpn = m + n;
    if (i >= mpn) return 0;
    if (j < mpn) {
        if (i < m && j < m)
            return i == j ? 0.0 : rbf_phs(dist_sum(cx[i], cy[i], cz[i], cx[j], cy[j], cz[j], dim), phs);
        if (i >= m && j >= m) return 0;
        const int pt = i < m ? i : j, k = (i < m ? j : i) - m;
        return monomial(cx[pt] - sx, cy[pt] - sy, cz[pt] - sz, a.px[k], a.py[k], a.pz[k]);
    }
    if (j >= mpn + nrhs) return 0;
    const int r = j - mpn, q = q0 + r;
    if (q < nops) return operator_rhs(ops[q], i, m, dim, phs, a.px, a.py, a.pz, cx, cy, cz, sx, sy, sz);
    return target_rhs(i, m, dim, phs, a.px, a.py, a.pz, cx, cy, cz, sx, sy, sz, tgt[r][0], tgt[r][1], tgt[r][2]);
}

// WPS: waves per SIMD the register budget aims at (1: no constraint).  Four
// systems per CU fit while the tile grid fits 128 registers (measured 20%
// faster than the unconstrained three at TILE 6; five spills inside the
// elimination loop).  The launcher picks TILE and WPS per system size by
// timing (lu_choose).
//
// One launch covers a batch: workgroup b solves system (level, node) =
// list[b], so every level and uneven system sizes (mpn + 4 <= 16 TILE) share
// it; the hardware dispatcher hands the workgroups to CUs as slots free up.
// (A persistent loop over a work counter instead costs registers: 39 spills
// at TILE 6 and 8% of the build rate, measured.)
template <int TILE, int WPS>
__global__ void __launch_bounds__(LU_T, WPS)
lu_kernel(WeightsBatch bt)
{
    constexpr int NC = 16 * TILE;
    __shared__ double cx[NC], cy[NC], cz[NC];
    __shared__ double colk[2][NC], rowp[2][NC], diag[NC];
    __shared__ int sp[2], pos[NC];
    __shared__ double *outp[NC];
    __shared__ double tgt[NC][3];
    __shared__ int s_bad;
    const int tid = threadIdx.x, ti = tid % 16, tj = tid / 16;
    {
    const int2 e = bt.list[blockIdx.x];
    const int s = __builtin_amdgcn_readfirstlane(e.y);
    const WeightsArgs &a = bt.lv[__builtin_amdgcn_readfirstlane(e.x)];
    const int m = a.m, n = a.n, mpn = m + n, dim = a.dim, phs = a.phs;
    const int *cl = a.cloud + (size_t)s * m;
    for (int i = tid; i < m; i += LU_T) {
        cx[i] = a.x[cl[i]];
        cy[i] = a.y[cl[i]];
        cz[i] = a.z[cl[i]];
    }
    // right-hand sides of this system: operators, then interpolation targets
    __shared__ int ops[4];
    const int nops = (a.Dx ? 1 : 0) + (a.Dy ? 1 : 0) + ((a.Dz && dim == 3) ? 1 : 0) + (a.lap ? 1 : 0);
    if (tid == 0) {
        int q = 0;
        if (a.Dx) ops[q++] = 0;
        if (a.Dy) ops[q++] = 1;
        if (a.Dz && dim == 3) ops[q++] = 2;
        if (a.lap) ops[q++] = 3;
    }
    const int nr = a.rt_off ? a.rt_off[s + 1] - a.rt_off[s] : 0;
    const int np = a.pt_off ? a.pt_off[s + 1] - a.pt_off[s] : 0;
    const int total = nops + nr + np;
    const int cap = NC - mpn;
    if (tid == 0) s_bad = 0;
    __syncthreads();
    const double sx = cx[0], sy = cy[0], sz = cz[0];
    for (int q0 = 0; q0 < total; q0 += cap) {
        const int nrhs = (total - q0 < cap) ? total - q0 : cap;
        // pivot positions: NC = row not yet used as pivot; the pivot-column
        // buffers hold 0 on padding rows (never written by the search)
        for (int i = tid; i < NC; i += LU_T) pos[i] = NC;
        for (int i = tid; i < 2 * NC; i += LU_T) colk[i / NC][i % NC] = 0.0;
        // output rows and target points of this pass
        for (int r = tid; r < nrhs; r += LU_T) {
            const int q = q0 + r;
            if (q < nops) {
                const int op = ops[q];
                outp[r] = (op == 0 ? a.Dx : op == 1 ? a.Dy : op == 2 ? a.Dz : a.lap) + (size_t)s * m;
            } else {
                const double *t;
                if (q < nops + nr) {
                    outp[r] = a.R_out + (size_t)a.rt_row[a.rt_off[s] + q - nops] * m;
                    t = a.rt_xyz + 3 * (size_t)(a.rt_off[s] + q - nops);
                } else {
                    outp[r] = a.P_out + (size_t)a.pt_row[a.pt_off[s] + q - nops - nr] * m;
                    t = a.pt_xyz + 3 * (size_t)(a.pt_off[s] + q - nops - nr);
                }
                tgt[r][0] = t[0]; tgt[r][1] = t[1]; tgt[r][2] = t[2];
            }
        }
        __syncthreads();
        // ---- assemble the tile (hf/rbf.h:42-77 and the right-hand sides) ----
        double A[TILE][TILE];
#pragma unroll
        for (int ai = 0; ai < TILE; ai++)
#pragma unroll
            for (int cj = 0; cj < TILE; cj++)
                A[ai][cj] = lu_entry(PolyTerms{a.px, a.py, a.pz}, dim, phs, ti + 16 * ai, tj + 16 * cj, m, n,
                                     nrhs, q0, nops, ops, cx, cy, cz, tgt, sx, sy, sz);
        const int ncols = mpn + nrhs;
        // ---- Gauss-Jordan elimination with partial pivoting ----
        // The 16 lanes holding column k (one lane group of one wave) extract it
        // and find its pivot among the rows not yet used as pivots; the pivot
        // row's owners publish it; then every other row takes -= l * (pivot
        // row), branch-free, which clears column k above and below the pivot.
        // Padding rows get l = 0, and columns past the right-hand sides hold
        // zeros in every row.  colk / rowp / sp alternate between two buffers,
        // so two barriers per step suffice.
        // The step loop runs tile column by tile column (kc unrolled), so the
        // pivot column is read from registers directly, the update is one
        // fused multiply-add per element, and tile columns left of kc are
        // skipped: their columns are eliminated and never read again (pivots
        // come from columns >= k, the solution from the right-hand sides).
#pragma unroll
        for (int kc = 0; kc < TILE; kc++) {
            for (int k16 = 0; k16 < 16; k16++) {
                const int k = kc * 16 + k16;
                if (k >= mpn) break;
                const int bf = k & 1;
                if (tj == k16) {
                    double best = -1;
                    int bi = NC;
#pragma unroll
                    for (int ai = 0; ai < TILE; ai++) {
                        const double v = A[ai][kc];
                        const int i = ti + 16 * ai;
                        if (i < mpn) {
                            colk[bf][i] = v;
                            const double av = pos[i] < k ? -1.0 : fabs(v);   // pivot rows are out
                            if (av > best) { best = av; bi = i; }
                        }
                    }
#pragma unroll
                    for (int o = 8; o > 0; o >>= 1) {   // within the 16-lane group
                        const double ob = __shfl_xor(best, o);
                        const int oi = __shfl_xor(bi, o);
                        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
                    }
                    if (ti == 0) {
                        if (!(best > 0)) s_bad = 1;
                        sp[bf] = bi;
                        pos[bi] = k;
                    }
                }
                __syncthreads();
                if (s_bad) break;
                const int p = sp[bf];
                // the pivot's tile row is the same for every thread: a scalar,
                // so the row is read straight from its registers
                const int pr = __builtin_amdgcn_readfirstlane(p / 16);
                if (ti == p % 16) {
#pragma unroll
                    for (int ai = 0; ai < TILE; ai++)
                        if (ai == pr) {
#pragma unroll
                            for (int cj = kc; cj < TILE; cj++) rowp[bf][tj + 16 * cj] = A[ai][cj];
                        }
                }
                if (tid == 0) {   // the pivot's own multiplier is 0: its row stays
                    diag[k] = colk[bf][p];
                    colk[bf][p] = 0.0;
                }
                __syncthreads();
                const double rinv = 1.0 / diag[k];
#pragma unroll
                for (int ai = 0; ai < TILE; ai++) {
                    const double l = colk[bf][ti + 16 * ai] * rinv;
#pragma unroll
                    for (int cj = kc; cj < TILE; cj++)
                        A[ai][cj] = __builtin_fma(-l, rowp[bf][tj + 16 * cj], A[ai][cj]);
                }
            }
            // s_bad is written only before a step's first barrier and read after
            // it, so every thread leaves both loops at the same step
            if (s_bad) break;
        }
        if (s_bad) {
            if (tid == 0) atomicOr(bt.status, 1);
            return;
        }
        // ---- solution: unknown k is the right-hand-side entries of its pivot
        // row over the pivot; only the first m unknowns (the stencil weights)
        // are written.  Every row is done by its own lanes, in parallel.
#pragma unroll
        for (int ai = 0; ai < TILE; ai++) {
            const int i = ti + 16 * ai;
            const int k = i < mpn ? pos[i] : NC;
            if (k < m) {
                const double d = diag[k];
#pragma unroll
                for (int cj = 0; cj < TILE; cj++) {
                    const int j = tj + 16 * cj;
                    if (j >= mpn && j < ncols) outp[j - mpn][k] = A[ai][cj] / d;
                }
            }
        }
        __syncthreads();   // LDS tables are rewritten by the next pass
    }
    }
}

}  // namespace

int weights_tile(int mpn)
{
    for (int t = 2; t <= 9; t++)
        if (16 * t >= mpn + 4) return t;
    return 0;
}

using LuFn = void (*)(WeightsBatch);

template <int W>
static LuFn lu_fn_w(int tile)
{
    switch (tile) {
    case 2: return lu_kernel<2, W>;
    case 3: return lu_kernel<3, W>;
    case 4: return lu_kernel<4, W>;
    case 5: return lu_kernel<5, W>;
    case 6: return lu_kernel<6, W>;
    case 7: return lu_kernel<7, W>;
    default: return lu_kernel<8, W>;
    }
}
// 9 x 16 tiles (systems of up to 140 unknowns, the wide boundary stencils):
// unconstrained register budget only
static LuFn lu_fn(int tile, int wps)
{
    if (tile == 9) return lu_kernel<9, 1>;
    return wps == 4 ? lu_fn_w<4>(tile) : lu_fn_w<1>(tile);
}

// One workgroup per system of the batch.
static int lu_run(LuFn f, WeightsBatch b, hipStream_t s)
{
    if (b.nsys <= 0) return 0;
    hipLaunchKernelGGL(f, dim3(b.nsys), dim3(LU_T), 0, s, b);
    HIP_TRY(hipGetLastError());
    return 0;
}

// Autotuner of the weight kernel: per system size mpn, the register-tile
// Gauss-Jordan kernel (FP64; the tile grid that holds [A | 4 RHS] and the
// next, register budget of four waves per SIMD or unconstrained) and the
// matrix-core kernel (lu32.hip: FP32 LU on MFMA, FP64 refinement) are timed
// on the first 8192 systems of the first build -- every variant writes the
// same rows to within the 1e-11 bar, and the sample is simply recomputed by
// the winner -- and the fastest is kept for the process (the matrix-core
// kernel only if it is 10% faster than the best Gauss-Jordan variant).  MGM_LU_TUNE=0 takes
// the smallest Gauss-Jordan grid and the four-wave budget; MGM_LU_KERNEL=gj or
// mfma restricts the choice to one kernel; MGM_LU_VARIANT=tile,wps forces a
// Gauss-Jordan variant.  A choice with wps == 0 is the matrix-core kernel
// (tile = its 16 x 16 tile count).
static std::map<int, std::pair<int, int>> g_lu_choice;
static std::map<int, std::pair<int, int>> g_lu_last;   // variant of the last build per size

static int run_choice(std::pair<int, int> c, const WeightsBatch &b, int rmax, hipStream_t s)
{
    if (c.second == 0) return launch_weights_lu32(b, c.first, rmax, s);
    return lu_run(lu_fn(c.first, c.second), b, s);
}

static int lu_choose(int mpn, const WeightsBatch &b, int rmax, hipStream_t s, int *tile, int *wps)
{
    const int t0 = weights_tile(mpn), t32 = lu32_tiles(mpn);
    const char *kind = getenv("MGM_LU_KERNEL");
    const bool only_gj = kind && strcmp(kind, "gj") == 0, only_mfma = kind && strcmp(kind, "mfma") == 0;
    if (kind && !only_gj && !only_mfma) MGM_FAIL(-1, "MGM_LU_KERNEL: expected gj or mfma");
    if (only_mfma) {
        if (!t32) MGM_FAIL(-1, "MGM_LU_KERNEL=mfma: systems wider than 128 unknowns");
        *tile = t32;
        *wps = 0;
        return 0;
    }
    // MGM_LU_VARIANT=tile,wps forces a variant (tests compare them), uncached
    if (const char *v = getenv("MGM_LU_VARIANT")) {
        int t = 0, w = 0;
        if (sscanf(v, "%d,%d", &t, &w) == 2 && t0 && t >= t0 && t <= 9 && (w == 1 || (w == 4 && t <= 6))) {
            *tile = t;
            *wps = w;
            return 0;
        }
        MGM_FAIL(-1, "MGM_LU_VARIANT: expected tile,wps with tile >= the smallest grid and wps 1 or 4");
    }
    const int key = only_gj ? -mpn : mpn;
    auto it = g_lu_choice.find(key);
    if (it != g_lu_choice.end()) { *tile = it->second.first; *wps = it->second.second; return 0; }
    if (!t0 && only_gj) MGM_FAIL(-1, "MGM_LU_KERNEL=gj: systems too wide for the register-tile kernel");
    // systems too wide for the Gauss-Jordan register tiles go to the matrix cores
    std::pair<int, int> best = t0 ? std::pair<int, int>{t0, t0 <= 6 ? 4 : 1} : std::pair<int, int>{t32, 0};
    const char *e = getenv("MGM_LU_TUNE");
    WeightsBatch sample = b;
    sample.nsys = b.nsys < 8192 ? b.nsys : 8192;
    if (t0 && !(e && e[0] == '0') && sample.nsys >= 1024) {
        std::vector<std::pair<int, int>> cand;
        for (int t = t0; t <= t0 + 1 && t <= 9; t++) {
            if (t <= 6) cand.push_back({t, 4});
            cand.push_back({t, 1});
        }
        if (t32 && !only_gj) cand.push_back({t32, 0});
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        float best_ms = 1e30f;
        for (auto c : cand) {
            int rc = run_choice(c, sample, rmax, s);   // first launch loads the code object
            if (!rc) HIP_TRY(hipEventRecord(e0, s));
            if (!rc) rc = run_choice(c, sample, rmax, s);
            if (!rc) HIP_TRY(hipEventRecord(e1, s));
            if (rc) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return rc; }
            HIP_TRY(hipEventSynchronize(e1));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            // the matrix-core kernel (tried last) must win by 10%: its rows
            // differ from the Gauss-Jordan kernel's in the last bits, so a
            // near tie would make the choice, and the rows, vary from run to run
            if (c.second == 0 ? ms < 0.9f * best_ms : ms < best_ms) { best_ms = ms; best = c; }
        }
        HIP_TRY(hipEventDestroy(e0));
        HIP_TRY(hipEventDestroy(e1));
    }
    g_lu_choice[key] = best;
    *tile = best.first;
    *wps = best.second;
    return 0;
}

int weights_kernel_choice(int mpn, int *tile, int *wps)
{
    auto it = g_lu_last.find(mpn);
    if (it == g_lu_last.end()) return -1;
    *tile = it->second.first;
    *wps = it->second.second;
    return 0;
}

namespace {
__global__ void __launch_bounds__(256) fill_list_kernel(int2 *list, int n, const int *sys, int lvl)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) list[i] = make_int2(lvl, sys[i]);
}
}  // namespace

int launch_weights_levels(const std::vector<WeightsArgs> &levels, hipStream_t s)
{
    int mpn = 0, rmax = 0;
    long total = 0;
    for (const WeightsArgs &a : levels) {
        mpn = a.m + a.n > mpn ? a.m + a.n : mpn;
        rmax = a.rhs_max > rmax ? a.rhs_max : rmax;
        total += a.nsys;
    }
    if (levels.empty() || total == 0) return 0;
    if (!weights_tile(mpn) && !lu32_tiles(mpn)) MGM_FAIL(-1, "weights: batch with systems over 128 unknowns");
    if (total >= (1L << 31)) MGM_FAIL(-1, "weights: too many systems for one batch");
    // descriptors, the (level, node) list and the work counter; kept between
    // builds and grown when a batch needs more
    static WeightsArgs *d_lv = nullptr;
    static size_t n_lv = 0;
    static int2 *d_list = nullptr;
    static size_t n_list = 0;
    static int *d_next = nullptr;
    if (n_lv < levels.size()) {
        if (d_lv) HIP_TRY(hipFree(d_lv));   // synchronises: no launch still reads it
        d_lv = nullptr;
        HIP_TRY(hipMalloc((void **)&d_lv, sizeof(WeightsArgs) * levels.size()));
        n_lv = levels.size();
    }
    if (n_list < (size_t)total) {
        if (d_list) HIP_TRY(hipFree(d_list));
        d_list = nullptr;
        HIP_TRY(hipMalloc((void **)&d_list, sizeof(int2) * total));
        n_list = total;
    }
    if (!d_next) HIP_TRY(hipMalloc((void **)&d_next, sizeof(int)));
    HIP_TRY(hipMemcpyAsync(d_lv, levels.data(), sizeof(WeightsArgs) * levels.size(), hipMemcpyHostToDevice, s));
    long off = 0;
    for (size_t l = 0; l < levels.size(); l++) {
        const int n = levels[l].nsys;
        if (n > 0) hipLaunchKernelGGL(fill_list_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_list + off, n,
                                      levels[l].sys, (int)l);
        HIP_TRY(hipGetLastError());
        off += n;
    }
    WeightsBatch b{d_lv, d_list, (int)total, d_next, levels[0].status};
    int tile = 0, wps = 0, rc;
    if ((rc = lu_choose(mpn, b, rmax, s, &tile, &wps))) return rc;
    for (const WeightsArgs &a : levels) g_lu_last[a.m + a.n] = {tile, wps};
    g_lu_last[mpn] = {tile, wps};
    return run_choice({tile, wps}, b, rmax, s);
}

int launch_weights(const WeightsArgs &a, hipStream_t s)
{
    if (weights_tile(a.m + a.n) || lu32_tiles(a.m + a.n)) return launch_weights_levels({a}, s);
    if (!a.ws) MGM_FAIL(-1, "weights: workspace missing for wide systems");
    for (int base = 0; base < a.nsys; base += a.batch) {
        int count = (a.nsys - base < a.batch) ? a.nsys - base : a.batch;
        hipLaunchKernelGGL(weights_kernel, dim3((count + 63) / 64), dim3(64), 0, s, a, base, count);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

}  // namespace mgm
