// lu32.hip -- batched local weight solves on the matrix cores: an FP32 LU
// factorisation with partial pivoting whose trailing updates are rank-4
// products on v_mfma_f32_16x16x4_f32, then FP64 iterative refinement of
// every right-hand side against the FP64 system (residual b - A x in FP64).
//
// It solves the same systems as weights.hip (A = [[Phi, P], [P^T, 0]] over a
// node's stencil, hf/rbf.h:42-77; right-hand sides = the four operators at
// the seed, hf/rbf.h:79-270, and the interpolation targets seeded at the
// node, hf/mesh_functions.h:533-670) and writes the same rows.  The
// reference inverts each system by Gauss-Jordan in FP64 (hf/mat_lib.h:423-491);
// here the O(mpn^3) work runs in FP32 on matrix cores and FP64 is spent only
// on O(mpn^2) residuals.  The refined solution is that of the FP64 system.
//
// One wave per system, no workgroup barriers.  The augmented matrix (padded
// to 16T x 16T) lives in registers as T x T tiles of 16 x 16 in the MFMA
// accumulator layout of the tile's transpose: lane l = 16 g + li holds, in
// register q of tile (ti, tj), A[16 ti + li][16 tj + 4 g + q].  A column is
// then 16 lanes (one group g) x T registers; a row is the 4 lanes li of its
// row index times 4T registers.
//
// Factorisation: columns in panels of four (one register group).  Within a
// panel the 16 lanes holding the pivot column pick the pivot (largest
// magnitude among rows not yet used as pivots), compute the multipliers and
// update the panel's remaining columns.  Rows are permuted logically (a row
// keeps its place; pos[] records the step that used it as pivot).  The
// panel's four pivot rows then take their triangular update in LDS, and one
// MFMA per 16 x 16 tile applies the rank-4 update
//     A[r][J] -= sum_q L[r][4c+q] U[q][J]
// to every trailing tile.  Solves walk the pivot order (forward: rows not yet
// pivot, backward: earlier pivots), one right-hand side per lane group.
//
// Refinement: x = 0, r = b; repeat { solve LU d = fl32(r); x += d; r = b - A x }
// until every right-hand side's correction is below 1e-12 of its largest
// weight (typically three refinements for PHS r^3 stencils, five for r^5).
// Phi and P are kept in FP64 in a per-wave global scratch slot (written once
// per system), so residuals stream them instead of recomputing square roots.
//
// Work distribution: a persistent grid of one-wave workgroups takes systems
// from a global counter; each workgroup owns one scratch slot.
#pragma clang fp contract(off)

#include <cstdio>
#include <cstdlib>

#include "device.hpp"
#include "rbf_entry.hpp"

namespace mgm {

namespace {

constexpr int kNotPivot = 1 << 20;
typedef float f4 __attribute__((ext_vector_type(4)));

// Per-wave LDS of a system of up to 16T unknowns.
template <int T>
struct Lu32Smem {
    // xs, bs and rs are contiguous: the matrix assembly stages 16 rows of
    // 16T + 4 floats there
    double cx[16 * T], cy[16 * T], cz[16 * T];
    double xs[4][16 * T];   // FP64 solution of each right-hand side, by unknown
    double bs[4][16 * T];   // FP64 right-hand sides, by row
    float rs[4][16 * T];    // FP32 residual, by row
    float U[4][16 * T];     // the panel's pivot rows
    f4 L[T][16];            // the panel's multipliers of each row (0 on pivot rows)
    float diag[16 * T];     // pivot value of each step
    int piv[16 * T];        // pivot row of each step
    double *outp[4];
};

__device__ inline float bcast(float v, int lane) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane)); }
__device__ inline int bcast(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Branch-free selections written with bit masks: the compiler turns select
// chains over a register array (on a uniform or a per-lane index alike) into
// an indexed load from a stack copy, which would move the whole array to
// scratch memory.
__device__ inline unsigned lane_mask(bool c) { return 0u - (unsigned)c; }

// v[ti] on the lanes that hold row p (16 ti + li == p), 0 elsewhere.
template <int T>
__device__ __attribute__((always_inline)) inline float row_reg(const float (&v)[T], int p, int li)
{
    unsigned r = 0;
#pragma unroll
    for (int i = 0; i < T; i++) r |= __float_as_uint(v[i]) & lane_mask(16 * i + li == p);
    return __uint_as_float(r);
}

// v[g] for this lane's group g (masks mg[i] = lane_mask(g == i)).
__device__ __attribute__((always_inline)) inline float group_sel(float v0, float v1, float v2, float v3,
                                                                  const unsigned (&mg)[4])
{
    return __uint_as_float((__float_as_uint(v0) & mg[0]) | (__float_as_uint(v1) & mg[1]) |
                           (__float_as_uint(v2) & mg[2]) | (__float_as_uint(v3) & mg[3]));
}

// Maximum over the 16 lanes of each row of the wave (DPP moves: quad swaps,
// then rotations by 4 and 8 within the row).
__device__ __attribute__((always_inline)) inline unsigned row16_max(unsigned v)
{
    unsigned o = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);   // quad_perm 1,0,3,2
    v = o > v ? o : v;
    o = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);            // quad_perm 2,3,0,1
    v = o > v ? o : v;
    o = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false);           // row_ror 4
    v = o > v ? o : v;
    o = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false);           // row_ror 8
    return o > v ? o : v;
}

// The value of row p of group g's right-hand side, for every lane of group g.
template <int T>
__device__ __attribute__((always_inline)) inline float row_value(const float (&rr)[T], int p, int li,
                                                                 const unsigned (&mg)[4])
{
    const float v = row_reg<T>(rr, p, li);
    const int lp = p & 15;
    return group_sel(bcast(v, lp), bcast(v, lp + 16), bcast(v, lp + 32), bcast(v, lp + 48), mg);
}

// The tile columns are rotated so that the one being worked on always sits in
// position 0: panel, solve and update code then exist once, with register
// indexes fixed at compile time.  (Selecting a tile column by a runtime index
// instead makes the compiler move the matrix to scratch memory; one copy of
// the code per tile column overflows the instruction cache.)  A full turn of
// T rotations restores the order.
template <int T>
__device__ __attribute__((always_inline)) inline void rot_left(f4 (&a)[T][T])
{
#pragma unroll
    for (int ti = 0; ti < T; ti++) {
        const f4 t0 = a[ti][0];
#pragma unroll
        for (int tj = 0; tj + 1 < T; tj++) a[ti][tj] = a[ti][tj + 1];
        a[ti][T - 1] = t0;
    }
}

template <int T>
__device__ __attribute__((always_inline)) inline void rot_right(f4 (&a)[T][T])
{
#pragma unroll
    for (int ti = 0; ti < T; ti++) {
        const f4 t1 = a[ti][T - 1];
#pragma unroll
        for (int tj = T - 1; tj > 0; tj--) a[ti][tj] = a[ti][tj - 1];
        a[ti][0] = t1;
    }
}

// One panel: columns 4c .. 4c+3 (c = 4 tc + gc), register group gc of the
// tile column in position 0; positions 1 .. T-1-tc hold the trailing tile
// columns, the rest the finished ones.  Returns false on a zero pivot column.
template <int T>
__device__ __attribute__((always_inline)) inline bool lu_panel(f4 (&a)[T][T], int (&pos)[T], int tc, int gc, int mpn,
                                                               Lu32Smem<T> &sm)
{
    const int lane = threadIdx.x & 63, li = lane & 15, g = lane >> 4;
    const unsigned mg[4] = {lane_mask(g == 0), lane_mask(g == 1), lane_mask(g == 2), lane_mask(g == 3)};
    const int c = 4 * tc + gc, nt = T - tc;   // nt: tile columns still active
    const int nq = mpn - 4 * c < 4 ? mpn - 4 * c : 4;
    int pq[4] = {-1, -1, -1, -1};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (q < nq) {
            const int k = 4 * c + q;
            // pivot: largest |a| among rows not yet pivot; key = |a| with its
            // low 7 bits replaced by 127 - row (near-ties go to the lower row)
            unsigned key = 0;
#pragma unroll
            for (int ti = 0; ti < T; ti++) {
                const unsigned kv = (__float_as_uint(fabsf(a[ti][0][q])) & ~127u) | (unsigned)(127 - (16 * ti + li));
                if (pos[ti] == kNotPivot && kv > key) key = kv;
            }
            key = row16_max(key);
            key = (unsigned)bcast((int)key, 16 * gc);
            if ((key & ~127u) == 0) return false;
            const int p = uni(127 - (int)(key & 127u)), src = (p & 15) + 16 * gc;
            pq[q] = p;
            float v0[T], v1[T], v2[T], v3[T];
#pragma unroll
            for (int ti = 0; ti < T; ti++) {
                v0[ti] = a[ti][0][0]; v1[ti] = a[ti][0][1]; v2[ti] = a[ti][0][2]; v3[ti] = a[ti][0][3];
            }
            // the pivot row's panel entries (the pivot itself is entry q)
            const float u0 = bcast(row_reg<T>(v0, p, li), src), u1 = bcast(row_reg<T>(v1, p, li), src);
            const float u2 = bcast(row_reg<T>(v2, p, li), src), u3 = bcast(row_reg<T>(v3, p, li), src);
            const float d = q == 0 ? u0 : q == 1 ? u1 : q == 2 ? u2 : u3;
#pragma unroll
            for (int ti = 0; ti < T; ti++)
                if (16 * ti + li == p) pos[ti] = k;
            if (lane == 0) {
                sm.piv[k] = p;
                sm.diag[k] = d;
            }
            if (g == gc) {
                const float rd = __builtin_amdgcn_rcpf(d);   // FP32 factors are refined anyway
#pragma unroll
                for (int ti = 0; ti < T; ti++) {
                    const bool upd = pos[ti] > k;
                    const float mlt = a[ti][0][q] * rd;
                    a[ti][0][q] = upd ? mlt : a[ti][0][q];
                    if (q < 1) a[ti][0][1] = upd ? __builtin_fmaf(-mlt, u1, a[ti][0][1]) : a[ti][0][1];
                    if (q < 2) a[ti][0][2] = upd ? __builtin_fmaf(-mlt, u2, a[ti][0][2]) : a[ti][0][2];
                    if (q < 3) a[ti][0][3] = upd ? __builtin_fmaf(-mlt, u3, a[ti][0][3]) : a[ti][0][3];
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    // the panel's pivot rows, in storage order, to LDS (4 lanes per row)
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int ti = 0; ti < T; ti++)
            if (16 * ti + li == pq[q]) {
#pragma unroll
                for (int tj = 0; tj < T; tj++)
                    *reinterpret_cast<f4 *>(&sm.U[q][16 * tj + 4 * g]) = a[ti][tj];
            }
    // the panel's multipliers, 0 on rows that are pivots by now
    if (g == gc) {
#pragma unroll
        for (int ti = 0; ti < T; ti++) {
            const f4 z = {0.f, 0.f, 0.f, 0.f};
            sm.L[ti][li] = pos[ti] > 4 * c + 3 ? a[ti][0] : z;
        }
    }
    __syncthreads();
    // triangular update of the pivot rows' trailing columns (storage columns
    // 4gc+4 .. 16nt-1): U[q][j] -= sum_{q' < q} L11[q][q'] U[q'][j],
    // L11[q][q'] = U[q][4gc + q']
    const int jb = 4 * gc;
    for (int j = jb + 4 + lane; j < 16 * nt; j += 64) {
        const float u0 = sm.U[0][j];
        if (nq > 1) {
            const float u1 = __builtin_fmaf(-sm.U[1][jb], u0, sm.U[1][j]);
            sm.U[1][j] = u1;
            if (nq > 2) {
                const float u2 = __builtin_fmaf(-sm.U[2][jb + 1], u1, __builtin_fmaf(-sm.U[2][jb], u0, sm.U[2][j]));
                sm.U[2][j] = u2;
                if (nq > 3)
                    sm.U[3][j] = __builtin_fmaf(-sm.U[3][jb + 2], u2,
                                                __builtin_fmaf(-sm.U[3][jb + 1], u1,
                                                               __builtin_fmaf(-sm.U[3][jb], u0, sm.U[3][j])));
            }
        }
    }
    __syncthreads();
    // pivot rows take their updated trailing columns back (the 4 lanes of
    // each pivot row, a branch per tile row)
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int ti = 0; ti < T; ti++)
            if (16 * ti + li == pq[q]) {
#pragma unroll
                for (int tj = 0; tj < T; tj++)
                    if ((tj > 0 && tj < nt) || (tj == 0 && g > gc))
                        a[ti][tj] = *reinterpret_cast<const f4 *>(&sm.U[q][16 * tj + 4 * g]);
            }
    // rank-4 update of the trailing tiles on the matrix cores:
    // C'[x][y] = A[16ti + y][16tj + x] -= sum_k U[k][16tj + x] L[16ti + y][4c + k]
    float bop[T];
#pragma unroll
    for (int ti = 0; ti < T; ti++) {
        const f4 l4 = sm.L[ti][li];
        bop[ti] = group_sel(l4[0], l4[1], l4[2], l4[3], mg);
    }
#pragma unroll
    for (int tj = 0; tj < T; tj++) {
        if (tj >= nt || (tj == 0 && gc == 3)) continue;
        const int j = 16 * tj + li;
        const float aop = (j >= jb + 4 && g < nq) ? -sm.U[g][j] : 0.f;
#pragma unroll
        for (int ti = 0; ti < T; ti++) a[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(aop, bop[ti], a[ti][tj], 0, 0, 0);
    }
    __syncthreads();   // sm.U / sm.L are rewritten by the next panel
    return true;
}

template <int T>
__device__ __attribute__((always_inline)) inline bool lu_factor(f4 (&a)[T][T], int (&pos)[T], int mpn,
                                                                Lu32Smem<T> &sm)
{
#pragma unroll 1
    for (int tc = 0; tc < T; tc++) {
#pragma unroll 1
        for (int gc = 0; gc < 4 && 16 * tc + 4 * gc < mpn; gc++)
            if (!lu_panel<T>(a, pos, tc, gc, mpn, sm)) return false;
        rot_left<T>(a);
    }
    return true;
}

// Columns 16 kt + 4 gk + q (q = 0..3) of the factors, the tile column in
// position 0, for the rows of this lane (16 ti + li): from the lanes of group gk.
template <int T>
__device__ __attribute__((always_inline)) inline void factor_cols(const f4 (&a)[T][T], int gk, float (&col)[4][T])
{
    const int src = (threadIdx.x & 15) + 16 * gk;
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int ti = 0; ti < T; ti++) col[q][ti] = __shfl(a[ti][0][q], src);
}

// Forward elimination of group g's right-hand side (rr: rows 16 ti + li):
// step k, rows not yet pivot take -= L[r][k] y_k, y_k the value of row piv[k].
template <int T>
__device__ __attribute__((always_inline)) inline void solve_fwd(f4 (&a)[T][T], const int (&pos)[T], float (&rr)[T],
                                                                int mpn, const Lu32Smem<T> &sm)
{
    const int lane = threadIdx.x & 63, li = lane & 15, g = lane >> 4;
    const unsigned mg[4] = {lane_mask(g == 0), lane_mask(g == 1), lane_mask(g == 2), lane_mask(g == 3)};
#pragma unroll 1
    for (int kt = 0; kt < T; kt++) {
#pragma unroll 1
        for (int gk = 0; gk < 4; gk++) {
            const int k0 = 16 * kt + 4 * gk;
            if (k0 >= mpn) break;
            float col[4][T];
            factor_cols<T>(a, gk, col);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int k = k0 + q;
                if (k < mpn) {
                    const int p = uni(sm.piv[k]);
                    const float y = row_value<T>(rr, p, li, mg);
#pragma unroll
                    for (int ti = 0; ti < T; ti++) {
                        const float nv = __builtin_fmaf(-col[q][ti], y, rr[ti]);
                        rr[ti] = pos[ti] > k ? nv : rr[ti];
                    }
                }
            }
        }
        rot_left<T>(a);
    }
}

// Back substitution: last step first, x_k = (value of row piv[k]) / diag[k]
// is written to that row, and earlier pivot rows take -= U[r][k] x_k.
template <int T>
__device__ __attribute__((always_inline)) inline void solve_bwd(f4 (&a)[T][T], const int (&pos)[T], float (&rr)[T],
                                                                int mpn, const Lu32Smem<T> &sm)
{
    const int lane = threadIdx.x & 63, li = lane & 15, g = lane >> 4;
    const unsigned mg[4] = {lane_mask(g == 0), lane_mask(g == 1), lane_mask(g == 2), lane_mask(g == 3)};
#pragma unroll 1
    for (int kt = T - 1; kt >= 0; kt--) {
        rot_right<T>(a);
#pragma unroll 1
        for (int gk = 3; gk >= 0; gk--) {
            const int k0 = 16 * kt + 4 * gk;
            if (k0 >= mpn) continue;
            float col[4][T];
            factor_cols<T>(a, gk, col);
#pragma unroll
            for (int q = 3; q >= 0; q--) {
                const int k = k0 + q;
                if (k < mpn) {
                    const int p = uni(sm.piv[k]);
                    const float x = row_value<T>(rr, p, li, mg) / sm.diag[k];
#pragma unroll
                    for (int ti = 0; ti < T; ti++) {
                        const float nv = __builtin_fmaf(-col[q][ti], x, rr[ti]);
                        rr[ti] = 16 * ti + li == p ? x : (pos[ti] < k ? nv : rr[ti]);
                    }
                }
            }
        }
    }
}

// Scratch slot of a wave, in doubles: the FP64 system A padded to 16T x 16T
// with zeros, column-major (A[i][j] at j * 16T + i), then the right-hand
// sides of the system [r][i] (rmax x 16T).
__host__ __device__ inline long lu32_slot(int tiles, int rmax)
{
    return (long)(16 * tiles) * (16 * tiles + rmax);
}

// The q-th operator present in mask (bit 0 Dx, 1 Dy, 2 Dz, 3 Laplacian).
__device__ inline int nth_op(unsigned mask, int q)
{
    for (int op = 0; op < 4; op++)
        if (mask & (1u << op)) {
            if (q == 0) return op;
            q--;
        }
    return 3;
}

// Output row and, for interpolation targets, the target point of right-hand
// side q of system s (operators first, then restriction, then prolongation
// targets).
__device__ inline double *rhs_out(const WeightsArgs &a, int s, int q, int nops, int nr, unsigned opmask,
                                  const double **tgt)
{
    *tgt = nullptr;
    if (q < nops) {
        const int op = nth_op(opmask, q);
        return (op == 0 ? a.Dx : op == 1 ? a.Dy : op == 2 ? a.Dz : a.lap) + (size_t)s * a.m;
    }
    if (q < nops + nr) {
        *tgt = a.rt_xyz + 3 * (size_t)(a.rt_off[s] + q - nops);
        return a.R_out + (size_t)a.rt_row[a.rt_off[s] + q - nops] * a.m;
    }
    *tgt = a.pt_xyz + 3 * (size_t)(a.pt_off[s] + q - nops - nr);
    return a.P_out + (size_t)a.pt_row[a.pt_off[s] + q - nops - nr] * a.m;
}

// The FP64 system (padded, column-major) and the system's right-hand sides
// to the scratch slot.  Out of line: its FP64 constants and loop invariants
// then stay out of the register budget of the factorisation.
template <int T>
__device__ __attribute__((noinline)) void fill_system(const WeightsArgs &a, int s, int nops, int nr, int total,
                                                      unsigned opmask, double *full, double *rb, const Lu32Smem<T> &sm)
{
    constexpr int NC = 16 * T;
    const int lane = threadIdx.x & 63;
    const int m = a.m, n = a.n, mpn = m + n, dim = a.dim, phs = a.phs;
    const double sx = sm.cx[0], sy = sm.cy[0], sz = sm.cz[0];
    // the FP64 system and its right-hand sides to the scratch slot
    for (int e = lane; e < NC * NC; e += 64) {
        const int j = e / NC, i = e - j * NC;
        double v = 0.0;
        if (i < m && j < m) {
            if (i != j)
                v = rbf_phs(dist_sum(sm.cx[i], sm.cy[i], sm.cz[i], sm.cx[j], sm.cy[j], sm.cz[j], dim), phs);
        } else if (i < mpn && j < mpn && (i < m || j < m)) {
            const int pt = i < m ? i : j, k = (i < m ? j : i) - m;
            v = monomial(sm.cx[pt] - sx, sm.cy[pt] - sy, sm.cz[pt] - sz, a.px[k], a.py[k], a.pz[k]);
        }
        full[e] = v;
    }
    for (int e = lane; e < total * NC; e += 64) {
        const int q = e / NC, i = e - q * NC;
        double v = 0.0;
        if (i < mpn) {
            const double *tg;
            const double *o = rhs_out(a, s, q, nops, nr, opmask, &tg);
            (void)o;
            v = q < nops ? operator_rhs(nth_op(opmask, q), i, m, dim, phs, a.px, a.py, a.pz, sm.cx, sm.cy, sm.cz,
                                        sx, sy, sz)
                         : target_rhs(i, m, dim, phs, a.px, a.py, a.pz, sm.cx, sm.cy, sm.cz, sx, sy, sz, tg[0],
                                      tg[1], tg[2]);
        }
        rb[e] = v;
    }
}

// WPS: waves per SIMD the register budget aims at.
template <int T, int WPS>
__global__ void __launch_bounds__(64, WPS)
lu32_kernel(WeightsBatch bt, double *scratch, int rmax, int nref)
{
    __shared__ Lu32Smem<T> sm;
    constexpr int NC = 16 * T;
    const int lane = threadIdx.x & 63, li = lane & 15, g = lane >> 4;
    double *full = scratch + (size_t)blockIdx.x * lu32_slot(T, rmax);   // [j][i], NC x NC
    double *rb = full + (size_t)NC * NC;                                // [r][i], rmax x NC
    for (;;) {
        int t = 0;
        if (lane == 0) t = atomicAdd(bt.next, 1);
        t = bcast(t, 0);
        if (t >= bt.nsys) break;
        const int2 e = bt.list[t];
        const int s = uni(e.y);
        const WeightsArgs &a = bt.lv[uni(e.x)];
        const int m = a.m, n = a.n, mpn = m + n, dim = a.dim;
        const unsigned opmask = (a.Dx ? 1u : 0u) | (a.Dy ? 2u : 0u) | ((a.Dz && dim == 3) ? 4u : 0u) | (a.lap ? 8u : 0u);
        const int nops = __builtin_popcount(opmask);
        const int *cl = a.cloud + (size_t)s * m;
        for (int i = lane; i < NC; i += 64) {
            const bool in = i < m;
            const int c = in ? cl[i] : 0;
            sm.cx[i] = in ? a.x[c] : 0.0;
            sm.cy[i] = in ? a.y[c] : 0.0;
            sm.cz[i] = in ? a.z[c] : 0.0;
        }
        __syncthreads();
        const double sx = sm.cx[0], sy = sm.cy[0], sz = sm.cz[0];
        const int nr = a.rt_off ? a.rt_off[s + 1] - a.rt_off[s] : 0;
        const int np = a.pt_off ? a.pt_off[s + 1] - a.pt_off[s] : 0;
        const int total = nops + nr + np;
        fill_system<T>(a, s, nops, nr, total, opmask, full, rb, sm);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // no stale cache lines of the slot
        // FP32 matrix in registers, one tile row at a time through LDS (xs, bs
        // and rs are free until the solves): rows 16 ti .. 16 ti + 15 staged
        // row-major with a padded stride, then each lane reads its 4T entries
        f4 A[T][T];
        int pos[T];
        float *stage = reinterpret_cast<float *>(&sm.xs[0][0]);
        constexpr int SST = NC + 4;
#pragma unroll
        for (int ti = 0; ti < T; ti++) {
            pos[ti] = 16 * ti + li < mpn ? kNotPivot : -1;
            for (int e = lane; e < 16 * NC; e += 64) {
                const int j = e >> 4, r = e & 15;
                stage[r * SST + j] = (float)full[(size_t)j * NC + 16 * ti + r];
            }
            __syncthreads();
#pragma unroll
            for (int tj = 0; tj < T; tj++) A[ti][tj] = *reinterpret_cast<const f4 *>(&stage[li * SST + 16 * tj + 4 * g]);
            __syncthreads();
        }
        if (!lu_factor<T>(A, pos, mpn, sm)) {
            if (lane == 0) atomicOr(bt.status, 1);
            __syncthreads();
            continue;
        }
        // right-hand sides in passes of four (one per lane group)
        for (int q0 = 0; q0 < total; q0 += 4) {
            const int nrhs = total - q0 < 4 ? total - q0 : 4;
            if (lane < nrhs) {
                const double *tg;
                sm.outp[lane] = rhs_out(a, s, q0 + lane, nops, nr, opmask, &tg);
            }
            for (int e = lane; e < 4 * NC; e += 64) {
                const int c = e / NC, i = e - c * NC;
                const double v = c < nrhs ? rb[(size_t)(q0 + c) * NC + i] : 0.0;
                sm.bs[c][i] = v;
                sm.xs[c][i] = 0.0;
                sm.rs[c][i] = (float)v;
            }
            __syncthreads();
            for (int it = 0;; it++) {
                // d = (LU)^-1 fl32(r), right-hand side g in lane group g
                float rr[T];
#pragma unroll
                for (int ti = 0; ti < T; ti++) rr[ti] = sm.rs[g][16 * ti + li];
                solve_fwd<T>(A, pos, rr, mpn, sm);
                solve_bwd<T>(A, pos, rr, mpn, sm);
                // x += d (row r holds unknown pos[r]); the largest correction
                // and the largest weight of each right-hand side
                double dmax = 0.0, xmax = 0.0;
#pragma unroll
                for (int ti = 0; ti < T; ti++)
                    if (16 * ti + li < mpn) {
                        const double xn = sm.xs[g][pos[ti]] + (double)rr[ti];
                        sm.xs[g][pos[ti]] = xn;
                        if (pos[ti] < m) {
                            dmax = fmax(dmax, fabs((double)rr[ti]));
                            xmax = fmax(xmax, fabs(xn));
                        }
                    }
#pragma unroll
                for (int o = 1; o < 16; o <<= 1) {
                    dmax = fmax(dmax, __shfl_xor(dmax, o));
                    xmax = fmax(xmax, __shfl_xor(xmax, o));
                }
                __syncthreads();
                // done when every right-hand side's last correction is below
                // 1e-12 of its largest weight (the next would be smaller still
                // by the contraction of the refinement), or after nref steps
                if (it == nref || __all(dmax <= 1e-12 * xmax)) break;
                // r = b - A x in FP64, one row per lane
                for (int i = lane; i < mpn; i += 64) {
                    double r0 = sm.bs[0][i], r1 = sm.bs[1][i], r2 = sm.bs[2][i], r3 = sm.bs[3][i];
                    const double *col = full + i;
#pragma unroll 8
                    for (int j = 0; j < mpn; j++) {
                        const double v = col[(size_t)j * NC];
                        r0 = __builtin_fma(-v, sm.xs[0][j], r0);
                        r1 = __builtin_fma(-v, sm.xs[1][j], r1);
                        r2 = __builtin_fma(-v, sm.xs[2][j], r2);
                        r3 = __builtin_fma(-v, sm.xs[3][j], r3);
                    }
                    sm.rs[0][i] = (float)r0; sm.rs[1][i] = (float)r1; sm.rs[2][i] = (float)r2; sm.rs[3][i] = (float)r3;
                }
                __syncthreads();
            }
            // the first m unknowns are the weights
            for (int e = lane; e < 4 * m; e += 64) {
                const int c = e / m, j = e - c * m;
                if (c < nrhs) sm.outp[c][j] = sm.xs[c][j];
            }
            __syncthreads();
        }
    }
}

using Lu32Fn = void (*)(WeightsBatch, double *, int, int);

}  // namespace

// Tile count of the matrix-core kernel for systems of size mpn (0: too wide).
int lu32_tiles(int mpn)
{
    for (int t = 2; t <= 8; t++)
        if (16 * t >= mpn) return t;
    return 0;
}

static Lu32Fn lu32_fn(int tiles)
{
#ifdef LU32_DEV_T
    return tiles == LU32_DEV_T ? lu32_kernel<LU32_DEV_T, LU32_DEV_W> : nullptr;
#else
    switch (tiles) {
    case 2: return lu32_kernel<2, 2>;
    case 3: return lu32_kernel<3, 2>;
    case 4: return lu32_kernel<4, 2>;
    case 5: return lu32_kernel<5, 2>;
    case 6: return lu32_kernel<6, 2>;
    case 7: return lu32_kernel<7, 1>;
    default: return lu32_kernel<8, 1>;
    }
#endif
}

// Most refinement steps after the first solve (MGM_LU32_REFINE overrides).
static int lu32_refinements()
{
    const char *e = getenv("MGM_LU32_REFINE");
    const int r = e ? atoi(e) : 8;
    return r < 0 ? 0 : (r > 8 ? 8 : r);
}

int launch_weights_lu32(const WeightsBatch &b, int tiles, int rmax, hipStream_t s)
{
    if (tiles < 2 || tiles > 8) MGM_FAIL(-1, "weights: the matrix-core kernel takes 2 to 8 tiles");
    int dev = 0, cus = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int waves_per_cu = tiles <= 6 ? 8 : 4;
    const int slots = cus * waves_per_cu;
    if (rmax < 1) rmax = 4;
    // scratch slots, kept for later builds (grown when a build needs more;
    // never freed, like the other per-process buffers)
    static double *scratch = nullptr;
    static size_t scratch_n = 0;
    const size_t need = (size_t)slots * lu32_slot(tiles, rmax);
    if (scratch_n < need) {
        if (scratch) HIP_TRY(hipFree(scratch));   // synchronises: no launch still reads it
        scratch = nullptr;
        scratch_n = 0;
        HIP_TRY(hipMalloc((void **)&scratch, sizeof(double) * need));
        scratch_n = need;
    }
    HIP_TRY(hipMemsetAsync(b.next, 0, sizeof(int), s));
    hipLaunchKernelGGL(lu32_fn(tiles), dim3(slots), dim3(64), 0, s, b, scratch, rmax, lu32_refinements());
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace mgm
