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
// panel's four pivot rows are copied to LDS and take their triangular update
// there, giving the operand U; one MFMA per 16 x 16 tile then applies
//     A[r][J] -= sum_q L[r][4c+q] U[q][J]
// to every trailing tile, the panel's pivot rows included: their L entries
// are the multipliers among the panel's pivots, which leaves U in place.  The factors then go to the wave's scratch slot,
// with the inverses of the 16 x 16 diagonal blocks of L and U in pivot order.
//
// Solves: up to 16 right-hand sides at once, one per lane li, in blocks of 16
// steps of the pivot order.  Forward, for block kt: the block's pivot rows of
// the running residual are gathered through LDS, multiplied by the inverse of
// L's diagonal block (four MFMA), and every row takes -L[:, kt] y_kt (one MFMA
// per tile and column group); backward the same with U from the last block.
// No step walks the pivots one by one.
//
// Refinement: x = 0, r = b; repeat { solve LU d = fl32(r); x += d; r = b - A x }
// until every right-hand side's correction is below 1e-12 of its largest
// weight (typically three refinements for PHS r^3 stencils, five for r^5).
// Phi and P are kept in FP64 in the slot (written once per system); the FP64
// residual runs on the FP64 matrix cores (v_mfma_f64_16x16x4_f64).
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

// Per-wave LDS of a system of up to 16T unknowns; the phases of a system
// share one buffer.
template <int T>
struct Lu32Smem {
    union {
        struct {                          // matrix assembly
            double cx[16 * T], cy[16 * T], cz[16 * T];
            float stage[16 * (16 * T + 4)];   // 16 rows of the FP32 matrix
        } as;
        struct {                          // factorisation
            float U[4][16 * T];           // the panel's pivot rows
            f4 L[T][16];                  // the panel's multipliers of each row (0 on pivot rows)
        } fa;
        float G[T * 256];                 // diagonal blocks of the factors in pivot order
        float R[16 * (16 * T + 4)];       // residual rows by right-hand side (solve gathers)
    };
    f4 ys[T * 64];                        // forward results of a solve, by block and lane
    float diag[16 * T];                   // pivot value of each step
    int piv[16 * T];                      // pivot row of each step
    double *outp[16];
};

__device__ inline float bcast(float v, int lane) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane)); }
__device__ inline int bcast(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// The workgroup is one wave: LDS accesses of a wave complete in program
// order, so ordering them needs only a compiler barrier, not s_barrier.
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

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
                    *reinterpret_cast<f4 *>(&sm.fa.U[q][16 * tj + 4 * g]) = a[ti][tj];
            }
    // the panel's multipliers of each row: column q of the panel where the
    // row was not yet pivot at step 4c + q, else 0.  The panel's own pivot
    // rows keep the multipliers of the earlier pivots among them, so the
    // rank-4 update below also turns their trailing entries into those of U
    // (the triangular solve with the panel's unit lower block).
    if (g == gc) {
#pragma unroll
        for (int ti = 0; ti < T; ti++) {
            f4 l;
#pragma unroll
            for (int q = 0; q < 4; q++) l[q] = pos[ti] > 4 * c + q ? a[ti][0][q] : 0.f;
            sm.fa.L[ti][li] = l;
        }
    }
    wave_sync();
    // triangular update of the pivot rows' trailing columns in LDS (storage
    // columns 4gc+4 .. 16nt-1): U[q][j] -= sum_{q' < q} L11[q][q'] U[q'][j],
    // L11[q][q'] = U[q][4gc + q'], the operand of the rank-4 update
    const int jb = 4 * gc;
    for (int j = jb + 4 + lane; j < 16 * nt; j += 64) {
        const float u0 = sm.fa.U[0][j];
        if (nq > 1) {
            const float u1 = __builtin_fmaf(-sm.fa.U[1][jb], u0, sm.fa.U[1][j]);
            sm.fa.U[1][j] = u1;
            if (nq > 2) {
                const float u2 = __builtin_fmaf(-sm.fa.U[2][jb + 1], u1, __builtin_fmaf(-sm.fa.U[2][jb], u0, sm.fa.U[2][j]));
                sm.fa.U[2][j] = u2;
                if (nq > 3)
                    sm.fa.U[3][j] = __builtin_fmaf(-sm.fa.U[3][jb + 2], u2,
                                                __builtin_fmaf(-sm.fa.U[3][jb + 1], u1,
                                                               __builtin_fmaf(-sm.fa.U[3][jb], u0, sm.fa.U[3][j])));
            }
        }
    }
    wave_sync();
    // rank-4 update of the trailing tiles on the matrix cores:
    // C'[x][y] = A[16ti + y][16tj + x] -= sum_k U[k][16tj + x] L[16ti + y][4c + k]
    float bop[T];
#pragma unroll
    for (int ti = 0; ti < T; ti++) {
        const f4 l4 = sm.fa.L[ti][li];
        bop[ti] = group_sel(l4[0], l4[1], l4[2], l4[3], mg);
    }
#pragma unroll
    for (int tj = 0; tj < T; tj++) {
        if (tj >= nt || (tj == 0 && gc == 3)) continue;
        const int j = 16 * tj + li;
        const float aop = (j >= jb + 4 && g < nq) ? -sm.fa.U[g][j] : 0.f;
#pragma unroll
        for (int ti = 0; ti < T; ti++) a[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(aop, bop[ti], a[ti][tj], 0, 0, 0);
    }
    wave_sync();   // sm.U / sm.L are rewritten by the next panel
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

// Right-hand sides and residuals in the solve layout: for the rows of tile
// ti, lane 16 g + li holds in register r the value of row 16 ti + 4 g + r of
// right-hand side li (the MFMA result layout), up to 16 right-hand sides.

// The rows piv(16 kt + 4 g + r) (packed pb, a byte per r) of the residual v
// in the solve layout, through LDS.
template <int T>
__device__ __attribute__((always_inline)) inline f4 gather_rows(const f4 (&v)[T], unsigned pb, float *R)
{
    constexpr int RS = 16 * T + 4;
    const int lane = threadIdx.x & 63, li = lane & 15, g = lane >> 4;
    wave_sync();   // the previous gather's reads are done
#pragma unroll
    for (int ti = 0; ti < T; ti++) *reinterpret_cast<f4 *>(&R[li * RS + 16 * ti + 4 * g]) = v[ti];
    wave_sync();
    f4 out;
#pragma unroll
    for (int r = 0; r < 4; r++) out[r] = R[li * RS + ((pb >> (8 * r)) & 255u)];
    return out;
}

// One solve LU d = r of up to 16 right-hand sides by 16 x 16 blocks of the
// pivot order, on the matrix cores.  Forward: for block kt, gather the block's
// pivot rows of the running residual, multiply by the inverse of the unit lower
// diagonal block (y_kt), subtract L[:, kt] y_kt from every row (one MFMA per
// tile and column group).  Backward: the same with the upper diagonal block's
// inverse, from the last block, subtracting U[:, kt] x_kt.  Rows a step has
// consumed are updated too; they are never gathered again.  The correction of
// block kt (unknowns 16 kt + 4 g + r) is passed to emit(kt, x).
template <int T, typename Emit>
__device__ __attribute__((always_inline)) inline void solve_blocks(const f4 *fac, f4 (&rr)[T], const unsigned (&pb)[T],
                                                                   const f4 *linv, const f4 *uinv, float *R, f4 *ys,
                                                                   Emit emit)
{
    const int lane = threadIdx.x & 63, li = lane & 15, g = lane >> 4;
#pragma unroll
    for (int kt = 0; kt < T; kt++) {
        __builtin_amdgcn_sched_barrier(0);
        f4 col[T];   // tile column kt of the factors
#pragma unroll
        for (int ti = 0; ti < T; ti++) col[ti] = fac[(ti * T + kt) * 64 + lane];
        const f4 lf = linv[kt * 64 + 4 * li + g];   // row li, columns 4g .. 4g+3
        const f4 yt = gather_rows<T>(rr, pb[kt], R);
        f4 y = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s2 = 0; s2 < 4; s2++) y = __builtin_amdgcn_mfma_f32_16x16x4f32(lf[s2], yt[s2], y, 0, 0, 0);
        ys[kt * 64 + lane] = y;
        if (kt + 1 < T) {
            const f4 ny = -y;
#pragma unroll
            for (int ti = 0; ti < T; ti++)
#pragma unroll
                for (int q = 0; q < 4; q++) rr[ti] = __builtin_amdgcn_mfma_f32_16x16x4f32(col[ti][q], ny[q], rr[ti], 0, 0, 0);
        }
    }
#pragma unroll
    for (int ti = 0; ti < T; ti++) rr[ti] = (f4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = T - 1; kt >= 0; kt--) {
        __builtin_amdgcn_sched_barrier(0);
        f4 col[T];
        if (kt > 0) {
#pragma unroll
            for (int ti = 0; ti < T; ti++) col[ti] = fac[(ti * T + kt) * 64 + lane];
        }
        const f4 uf = uinv[kt * 64 + 4 * li + g];
        f4 yt = ys[kt * 64 + lane];
        if (kt + 1 < T) yt += gather_rows<T>(rr, pb[kt], R);
        f4 x = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s2 = 0; s2 < 4; s2++) x = __builtin_amdgcn_mfma_f32_16x16x4f32(uf[s2], yt[s2], x, 0, 0, 0);
        if (kt > 0) {
            const f4 nx = -x;
#pragma unroll
            for (int ti = 0; ti < T; ti++)
#pragma unroll
                for (int q = 0; q < 4; q++) rr[ti] = __builtin_amdgcn_mfma_f32_16x16x4f32(col[ti][q], nx[q], rr[ti], 0, 0, 0);
        }
        emit(kt, x);
    }
}

// Inverses of the diagonal 16 x 16 blocks of the factors in pivot order,
// G_kt[i][j] = A[piv(16 kt + i)][16 kt + j] (L strictly below the diagonal,
// with a unit diagonal; U on and above it): one lane per (block, column),
// substitution against the unit vector.  Written row-major, 256 floats per
// block: the inverses of L's blocks, then of U's.
template <int T>
__device__ __attribute__((always_inline)) inline void block_inverses(const float *G, float *inv)
{
    int t0 = threadIdx.x & 63;
    asm volatile("" : "+v"(t0));   // keeps the unit-vector selects inside the loop over systems
#pragma unroll 1
    for (int t = t0; t < 16 * T; t += 64) {
        const float *Gb = G + (t >> 4) * 256;
        const int j = t & 15;
        float y[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            float v = i == j ? 1.f : 0.f;
#pragma unroll
            for (int k = 0; k < i; k++) v = __builtin_fmaf(-Gb[i * 16 + k], y[k], v);
            y[i] = v;
        }
        float *Li = inv + (t >> 4) * 256 + j;
#pragma unroll
        for (int i = 0; i < 16; i++) Li[i * 16] = y[i];
#pragma unroll
        for (int i = 15; i >= 0; i--) {
            float v = i == j ? 1.f : 0.f;
#pragma unroll
            for (int k = i + 1; k < 16; k++) v = __builtin_fmaf(-Gb[i * 16 + k], y[k], v);
            y[i] = v * __builtin_amdgcn_rcpf(Gb[i * 17]);   // FP32 factors are refined anyway
        }
        float *Ui = inv + (T + (t >> 4)) * 256 + j;
#pragma unroll
        for (int i = 0; i < 16; i++) Ui[i * 16] = y[i];
    }
}

// Scratch slot of a wave, in doubles: the FP64 system A padded to 16T x 16T
// with zeros, column-major (A[i][j] at j * 16T + i); the right-hand sides of
// the system [r][i] (rmax x 16T); the FP64 solution of the current pass in
// the solve layout (T x 64 lanes x 4); the inverses of the diagonal blocks
// (2T x 256 floats); the FP32 factors, tile (ti, tj) as 64 lanes x 4 floats
// (T^2 x 256 floats).
__host__ __device__ inline long lu32_slot(int tiles, int rmax)
{
    return (long)(16 * tiles) * (16 * tiles + rmax) + 512L * tiles + 128L * tiles * tiles;
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
    const double *cx = sm.as.cx, *cy = sm.as.cy, *cz = sm.as.cz;
    const double sx = cx[0], sy = cy[0], sz = cz[0];
    for (int e = lane; e < NC * NC; e += 64) {
        const int j = e / NC, i = e - j * NC;
        double v = 0.0;
        if (i < m && j < m) {
            if (i != j) v = rbf_phs(dist_sum(cx[i], cy[i], cz[i], cx[j], cy[j], cz[j], dim), phs);
        } else if (i < mpn && j < mpn && (i < m || j < m)) {
            const int pt = i < m ? i : j, k = (i < m ? j : i) - m;
            v = monomial(cx[pt] - sx, cy[pt] - sy, cz[pt] - sz, a.px[k], a.py[k], a.pz[k]);
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
            v = q < nops ? operator_rhs(nth_op(opmask, q), i, m, dim, phs, a.px, a.py, a.pz, cx, cy, cz, sx, sy, sz)
                         : target_rhs(i, m, dim, phs, a.px, a.py, a.pz, cx, cy, cz, sx, sy, sz, tg[0], tg[1], tg[2]);
        }
        rb[e] = v;
    }
}

typedef double d4 __attribute__((ext_vector_type(4)));

// WPS: waves per SIMD the register budget aims at.
template <int T, int WPS>
__global__ void __launch_bounds__(64, WPS)
lu32_kernel(WeightsBatch bt, double *scratch, int rmax, int nref)
{
    __shared__ Lu32Smem<T> sm;
    constexpr int NC = 16 * T;
    const int lane = threadIdx.x & 63, li = lane & 15, g = lane >> 4;
    double *const slot = scratch + (size_t)blockIdx.x * lu32_slot(T, rmax);
    for (;;) {
        // the slot's addresses are formed per system: hoisted out of this loop
        // (they are loop-invariant) they would not fit the registers
        double *full = slot;                                            // [j][i], NC x NC
        asm volatile("" : "+s"(full));
        double *rb = full + (size_t)NC * NC;                            // [r][i], rmax x NC
        d4 *xg = reinterpret_cast<d4 *>(rb + (size_t)rmax * NC);        // [kt][lane], solve layout
        float *inv = reinterpret_cast<float *>(xg + 64 * T);            // 2T blocks of 16 x 16
        const f4 *linv = reinterpret_cast<const f4 *>(inv), *uinv = linv + 64 * T;
        f4 *fac = reinterpret_cast<f4 *>(inv + 512 * T);                // [ti][tj][lane]
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
            sm.as.cx[i] = in ? a.x[c] : 0.0;
            sm.as.cy[i] = in ? a.y[c] : 0.0;
            sm.as.cz[i] = in ? a.z[c] : 0.0;
        }
        wave_sync();
        const int nr = a.rt_off ? a.rt_off[s + 1] - a.rt_off[s] : 0;
        const int np = a.pt_off ? a.pt_off[s + 1] - a.pt_off[s] : 0;
        const int total = nops + nr + np;
        fill_system<T>(a, s, nops, nr, total, opmask, full, rb, sm);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        wave_sync();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // no stale cache lines of the slot
        // FP32 matrix in registers, one tile row at a time through LDS: rows
        // 16 ti .. 16 ti + 15 staged row-major with a padded stride, then each
        // lane reads its 4T entries
        f4 A[T][T];
        int pos[T];
        float *stage = sm.as.stage;
        constexpr int SST = NC + 4;
#pragma unroll
        for (int ti = 0; ti < T; ti++) {
            pos[ti] = 16 * ti + li < mpn ? kNotPivot : -1;
            for (int e2 = lane; e2 < 16 * NC; e2 += 64) {
                const int j = e2 >> 4, r = e2 & 15;
                stage[r * SST + j] = (float)full[(size_t)j * NC + 16 * ti + r];
            }
            wave_sync();
#pragma unroll
            for (int tj = 0; tj < T; tj++) A[ti][tj] = *reinterpret_cast<const f4 *>(&stage[li * SST + 16 * tj + 4 * g]);
            wave_sync();
        }
        if (!lu_factor<T>(A, pos, mpn, sm)) {
            if (lane == 0) atomicOr(bt.status, 1);
            wave_sync();
            continue;
        }
        // padding rows and columns are zero: step k >= mpn takes row k
#pragma unroll
        for (int ti = 0; ti < T; ti++)
            if (pos[ti] < 0) pos[ti] = 16 * ti + li;
        if (lane < NC - mpn) sm.piv[mpn + lane] = mpn + lane;
        // the diagonal blocks in pivot order to LDS, their inverses to the slot
#pragma unroll
        for (int ti = 0; ti < T; ti++) {
            const int kb = pos[ti] >> 4, i = pos[ti] & 15;
#pragma unroll
            for (int tj = 0; tj < T; tj++)
                if (kb == tj) *reinterpret_cast<f4 *>(&sm.G[tj * 256 + i * 16 + 4 * g]) = A[ti][tj];
        }
        wave_sync();
        if (lane < NC - mpn) {
            const int k = mpn + lane;
            sm.G[(k >> 4) * 256 + (k & 15) * 17] = 1.f;   // unit pivots of the padding steps
        }
        // the factors leave the registers: the solves read them tile column by
        // tile column from the slot
#pragma unroll
        for (int ti = 0; ti < T; ti++)
#pragma unroll
            for (int tj = 0; tj < T; tj++) fac[(ti * T + tj) * 64 + lane] = A[ti][tj];
        wave_sync();
        block_inverses<T>(sm.G, inv);
        unsigned pb[T];   // pivot rows of steps 16 kt + 4 g + r, a byte per r
#pragma unroll
        for (int kt = 0; kt < T; kt++) {
            const int4 pv = *reinterpret_cast<const int4 *>(&sm.piv[16 * kt + 4 * g]);
            pb[kt] = (unsigned)pv.x | ((unsigned)pv.y << 8) | ((unsigned)pv.z << 16) | ((unsigned)pv.w << 24);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        wave_sync();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // the inverses, written by other lanes
        // right-hand sides in passes of 16 (one per lane li)
        for (int q0 = 0; q0 < total; q0 += 16) {
            const int nrhs = total - q0 < 16 ? total - q0 : 16;
            const bool has = li < nrhs;
            if (g == 0 && has) {
                const double *tg;
                sm.outp[li] = rhs_out(a, s, q0 + li, nops, nr, opmask, &tg);
            }
            const d4 *b4 = reinterpret_cast<const d4 *>(rb + (size_t)(q0 + (has ? li : 0)) * NC);
            f4 rr[T];
#pragma unroll
            for (int ti = 0; ti < T; ti++) {
                const d4 bv = has ? b4[4 * ti + g] : (d4){0.0, 0.0, 0.0, 0.0};
                rr[ti] = (f4){(float)bv[0], (float)bv[1], (float)bv[2], (float)bv[3]};
                xg[ti * 64 + lane] = (d4){0.0, 0.0, 0.0, 0.0};
            }
            for (int it = 0;; it++) {
                // the slot's addresses are formed here, each step: hoisted out
                // of the loop over systems they would not fit the registers
                const f4 *fac_i = fac, *linv_i = linv, *uinv_i = uinv;
                const double *full_i = full;
                asm volatile("" : "+s"(fac_i), "+s"(linv_i), "+s"(uinv_i), "+s"(full_i));
                // x += (LU)^-1 fl32(r); the largest correction and the largest
                // weight of each right-hand side (first m unknowns)
                double dmax = 0.0, xmax = 0.0;
                solve_blocks<T>(fac_i, rr, pb, linv_i, uinv_i, sm.R, sm.ys, [&](int kt, f4 x) {
                    const d4 xo = xg[kt * 64 + lane];
                    d4 xn;
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        xn[r] = xo[r] + (double)x[r];
                        if (16 * kt + 4 * g + r < m) {
                            dmax = fmax(dmax, fabs((double)x[r]));
                            xmax = fmax(xmax, fabs(xn[r]));
                        }
                    }
                    xg[kt * 64 + lane] = xn;
                });
                dmax = fmax(dmax, __shfl_xor(dmax, 16));
                dmax = fmax(dmax, __shfl_xor(dmax, 32));
                xmax = fmax(xmax, __shfl_xor(xmax, 16));
                xmax = fmax(xmax, __shfl_xor(xmax, 32));
                // done when every right-hand side's last correction is below
                // 1e-12 of its largest weight (the next would be smaller still
                // by the contraction of the refinement), or after nref steps
                if (it == nref || __all(dmax <= 1e-12 * xmax)) break;
                // r = b - A x in FP64 on the matrix cores: for row tile bi,
                // k-slices of four unknowns 16 kt + 4 g + r (r fixed).  The
                // FP64 MFMA leaves result row i in lane group i mod 4, register
                // i / 4, so the operand's row li is matrix row
                // 16 bi + 4 (li mod 4) + li / 4: the result is then in the solve
                // layout (row 16 bi + 4 g + r in lane group g, register r)
                d4 xk[T];
#pragma unroll
                for (int kt = 0; kt < T; kt++) xk[kt] = -xg[kt * 64 + lane];
#pragma unroll
                for (int bi = 0; bi < T; bi++) {
                    __builtin_amdgcn_sched_barrier(0);   // one row tile's loads in flight at a time
                    d4 acc = has ? b4[4 * bi + g] : (d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int kt = 0; kt < T; kt++)
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const double av = full_i[(size_t)(16 * kt + 4 * g + r) * NC + 16 * bi + 4 * (li & 3) + (li >> 2)];
                            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, xk[kt][r], acc, 0, 0, 0);
                        }
                    rr[bi] = (f4){(float)acc[0], (float)acc[1], (float)acc[2], (float)acc[3]};
                }
            }
            // the first m unknowns are the weights
            wave_sync();
#pragma unroll
            for (int kt = 0; kt < T; kt++) {
                const d4 xv = xg[kt * 64 + lane];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int j = 16 * kt + 4 * g + r;
                    if (has && j < m) sm.outp[li][j] = xv[r];
                }
            }
            wave_sync();
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
