// stencil.hip -- FP64 fixed-width (sliced-ELL) stencil kernels of the V-cycle.
//
// Layout: rows in slices of S consecutive device rows; within a slice the m
// columns follow each other and each column holds the slice's S entries
// contiguously (ell_at in device.hpp).  Device order: Morton-ranked blocks of
// rows, colour classes inside a block, Morton order inside a class, each
// (block, colour) chunk padded to whole slices; a colour step walks that
// colour's slice list and its neighbour gathers stay inside a few blocks.
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

#include <cstdlib>

#include "device.hpp"

namespace mgm {

namespace {

constexpr int BLK = 256;
constexpr int WAVES = BLK / 64;
constexpr int SMALL_BLK_MAX = 1024;
constexpr int GS_BLK = 256;   // gather_sort_kernel workgroup
// dynamic LDS of the single-workgroup relax: gfx950 gives a workgroup 160 KB
constexpr size_t kSmallLdsMax = 160 * 1024 - 1024;
// Staging slots per row in relax_small_kernel: the wider of the two
// operators, and at least the MT slots its sums read.
__host__ __device__ inline int small_slots(int mw, int mt) { return mw > mt ? mw : mt; }

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
    M_JACOBI = 8,    // the SOR update of every row from the previous iterate x, into y
    M_RESJAC = 9,    // M_RES0 into y2 and M_JACOBI into y, from one set of products
};

__device__ inline bool skips_boundary(int mode)
{
    return mode == M_RES0 || mode == M_RES1 || mode == M_RES3 || mode == M_RESTRICT || mode == M_PROLONG;
}

// Wave-scope barrier: LDS written by some lanes of a wave is visible to the
// other lanes of that wave afterwards.  Each wave stages into its own LDS
// region, so waves never wait for each other.
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Non-temporal load (streamed slab entries, read once per pass); BF16 structs
// load plainly.
template <typename T> __device__ inline T ldnt(const T *p) { return __builtin_nontemporal_load(p); }
template <> __device__ inline bf16 ldnt(const bf16 *p) { return *p; }

// Weight conversion: FP64 and FP32 weights are used as stored, BF16 weights
// (upper half of an FP32) widen exactly to FP32.
__device__ inline double wv(double v) { return v; }
__device__ inline float wv(float v) { return v; }
__device__ inline float wv(bf16 v) { return __uint_as_float((unsigned)v.bits << 16); }

// Stage the products w*x of a lane's stencil entries j = jo, jo+LPR, ... of
// its row in LDS at P[j*S + ri] (base = the row's entry 0 in the slab).
// NIT == 0: loop unrolled by 4, loads and gathers interleaved -- best for
// streaming passes, many waves per CU hide the latency.  NIT > 0: all weight
// and column loads of a chunk of NIT entries first, then all gathers -- one
// memory round trip per chunk, for latency-bound single-workgroup sweeps.
template <int S, int NIT, typename TW, typename TV>
__device__ inline void stage_products(const EllT<TW> &A, size_t base, int ri, int jo, const TV *x, TV *P)
{
    constexpr int LPR = 64 / S;
    const int m = A.m, ms = prod_stride(m);
    if (A.sorted) {
        // Gather-sorted slab: lane l takes the slice's entries l, l+64, ...
        // in column order and writes each product to the place of its plain
        // slot j*S + ri: row ri, entry j.
        const int n = S * m;
        const size_t sb = base - ri;
        const int lane = jo * S + ri;
        const TW *wp = A.w + sb;
        const unsigned *cp = reinterpret_cast<const unsigned *>(A.col) + sb;
        const int sh = A.shift;
        const unsigned mask = (1u << sh) - 1;
        if (A.stream) {
#pragma unroll 4
            for (int q = lane; q < n; q += 64) {
                const unsigned v = ldnt(cp + q), slot = v >> sh;
                P[(slot % S) * ms + slot / S] = wv(ldnt(wp + q)) * x[v & mask];
            }
        } else {
#pragma unroll 4
            for (int q = lane; q < n; q += 64) {
                const unsigned v = cp[q], slot = v >> sh;
                P[(slot % S) * ms + slot / S] = wv(wp[q]) * x[v & mask];
            }
        }
        return;
    }
    // entry j = t*LPR + jo sits at base + j*S = lane base + t*64; its product
    // at row ri, place j
    const TW *wp = A.w + base + (size_t)jo * S;
    const int *cp = A.col + base + (size_t)jo * S;
    TV *pp = P + ri * ms + jo;
    const int nt = (m - jo + LPR - 1) / LPR;   // entries of this lane
    if constexpr (NIT == 0) {
        if (A.stream) {
#pragma unroll 4
            for (int t = 0; t < nt; t++) pp[t * LPR] = wv(ldnt(wp + t * 64)) * x[ldnt(cp + t * 64)];
        } else {
#pragma unroll 4
            for (int t = 0; t < nt; t++) pp[t * LPR] = wv(wp[t * 64]) * x[cp[t * 64]];
        }
    } else {
        for (int t0 = 0; t0 < nt; t0 += NIT) {
            TW wq[NIT];
            int cq[NIT];
            TV xq[NIT];
#pragma unroll
            for (int t = 0; t < NIT; t++)
                if (t0 + t < nt) {
                    wq[t] = wp[(t0 + t) * 64];
                    cq[t] = cp[(t0 + t) * 64];
                }
#pragma unroll
            for (int t = 0; t < NIT; t++)
                if (t0 + t < nt) xq[t] = x[cq[t]];
#pragma unroll
            for (int t = 0; t < NIT; t++)
                if (t0 + t < nt) pp[(t0 + t) * LPR] = wv(wq[t]) * xq[t];
        }
    }
}

// acc (op) P[j*S + ri] for j = j0, j0+1, ..., m-1 in that order, op = + or -:
// the reference's left-to-right row sum.  Products are read eight at a time
// (one LDS wait per eight dependent adds, no selects on the chain), the tail
// one by one.
template <bool SUBTRACT, typename TV>
__device__ inline TV ordered_sum(const TV *P, int ri, int j0, int m, TV acc)
{
    // row ri's products, added (or subtracted) one after another in slot order
    // from j0; read 16 bytes at a time where aligned, eight slots per wait
    constexpr int NV = 16 / (int)sizeof(TV);
    typedef TV vec __attribute__((ext_vector_type(NV)));
    const TV *q = P + (size_t)ri * prod_stride(m);
    int j = j0;
    for (; j < m && (j & (NV - 1)); j++) acc = SUBTRACT ? acc - q[j] : acc + q[j];
    for (; j + 8 <= m; j += 8) {
        TV v[8];
#pragma unroll
        for (int u = 0; u < 8 / NV; u++) {
            const vec t = *reinterpret_cast<const vec *>(q + j + u * NV);
#pragma unroll
            for (int e = 0; e < NV; e++) v[u * NV + e] = t[e];
        }
#pragma unroll
        for (int t = 0; t < 8; t++) acc = SUBTRACT ? acc - v[t] : acc + v[t];
    }
    for (; j < m; j++) acc = SUBTRACT ? acc - q[j] : acc + q[j];
    return acc;
}

// The sums ordered_sum<false> forms from +0 over slots 0..m-1 (*a) and over
// slots 1..m-1 (*b), in one walk over the slots.
template <typename TV>
__device__ inline void ordered_sum2(const TV *P, int ri, int m, TV *a, TV *b)
{
    constexpr int NV = 16 / (int)sizeof(TV);
    typedef TV vec __attribute__((ext_vector_type(NV)));
    const TV *q = P + (size_t)ri * prod_stride(m);
    TV sa = TV(0) + q[0], sb = TV(0);
    int j = 1;
    for (; j < m && (j & (NV - 1)); j++) {
        sa = sa + q[j];
        sb = sb + q[j];
    }
    for (; j + 8 <= m; j += 8) {
        TV v[8];
#pragma unroll
        for (int u = 0; u < 8 / NV; u++) {
            const vec t = *reinterpret_cast<const vec *>(q + j + u * NV);
#pragma unroll
            for (int e = 0; e < NV; e++) v[u * NV + e] = t[e];
        }
#pragma unroll
        for (int t = 0; t < 8; t++) {
            sa = sa + v[t];
            sb = sb + v[t];
        }
    }
    for (; j < m; j++) {
        sa = sa + q[j];
        sb = sb + q[j];
    }
    *a = sa;
    *b = sb;
}

// One wave applies the rows of one slice: lanes (ri, jo) stage the products
// of row ri (stage_products); after a wave barrier lane jo == 0 combines its
// row in stencil order.  TW is the stored weight type, TV the type of
// vectors, products and sums.
template <int S, typename TW, typename TV, int NIT = 0>
__device__ inline void slice_rows(const EllT<TW> &A, long slice, int nrows, int mode,
                                  const uint8_t *__restrict__ btag, const TV *__restrict__ src,
                                  const TV *x, TV *y, TV w, TV c, int keep_boundary, TV *P)
{
    const int lane = threadIdx.x & 63;
    const int ri = lane % S, jo = lane / S;
    const int m = A.m;
    const long r = slice * S + ri;
    const bool inr = r < nrows;
    const uint8_t tag = (inr && btag) ? btag[r] : (uint8_t)(inr ? ROW_INTERIOR : ROW_PAD);
    const bool act = tag != ROW_PAD;
    const bool bnd = tag == ROW_BOUNDARY;
    // Modes that skip boundary rows wait for the tag before loading (saves the
    // boundary rows' traffic); the others load at once -- padding rows hold
    // weight 0 and column 0, so their loads are harmless and never combined.
    const bool need = skips_boundary(mode) ? (act && !bnd) : inr;
    const size_t base = (size_t)slice * S * m + ri;
    // the combining lane's row operands, loaded alongside the stencil
    const bool comb = jo == 0 && act;
    TV sv = 0, yv = 0, dg = 1;
    if (comb && mode != M_SPMV && mode != M_RESTRICT && mode != M_PROLONG && !(bnd && skips_boundary(mode)))
        sv = src[r];
    if (comb && (mode == M_SOR || (mode == M_PROLONG && !bnd))) yv = y[r];
    if (comb && mode == M_JACOBI) yv = x[r];
    if (comb && (mode == M_SOR || mode == M_JACOBI)) dg = wv((A.sorted ? A.w0 : A.w)[base]);
    // a gather-sorted slice mixes its rows over the lanes: every lane loads
    if (need || A.sorted) stage_products<S, NIT, TW, TV>(A, base, ri, jo, x, P);
    wave_sync();
    if (jo == 0 && act) {
        switch (mode) {
        case M_SPMV:
            y[r] = ordered_sum<false>(P, ri, 0, m, TV(0));
            break;
        case M_SOR:
        case M_JACOBI: {
            const TV sum = ordered_sum<false>(P, ri, 1, m, TV(0));
            y[r] = w * ((sv - sum) / dg) + c * yv;
            break;
        }
        case M_RES0:
        case M_RES3: {
            if (bnd) {
                if (mode == M_RES0) y[r] = 0;
                break;
            }
            y[r] = sv - ordered_sum<false>(P, ri, 0, m, TV(0));
            break;
        }
        case M_RES1:
        case M_RES2: {
            if (mode == M_RES1 && bnd) { y[r] = 0; break; }
            y[r] = ordered_sum<true>(P, ri, 0, m, sv);
            break;
        }
        case M_RESTRICT: {
            if (bnd) {
                if (!keep_boundary) y[r] = 0;
                break;
            }
            y[r] = ordered_sum<false>(P, ri, 0, m, TV(0));
            break;
        }
        case M_PROLONG:
            if (bnd) break;
            y[r] = yv + ordered_sum<false>(P, ri, 0, m, TV(0));
            break;
        }
    }
    wave_sync();  // P is reused by the wave's next slice
}

// Slice k of the launch is slices[k] when a slice list is given (one colour
// class), else first + k (a contiguous colour class, or a full pass).  Rows at or beyond nrows and padding
// rows (tag ROW_PAD) are neither read nor written.
template <int S, typename TW, typename TV, int W>
__global__ void __launch_bounds__(64 * W)
rows_kernel(EllT<TW> A, const int *__restrict__ slices, int first, int nslices, int nrows, int mode,
            const uint8_t *__restrict__ btag, const TV *__restrict__ src, const TV *x, TV *y, TV w,
            TV c, int keep_boundary, int nblocks)
{
    // Make the slab and vector pointers live before the early exit below, so
    // their kernel-argument loads issue with the others at wave start instead
    // of after it: one scalar-load wait per wave instead of two in series.
    // The statement emits no instruction.  Colour-step launches are short
    // enough for this to show: 2.7% over a whole solve (measured).
    asm volatile("" ::"s"(A.col), "s"(A.w), "s"(x), "s"(y));
    extern __shared__ unsigned char lds_raw[];
    TV *prod = reinterpret_cast<TV *>(lds_raw);   // [W][S][prod_stride(m)]
    const int wave = threadIdx.x >> 6;
    const long k = xcd_block(blockIdx.x, nblocks) * W + wave;   // nblocks = gridDim.x
    if (k >= nslices) return;
    const long slice = slices ? (long)slices[k] : (long)first + k;
    slice_rows<S, TW, TV>(A, slice, nrows, mode, btag, src, x, y, w, c, keep_boundary,
                          prod + (size_t)wave * prod_stride(A.m) * S);
}

// One slice of a full pass in the LDS-window form: as slice_rows on the
// plain slab order, with x read from the block's window in LDS (win_col holds
// each entry's place there).  Rows that the mode skips load nothing.
template <int S, typename TW, typename TV>
__device__ inline void window_slice(const EllT<TW> &A, const TW *wp, long slice, int mode,
                                    const uint8_t *__restrict__ btag, const TV *__restrict__ src, const TV *win,
                                    int r0, TV *y, TV *y2, TV *P, TV wf, TV cf)
{
    constexpr int LPR = 64 / S;
    const int lane = threadIdx.x & 63;
    const int ri = lane % S, jo = lane / S;
    const int m = A.m;
    const long r = slice * S + ri;
    const uint8_t tag = btag[r];
    const bool act = tag != ROW_PAD, bnd = tag == ROW_BOUNDARY;
    const bool skip = bnd && skips_boundary(mode);
    const bool comb = jo == 0 && act;
    TV sv = 0, yv = 0, dg = 1;
    if (comb && mode != M_SPMV && mode != M_PROLONG && mode != M_RESTRICT && !skip) sv = src[r];
    if (comb && mode == M_PROLONG && !bnd) yv = y[r];
    if (comb && (mode == M_JACOBI || mode == M_RESJAC)) {
        yv = win[r - r0];   // the row's own previous value: own rows lead the window
        dg = wv(wp[(size_t)slice * S * m + ri]);
    }
    if (act && !skip) {
        const size_t base = (size_t)slice * S * m + (size_t)jo * S + ri;
        const TW *w = wp + base;
        const uint16_t *c = A.win_col + base;
        TV *pp = P + ri * prod_stride(m) + jo;
        const int nt = (m - jo + LPR - 1) / LPR;
        if (A.stream) {
#pragma unroll 4
            for (int t = 0; t < nt; t++) pp[t * LPR] = wv(ldnt(w + t * 64)) * win[ldnt(c + t * 64)];
        } else {
#pragma unroll 4
            for (int t = 0; t < nt; t++) pp[t * LPR] = wv(w[t * 64]) * win[c[t * 64]];
        }
    }
    wave_sync();
    if (comb) {
        switch (mode) {
        case M_SPMV:
            y[r] = ordered_sum<false>(P, ri, 0, m, TV(0));
            break;
        case M_RES0:
        case M_RES3:
            if (bnd) {
                if (mode == M_RES0) y[r] = 0;
                break;
            }
            y[r] = sv - ordered_sum<false>(P, ri, 0, m, TV(0));
            break;
        case M_RES1:
        case M_RES2:
            if (mode == M_RES1 && bnd) {
                y[r] = 0;
                break;
            }
            y[r] = ordered_sum<true>(P, ri, 0, m, sv);
            break;
        case M_JACOBI: {
            const TV sum = ordered_sum<false>(P, ri, 1, m, TV(0));
            y[r] = wf * ((sv - sum) / dg) + cf * yv;
            break;
        }
        case M_RESJAC: {
            TV all, off;
            ordered_sum2(P, ri, m, &all, &off);
            y[r] = wf * ((sv - off) / dg) + cf * yv;
            y2[r] = bnd ? TV(0) : sv - all;
            break;
        }
        case M_PROLONG:
            if (bnd) break;
            y[r] = yv + ordered_sum<false>(P, ri, 0, m, TV(0));
            break;
        case M_RESTRICT:   // wf != 0: keep boundary rows (launch_rows passes the flag there)
            if (bnd) {
                if (wf == TV(0)) y[r] = 0;
                break;
            }
            y[r] = ordered_sum<false>(P, ri, 0, m, TV(0));
            break;
        }
    }
    wave_sync();   // P is reused by the wave's next slice
}

// Full pass (SpMV or a residual form) over a level in Morton blocks, one
// workgroup per block: the block's x window -- its own rows, then its halo
// rows -- is loaded into LDS once (coalesced own rows, one gather per halo
// row), and every stencil entry reads x there instead of gathering it from
// memory: a block of 4096 rows gathers ~4k halo values instead of ~290k
// stencil entries.  The products and ordered sums are those of rows_kernel,
// so the results are bitwise the same.
template <int S, typename TW, typename TV>
__global__ void __launch_bounds__(64 * kWindowWaves)
window_rows_kernel(EllT<TW> A, int mode, const uint8_t *__restrict__ btag, const TV *__restrict__ src,
                   const TV *__restrict__ x, TV *y, TV *y2, TV wf, TV cf)
{
    extern __shared__ unsigned char lds_raw[];
    constexpr int W = kWindowWaves;
    const int m = A.m;
    const int wave = threadIdx.x >> 6;
    TV *P = reinterpret_cast<TV *>(lds_raw) + (size_t)wave * prod_stride(m) * S;
    TV *win = reinterpret_cast<TV *>(lds_raw) + (size_t)W * prod_stride(m) * S;
    const int b = (int)xcd_block(blockIdx.x, gridDim.x);
    const int r0 = A.win_row0[b], r1 = A.win_row0[b + 1];
    const int h0 = A.win_hoff[b], nh = A.win_hoff[b + 1] - h0, nown = A.win_self ? r1 - r0 : 0;
    for (int i = threadIdx.x; i < nown; i += 64 * W) win[i] = x[r0 + i];
    for (int i = threadIdx.x; i < nh; i += 64 * W) win[nown + i] = x[A.win_halo[h0 + i]];
    __syncthreads();
    const TW *wp = A.sorted ? A.w0 : A.w;
    for (int sl = r0 / S + wave; sl < r1 / S; sl += W)
        window_slice<S>(A, wp, sl, mode, btag, src, win, r0, y, y2, P, wf, cf);
}

// Weights and columns of this lane's entries j = t*LPR + jo < MT of one
// slice (entry j of row ri sits at slice*S*m + j*S + ri).  Slots j >= m load
// entry m-1 again: every load is in the row's range and none is guarded, so
// they issue back to back; sor_stage stores +0.0 for those slots.
template <int S, int NIT, int MT, typename TW, typename TV>
__device__ inline void load_slice(const EllT<TW> &A, int slice, int ri, int jo, TV *wq, int *cq)
{
    constexpr int LPR = 64 / S;
    const size_t off = (size_t)slice * S * A.m + ri;
    const TW *wp = A.w + off;
    const int *cp = A.col + off;
#pragma unroll
    for (int t = 0; t < NIT; t++)
        if (t * LPR < MT) {
            const int j = t * LPR + jo < A.m ? t * LPR + jo : A.m - 1;
            wq[t] = wv(wp[(size_t)j * S]);
            cq[t] = cp[(size_t)j * S];
        }
}

// One SOR update of a slice's rows, p, source and tags in LDS, the slice's
// weights and columns already in registers (wq[0] of lane jo == 0 is the
// diagonal).  Same arithmetic as slice_rows in mode M_SOR, in two halves so
// the caller can issue the loads of its next slice in between.
//
// sor_stage writes the products of entries j < m, and +0.0 into the slots
// m <= j < MT, to the wave's LDS region (MT slots of S rows); it returns the
// diagonal.
template <int S, int NIT, int MT, typename TV>
__device__ inline TV sor_stage(int slice, int nrows, int m, const TV *xs, const TV *wq, const int *cq, TV *P)
{
    constexpr int LPR = 64 / S;
    const int lane = threadIdx.x & 63;
    const int ri = lane % S, jo = lane / S;
    const bool inr = slice * S + ri < nrows;
    TV xv[NIT];
#pragma unroll
    for (int t = 0; t < NIT; t++)
        if (t * LPR < MT) xv[t] = xs[cq[t]];
#pragma unroll
    for (int t = 0; t < NIT; t++) {
        const int j = t * LPR + jo;
        // (the slot test is needed only where a slot group straddles MT)
        if (t * LPR < MT && (t * LPR + LPR <= MT || j < MT) && inr) P[j * S + ri] = j < m ? wq[t] * xv[t] : TV(0);
    }
    return wq[0];
}

// sor_finish sums slots 1..MT-1 of each row in order and updates the row.  The
// sum has a fixed trip count, so every LDS read can be issued ahead of the
// adds.  It starts at +0.0, and a round-to-nearest sum that starts at +0.0 is
// never -0.0, so adding the +0.0 slots past m leaves it unchanged: the result
// is bitwise the sum over 1 <= j < m.
template <int S, int MT, int B, typename TV>
__device__ inline void sor_finish(int slice, int nrows, const uint8_t *tg, const TV *sr, TV *xs, TV w, TV c,
                                  TV dg, const TV *P)
{
    const int lane = threadIdx.x & 63;
    const int ri = lane % S, jo = lane / S;
    const int r = slice * S + ri;
    wave_sync();
    if (jo == 0 && r < nrows && tg[r] != ROW_PAD) {
        const TV *q = P + ri;
        TV sum = TV(0);
        // slots in batches of B, each batch's reads issued before the previous
        // batch's adds (the scheduling barriers keep that order), so the LDS
        // latency overlaps the dependent adds
        constexpr int NB = (MT - 1 + B - 1) / B;
        TV v[2][B];
#pragma unroll
        for (int t = 0; t < B; t++)
            if (1 + t < MT) v[0][t] = q[(1 + t) * S];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            if (b + 1 < NB) {
#pragma unroll
                for (int t = 0; t < B; t++)
                    if (1 + (b + 1) * B + t < MT) v[(b + 1) & 1][t] = q[(1 + (b + 1) * B + t) * S];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = 0; t < B; t++)
                if (1 + b * B + t < MT) sum += v[b & 1][t];
            __builtin_amdgcn_sched_barrier(0);
        }
        xs[r] = w * ((sr[r] - sum) / dg) + c * xs[r];
    }
    wave_sync();
}

// Residual form 1 (s - A p, zero on boundary rows) or 2 (s - A p on every
// row) of a small level, p and s in LDS, one row per thread: slice_rows'
// arithmetic (products in TV, subtracted from s left to right), with a row's
// loads and gathers issued eight entries at a time ahead of its subtractions.
template <int S, typename TW, typename TV>
__device__ inline void residual_rows_lds(const EllT<TW> &R, int nrows, int form, const uint8_t *tg,
                                         const TV *sr, const TV *xs, TV *res)
{
    const int m = R.m;
    for (int r = threadIdx.x; r < nrows; r += blockDim.x) {
        const uint8_t tag = tg[r];
        if (tag == ROW_PAD) continue;
        if (form == M_RES1 && tag == ROW_BOUNDARY) {
            res[r] = TV(0);
            continue;
        }
        const size_t base = (size_t)(r / S) * S * m + (r % S);
        const TW *wp = R.w + base;
        const int *cp = R.col + base;
        TV acc = sr[r];
        int j = 0;
        for (; j + 8 <= m; j += 8) {
            TV wq[8], xq[8];
#pragma unroll
            for (int t = 0; t < 8; t++) wq[t] = wv(wp[(size_t)(j + t) * S]);
#pragma unroll
            for (int t = 0; t < 8; t++) xq[t] = xs[cp[(size_t)(j + t) * S]];
#pragma unroll
            for (int t = 0; t < 8; t++) acc = acc - wq[t] * xq[t];
        }
        for (; j < m; j++) acc = acc - wv(wp[(size_t)j * S]) * xs[cp[(size_t)j * S]];
        res[r] = acc;
    }
}

// Whole relaxation of a small level in one workgroup, with the level's p,
// source and tags held in LDS: num_relax sweeps over the colour classes in
// order (workgroup barrier between classes), each sweep followed, for the T
// and F variants, by p -= p[node 0] and the residual form 1 (T) or 2 (F, with
// the lap operator A2).  Weights and columns stream from L2; each wave loads
// its first slice of the next colour step before the barrier, so a step costs
// the LDS gathers and the ordered sums.  Identical arithmetic to the
// per-colour launches.
template <int S, bool PIN, int MT, typename TW, typename TV>
__global__ void __launch_bounds__(SMALL_BLK_MAX)
relax_small_kernel(EllT<TW> A, EllT<TW> A2, const int *__restrict__ col_slices, const int *__restrict__ col_soff,
                   int ncol, int nrows, const uint8_t *__restrict__ btag, const int *__restrict__ ref0,
                   const TV *__restrict__ src, TV *p, TV *res, TV w, TV c, int num_relax, int variant)
{
    // register slots per lane: stencils up to 72 wide (36 at S = 16, the
    // slice height used only for stencils of at most 36 entries)
    constexpr int NIT = S == 16 ? 9 : (72 * S + 63) / 64;
    extern __shared__ unsigned char lds_raw[];
    const int NW = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int ri = lane % S, jo = lane / S;
    const int mw = small_slots(A.m > A2.m ? A.m : A2.m, MT);
    const int nlist = col_soff[ncol];
    TV *xs = reinterpret_cast<TV *>(lds_raw);                         // [nrows] p
    TV *sr = xs + nrows;                                              // [nrows] source
    TV *P = sr + nrows + (size_t)wave * mw * S;                       // [NW][mw][S]
    int *soff = reinterpret_cast<int *>(sr + nrows + (size_t)NW * mw * S);  // [ncol+1]
    int *slist = soff + ncol + 1;                                     // [nlist]
    uint8_t *tg = reinterpret_cast<uint8_t *>(slist + nlist);         // [nrows]
    __shared__ TV pref;
    for (int r = threadIdx.x; r < nrows; r += blockDim.x) {
        xs[r] = p[r];
        sr[r] = src[r];
        tg[r] = btag[r];
    }
    for (int k = threadIdx.x; k <= ncol; k += blockDim.x) soff[k] = col_soff[k];
    for (int k = threadIdx.x; k < nlist; k += blockDim.x) slist[k] = col_slices[k];
    __syncthreads();
    static_assert(MT <= NIT * (64 / S), "a lane's registers hold the MT slots of its rows");
    TV wq[NIT];
    int cq[NIT];
    auto first_slice = [&](int cl) { const int k = soff[cl] + wave; return k < soff[cl + 1] ? slist[k] : -1; };
    if (num_relax > 0) {
        const int f = first_slice(0);
        if (f >= 0) load_slice<S, NIT, MT>(A, f, ri, jo, wq, cq);
    }
    for (int it = 0; it < num_relax; it++) {
        for (int cl = 0; cl < ncol; cl++) {
            const int s0 = soff[cl], s1 = soff[cl + 1];
            // the first slice of the next colour step is loaded ahead too
            // (across a sweep boundary only when no pin/residual pass intervenes)
            const bool ahead = cl + 1 < ncol || (it + 1 < num_relax && !PIN);
            const int fnext = ahead ? first_slice(cl + 1 < ncol ? cl + 1 : 0) : -1;
            for (int k = s0 + wave; k < s1; k += NW) {
                const int sl = slist[k];
                const TV dg = sor_stage<S, NIT, MT>(sl, nrows, A.m, xs, wq, cq, P);
                // this wave's next slice: the registers are free once staged
                const int nx = k + NW < s1 ? slist[k + NW] : fnext;
                if (nx >= 0) load_slice<S, NIT, MT>(A, nx, ri, jo, wq, cq);
                // (batches of 8 in FP64, which would spill with 16)
                sor_finish<S, MT, sizeof(TV) == 8 ? 8 : 16>(sl, nrows, tg, sr, xs, w, c, dg, P);
            }
            if (s0 + wave >= s1 && fnext >= 0) load_slice<S, NIT, MT>(A, fnext, ri, jo, wq, cq);
            __syncthreads();
        }
        if constexpr (!PIN) continue;   // variant H: no pin, no residual pass
        if (threadIdx.x == 0) pref = xs[*ref0];
        __syncthreads();
        const TV pr = pref;
        for (int r = threadIdx.x; r < nrows; r += blockDim.x)
            if (tg[r] != ROW_PAD) xs[r] = xs[r] - pr;
        __syncthreads();
        if (variant == MGM_VAR_T)
            residual_rows_lds<S>(A, nrows, M_RES1, tg, sr, xs, res);
        else
            residual_rows_lds<S>(A2, nrows, M_RES2, tg, sr, xs, res);
        if (it + 1 < num_relax) {
            const int f = first_slice(0);
            if (f >= 0) load_slice<S, NIT, MT>(A, f, ri, jo, wq, cq);
        }
        __syncthreads();
    }
    for (int r = threadIdx.x; r < nrows; r += blockDim.x) p[r] = xs[r];
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

template <typename TV>
__global__ void __launch_bounds__(BLK)
zero_kernel(int N, const uint8_t *__restrict__ btag, int interior_only, TV *p)
{
    int r = blockIdx.x * BLK + threadIdx.x;
    if (r >= N || btag[r] == ROW_PAD) return;
    if (!interior_only || !btag[r]) p[r] = 0;
}

// One weighted-Jacobi sweep from x = +0 on every row: each neighbour product
// is a signed zero, so the ordered sum is +0, and y = w * ((s - 0) / a_rr) +
// c * 0 -- the same operations in the same order as the full pass, hence its
// result bit for bit -- on the rows it combines (all but padding), without
// reading the stencils.
template <typename TW, typename TV>
__global__ void __launch_bounds__(BLK)
jacobi_zero_kernel(EllT<TW> A, int N, const uint8_t *__restrict__ btag, const TV *__restrict__ src, TV *y, TV w,
                   TV c)
{
    const int r = blockIdx.x * BLK + threadIdx.x;
    if (r >= N || (btag && btag[r] == ROW_PAD)) return;
    const TW *wp = A.sorted ? A.w0 : A.w;
    const TV dg = wv(wp[ell_at(A.m, A.S, r, 0)]);
    const TV sum = TV(0), yv = TV(0);
    y[r] = w * ((src[r] - sum) / dg) + c * yv;
}

template <typename TV>
__global__ void copy_scalar_kernel(const int *ref0, const TV *p, double *dst)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = (double)p[*ref0];
}

// p -= p[node 0] (hf/timple_solver.h:248-252, hf/fractional_step_solver.h:180-183)
template <typename TV>
__global__ void __launch_bounds__(BLK)
pin_kernel(int N, const uint8_t *__restrict__ btag, const double *__restrict__ pref, TV *p)
{
    int r = blockIdx.x * BLK + threadIdx.x;
    if (r >= N || btag[r] == ROW_PAD) return;
    p[r] = p[r] - (TV)*pref;
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

// Narrowing copies for the mixed-precision V-cycle.  BF16 rounds to nearest,
// ties to even (weights are finite).
__global__ void __launch_bounds__(BLK) narrow_f32_kernel(long n, const double *__restrict__ in,
                                                          float *__restrict__ out)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t < n) out[t] = (float)in[t];
}
__global__ void __launch_bounds__(BLK) narrow_bf16_kernel(long n, const double *__restrict__ in,
                                                           bf16 *__restrict__ out)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t >= n) return;
    unsigned u = __float_as_uint((float)in[t]);
    u += 0x7fffu + ((u >> 16) & 1u);
    out[t].bits = (uint16_t)(u >> 16);
}
// p += e on every non-padding row
__global__ void __launch_bounds__(BLK) add_correction_kernel(int N, const uint8_t *__restrict__ btag,
                                                              const float *__restrict__ e,
                                                              double *__restrict__ p)
{
    int r = blockIdx.x * BLK + threadIdx.x;
    if (r >= N || btag[r] == ROW_PAD) return;
    p[r] = p[r] + (double)e[r];
}

// Streaming copy (one 16-byte load and store per thread; measured the fastest
// of grid-stride, unrolled and non-temporal forms): the HBM reference the
// stencil kernels are compared with.
__global__ void __launch_bounds__(BLK) stream_copy_kernel(long n16, const uint4 *__restrict__ in,
                                                           uint4 *__restrict__ out)
{
    const long i = (long)blockIdx.x * BLK + threadIdx.x;
    if (i < n16) out[i] = in[i];
}

// Row-major reference-order rows -> sliced-ELL device-order slab.
// One workgroup per slice: bitonic sort of the slice's (column << b | slot)
// keys in LDS (b = slot_bits; padded to 1024 with keys above every real one),
// written back as (slot << (32 - b)) | column.  Requires slab_sortable.
__global__ void __launch_bounds__(GS_BLK)
gather_sort_kernel(int m, int S, int b, const int *__restrict__ col, int *__restrict__ gcol)
{
    __shared__ unsigned key[1024];
    const int n = S * m;
    const size_t sb = (size_t)blockIdx.x * n;
    for (int i = threadIdx.x; i < 1024; i += GS_BLK)
        key[i] = i < n ? ((unsigned)col[sb + i] << b) | (unsigned)i : 0xFFFFFFFFu;
    __syncthreads();
    for (int k = 2; k <= 1024; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < 1024; i += GS_BLK) {
                const int o = i ^ j;
                if (o > i) {
                    const unsigned a = key[i], c = key[o];
                    if ((a > c) == ((i & k) == 0)) {
                        key[i] = c;
                        key[o] = a;
                    }
                }
            }
            __syncthreads();
        }
    const unsigned smask = (1u << b) - 1;
    for (int q = threadIdx.x; q < n; q += GS_BLK) {
        const unsigned v = key[q];
        gcol[sb + q] = (int)(((v & smask) << (32 - b)) | (v >> b));
    }
}

__global__ void __launch_bounds__(BLK)
permute_slab_kernel(long n, int per, int shift, const int *__restrict__ gcol, const double *__restrict__ w,
                    double *__restrict__ wg)
{
    const long e = (long)blockIdx.x * BLK + threadIdx.x;
    if (e >= n) return;
    const long sb = e - e % per;
    wg[e] = w[sb + ((unsigned)gcol[e] >> shift)];
}

__global__ void __launch_bounds__(BLK)
gather_rows_kernel(int N, int m, int S, const double *__restrict__ rows, const int *__restrict__ iperm,
                   const int *__restrict__ cloud, double *__restrict__ w_out,
                   int *__restrict__ col_out)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t >= (long)N * m) return;
    int ref = (int)(t / m), j = (int)(t % m);
    size_t o = ell_at(m, S, iperm[ref], j);
    if (w_out) w_out[o] = rows[t];
    if (col_out) col_out[o] = iperm[cloud[t]];
}

// Transfer rows (reference order, width m of the source level) -> slab over the
// target level's device order, columns pre-composed through the source cloud.
__global__ void __launch_bounds__(BLK)
transfer_slab_kernel(int Nt, int m, int S, const double *__restrict__ rows,
                     const int *__restrict__ iperm_t, const int *__restrict__ pts,
                     const int *__restrict__ cloud_src, const int *__restrict__ iperm_src,
                     double *__restrict__ w_out, int *__restrict__ col_out)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t >= (long)Nt * m) return;
    int ref = (int)(t / m), j = (int)(t % m);
    size_t o = ell_at(m, S, iperm_t[ref], j);
    w_out[o] = rows[t];
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

// Rows with wide boundary stencils (reference order: node bnodes[t], its mb
// points cloud_b[t * mb ..], weights w[t * mb ..]) applied to a device-order
// vector: the ordered sum from 0.0 of the rounded products, one thread per row.
__global__ void __launch_bounds__(BLK)
wide_rows_kernel(int nbw, int mb, const int *__restrict__ bnodes, const int *__restrict__ cloud_b,
                 const double *__restrict__ w, const int *__restrict__ iperm, const double *__restrict__ x,
                 double *__restrict__ y)
{
    const int t = blockIdx.x * BLK + threadIdx.x;
    if (t >= nbw) return;
    const int *c = cloud_b + (size_t)t * mb;
    const double *ww = w + (size_t)t * mb;
    double s = 0.0;
    for (int j = 0; j < mb; j++) s += ww[j] * x[iperm[c[j]]];
    y[iperm[bnodes[t]]] = s;
}

__global__ void __launch_bounds__(BLK)
dirichlet_rows_kernel(int N, int m, const uint8_t *__restrict__ btag, double *__restrict__ rows)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t >= (long)N * m) return;
    int i = (int)(t / m), j = (int)(t % m);
    if (btag[i]) rows[t] = (j == 0) ? 1.0 : 0.0;
}

// W waves (slices) per workgroup
template <int W, typename TW, typename TV>
int launch_rows_w(EllT<TW> A, const int *slices, int first, int nslices, int nrows, int mode,
                  const uint8_t *btag, const TV *src, const TV *x, TV *y, TV w, TV c, int keep, hipStream_t s)
{
    const int blocks = (nslices + W - 1) / W;
    const size_t shm = (size_t)W * prod_stride(A.m) * A.S * sizeof(TV);
    const dim3 grid(blocks), block(64 * W);
    if (A.S == 4)
        hipLaunchKernelGGL((rows_kernel<4, TW, TV, W>), grid, block, shm, s, A, slices, first, nslices, nrows,
                           mode, btag, src, x, y, w, c, keep, blocks);
    else if (A.S == 8)
        hipLaunchKernelGGL((rows_kernel<8, TW, TV, W>), grid, block, shm, s, A, slices, first, nslices, nrows,
                           mode, btag, src, x, y, w, c, keep, blocks);
    else if (A.S == 16)
        hipLaunchKernelGGL((rows_kernel<16, TW, TV, W>), grid, block, shm, s, A, slices, first, nslices, nrows,
                           mode, btag, src, x, y, w, c, keep, blocks);
    else
        MGM_FAIL(-1, "unsupported slice height");
    HIP_TRY(hipGetLastError());
    return 0;
}

// slices == nullptr: all slices covering rows [0, nrows).  SOR colour steps
// run one wave per workgroup, the full passes four: the colour steps are
// latency-bound and gain from the finer spread (1M-point solve 0.313 s at
// four waves, 0.304 s at two, 0.299 s at one), while the full-pass apply
// loses 2.5% with smaller workgroups.
template <typename TW, typename TV>
size_t window_lds(const EllT<TW> &A)
{
    return ((size_t)kWindowWaves * prod_stride(A.m) * A.S + (size_t)A.win_max) * sizeof(TV);
}
template <typename TW, typename TV>
int launch_window(EllT<TW> A, int mode, const uint8_t *btag, const TV *src, const TV *x, TV *y, TV w, TV c,
                  hipStream_t s, TV *y2 = nullptr)
{
    const dim3 grid(A.win_blocks), block(64 * kWindowWaves);
    const size_t shm = window_lds<TW, TV>(A);
    if (A.S == 4)
        hipLaunchKernelGGL((window_rows_kernel<4, TW, TV>), grid, block, shm, s, A, mode, btag, src, x, y, y2, w, c);
    else if (A.S == 8)
        hipLaunchKernelGGL((window_rows_kernel<8, TW, TV>), grid, block, shm, s, A, mode, btag, src, x, y, y2, w, c);
    else if (A.S == 16)
        hipLaunchKernelGGL((window_rows_kernel<16, TW, TV>), grid, block, shm, s, A, mode, btag, src, x, y, y2, w, c);
    else
        MGM_FAIL(-1, "unsupported slice height");
    HIP_TRY(hipGetLastError());
    return 0;
}

template <typename TW, typename TV>
int launch_rows(EllT<TW> A, const int *slices, int first, int nslices, int nrows, int mode,
                const uint8_t *btag, const TV *src, const TV *x, TV *y, TV w, TV c, int keep,
                hipStream_t s)
{
    // full SpMV and residual passes over a level with LDS windows
    if (A.win_blocks && btag && !slices && first == 0 && nslices < 0 &&
        (mode == M_SPMV || mode == M_RES0 || mode == M_RES1 || mode == M_RES2 || mode == M_RES3 ||
         mode == M_JACOBI || mode == M_PROLONG || mode == M_RESTRICT) &&
        window_lds<TW, TV>(A) <= kSmallLdsMax)
        return launch_window(A, mode, btag, src, x, y, mode == M_RESTRICT ? TV(keep ? 1 : 0) : w, c, s);
    if (nslices < 0) nslices = (nrows + A.S - 1) / A.S;  // full pass
    if (nslices <= 0 || nrows <= 0) return 0;
    if (mode == M_SOR)
        return launch_rows_w<1>(A, slices, first, nslices, nrows, mode, btag, src, x, y, w, c, keep, s);
    return launch_rows_w<WAVES>(A, slices, first, nslices, nrows, mode, btag, src, x, y, w, c, keep, s);
}

}  // namespace

#define LAUNCH(kern, n, ...)                                                         \
    do {                                                                             \
        if ((n) > 0) {                                                               \
            hipLaunchKernelGGL(kern, dim3(nblk(n)), dim3(BLK), 0, s, __VA_ARGS__);   \
            HIP_TRY(hipGetLastError());                                              \
        }                                                                            \
    } while (0)

int launch_gather_sort(int nslices, int m, int S, const int *col, int *gcol, hipStream_t s)
{
    if (!slab_sortable(S, m, 0)) MGM_FAIL(-1, "gather_sort: slice of more than 1024 entries");
    if (nslices <= 0) return 0;
    hipLaunchKernelGGL(gather_sort_kernel, dim3(nslices), dim3(GS_BLK), 0, s, m, S, slot_bits(S, m), col, gcol);
    HIP_TRY(hipGetLastError());
    return 0;
}
int launch_permute_slab(long n, int m, int S, const int *gcol, const double *w, double *wg, hipStream_t s)
{
    LAUNCH(permute_slab_kernel, n, n, S * m, 32 - slot_bits(S, m), gcol, w, wg);
    return 0;
}

int slice_height(int m)
{
    static const int forced = [] {
        const char *e = getenv("MGM_SLICE_HEIGHT");
        int v = e ? atoi(e) : 0;
        return (v == 4 || v == 8 || v == 16) ? v : 0;
    }();
    if (forced) return forced;
    return m > 36 ? 8 : 16;
}

template <typename TW, typename TV>
int launch_spmv(EllT<TW> A, int nrows, const uint8_t *btag, const TV *x, TV *y, hipStream_t s)
{
    return launch_rows<TW, TV>(A, nullptr, 0, -1, nrows, M_SPMV, btag, nullptr, x, y, TV(0), TV(0), 0, s);
}
// One weighted-Jacobi sweep over every row: y = the SOR update of x with
// every neighbour taken from x (x and y distinct).
template <typename TW, typename TV>
int launch_jacobi(EllT<TW> A, int N, const uint8_t *btag, const TV *src, const TV *x, TV *y, TV w, TV c,
                  hipStream_t s)
{
    return launch_rows<TW, TV>(A, nullptr, 0, -1, N, M_JACOBI, btag, src, x, y, w, c, 0, s);
}
template <typename TW, typename TV>
int launch_jacobi_zero(EllT<TW> A, int N, const uint8_t *btag, const TV *src, TV *y, TV w, TV c, hipStream_t s)
{
    LAUNCH((jacobi_zero_kernel<TW, TV>), N, A, N, btag, src, y, w, c);
    return 0;
}
template <typename TW, typename TV>
int launch_residual_jacobi(EllT<TW> A, const uint8_t *btag, const TV *src, const TV *x, TV *y, TV *res, TV w, TV c,
                           hipStream_t s)
{
    if (!A.win_blocks || !btag || window_lds<TW, TV>(A) > kSmallLdsMax)
        MGM_FAIL(-1, "launch_residual_jacobi: the level has no LDS windows");
    return launch_window(A, M_RESJAC, btag, src, x, y, w, c, s, res);
}
bool has_windows(const Ell &A) { return A.win_blocks && window_lds<double, double>(A) <= kSmallLdsMax; }
template <typename TW, typename TV>
int launch_sor_slices(EllT<TW> A, const int *slices, int first, int nslices, const uint8_t *btag,
                      const TV *src, TV *p, TV w, TV c, hipStream_t s)
{
    if (nslices <= 0) return 0;
    return launch_rows<TW, TV>(A, slices, first, nslices, 0x7fffffff, M_SOR, btag, src, p, p, w, c, 0, s);
}
// LDS bytes of relax_small_kernel with nw waves: p, source, staging, the
// colour tables (ncol + 1 + nrows/S ints, bounded by 2*nrows + 1) and the
// tags; 0 if it does not fit
// Stencil width the sums of relax_small_kernel are unrolled for: 24, 40 or 72
// (at most 36 at S = 16, where a lane holds 9 entries of 4 per row).
static int small_mt(int m, int S)
{
    const int mt = m <= 24 ? 24 : m <= 40 ? 40 : 72;
    return S == 16 && mt > 36 ? 36 : mt;
}
static size_t relax_small_lds(int nrows, int mw, int S, int nw, int vbytes)
{
    size_t b = (2 * (size_t)nrows + (size_t)nw * mw * S) * (size_t)vbytes +
               (2 * (size_t)nrows + 1) * sizeof(int) + (size_t)nrows;
    return b <= kSmallLdsMax ? b : 0;
}
static int relax_small_waves(int nrows, int mw, int S, int vbytes)
{
    for (int nw = SMALL_BLK_MAX / 64; nw >= 4; nw /= 2)
        if (relax_small_lds(nrows, mw, S, nw, vbytes)) return nw;
    return 0;
}
bool relax_small_fits(int nrows, int m, int m2, int S, int vbytes)
{
    if (m > small_mt(m, S)) return false;
    return relax_small_waves(nrows, small_slots(m > m2 ? m : m2, small_mt(m, S)), S, vbytes) > 0;
}
template <int S, bool PIN, int MT, typename TW, typename TV>
static int launch_relax_small_v(EllT<TW> A, EllT<TW> A2, const int *col_slices, const int *col_soff, int ncol,
                                int nrows, const uint8_t *btag, const int *ref0, const TV *src, TV *p, TV *res,
                                TV w, TV c, int num_relax, int variant, int nw, size_t shm, hipStream_t s)
{
    hipLaunchKernelGGL((relax_small_kernel<S, PIN, MT, TW, TV>), dim3(1), dim3(64 * nw), shm, s, A, A2, col_slices,
                       col_soff, ncol, nrows, btag, ref0, src, p, res, w, c, num_relax, variant);
    HIP_TRY(hipGetLastError());
    return 0;
}
// The H variant has no pin or residual pass: its own instance, without their
// registers.
// (BF16 weights serve the H variant only.)
template <int S, int MT, typename TW, typename TV>
static int launch_relax_small_m(EllT<TW> A, EllT<TW> A2, const int *col_slices, const int *col_soff, int ncol,
                                int nrows, const uint8_t *btag, const int *ref0, const TV *src, TV *p, TV *res,
                                TV w, TV c, int num_relax, int variant, int nw, size_t shm, hipStream_t s)
{
    if (variant == MGM_VAR_H)
        return launch_relax_small_v<S, false, MT>(A, A2, col_slices, col_soff, ncol, nrows, btag, ref0, src, p,
                                                  res, w, c, num_relax, variant, nw, shm, s);
    if constexpr (std::is_same<TW, bf16>::value) {
        MGM_FAIL(-1, "relax_small: BF16 weights only for the H variant");
    } else {
        return launch_relax_small_v<S, true, MT>(A, A2, col_slices, col_soff, ncol, nrows, btag, ref0, src, p,
                                                 res, w, c, num_relax, variant, nw, shm, s);
    }
}
template <int S, typename TW, typename TV>
static int launch_relax_small_t(EllT<TW> A, EllT<TW> A2, const int *col_slices, const int *col_soff, int ncol,
                                int nrows, const uint8_t *btag, const int *ref0, const TV *src, TV *p, TV *res,
                                TV w, TV c, int num_relax, int variant, int nw, size_t shm, hipStream_t s)
{
    switch (small_mt(A.m, S)) {
    case 24: return launch_relax_small_m<S, 24>(A, A2, col_slices, col_soff, ncol, nrows, btag, ref0, src, p,
                                                res, w, c, num_relax, variant, nw, shm, s);
    case 40: return launch_relax_small_m<S, S == 16 ? 36 : 40>(A, A2, col_slices, col_soff, ncol, nrows, btag, ref0, src, p,
                                                res, w, c, num_relax, variant, nw, shm, s);
    }
    constexpr int MTMAX = S == 16 ? 36 : 72;
    return launch_relax_small_m<S, MTMAX>(A, A2, col_slices, col_soff, ncol, nrows, btag, ref0, src, p, res,
                                          w, c, num_relax, variant, nw, shm, s);
}
template <typename TW, typename TV>
int launch_relax_small(EllT<TW> A, EllT<TW> A2, const int *col_slices, const int *col_soff, int ncol, int nrows,
                       const uint8_t *btag, const int *ref0, const TV *src, TV *p, TV *res, TV w, TV c,
                       int num_relax, int variant, hipStream_t s)
{
    if (A.S != A2.S) MGM_FAIL(-1, "relax_small: operators differ in slice height");
    if (A.sorted || A2.sorted) MGM_FAIL(-1, "relax_small: needs plain slabs");
    if (A.m > small_mt(A.m, A.S)) MGM_FAIL(-1, "relax_small: stencil too wide");
    const int mw = small_slots(A.m > A2.m ? A.m : A2.m, small_mt(A.m, A.S));
    const int nw = relax_small_waves(nrows, mw, A.S, sizeof(TV));
    if (!nw) MGM_FAIL(-1, "relax_small: level does not fit in LDS");
    const size_t shm = relax_small_lds(nrows, mw, A.S, nw, sizeof(TV));
    switch (A.S) {
    case 4: return launch_relax_small_t<4>(A, A2, col_slices, col_soff, ncol, nrows, btag, ref0, src, p, res,
                                           w, c, num_relax, variant, nw, shm, s);
    case 8: return launch_relax_small_t<8>(A, A2, col_slices, col_soff, ncol, nrows, btag, ref0, src, p, res,
                                           w, c, num_relax, variant, nw, shm, s);
    case 16: return launch_relax_small_t<16>(A, A2, col_slices, col_soff, ncol, nrows, btag, ref0, src, p,
                                             res, w, c, num_relax, variant, nw, shm, s);
    }
    MGM_FAIL(-1, "unsupported slice height");
}
#define RELAX_SMALL_INST(TW, TV)                                                                          \
    template int launch_relax_small<TW, TV>(EllT<TW>, EllT<TW>, const int *, const int *, int, int,           \
                                            const uint8_t *, const int *, const TV *, TV *, TV *, TV, TV, int, \
                                            int, hipStream_t);
RELAX_SMALL_INST(double, double)
RELAX_SMALL_INST(float, float)
RELAX_SMALL_INST(bf16, float)
#undef RELAX_SMALL_INST
// Dynamic-LDS limit of the single-workgroup relax (up to 159 KB of the 160 KB
// a workgroup may have), set for every instance once per device by mgm_init:
// outside any stream capture, so a captured relax launches nothing else.
template <typename TW, typename TV>
static int small_lds_attrs()
{
    const int mx = (int)kSmallLdsMax;
#define SMALL_ATTR(S, PIN, MT)                                                         \
    HIP_TRY(hipFuncSetAttribute((const void *)relax_small_kernel<S, PIN, MT, TW, TV>, \
                                hipFuncAttributeMaxDynamicSharedMemorySize, mx))
    SMALL_ATTR(4, false, 24); SMALL_ATTR(4, false, 40); SMALL_ATTR(4, false, 72);
    SMALL_ATTR(8, false, 24); SMALL_ATTR(8, false, 40); SMALL_ATTR(8, false, 72);
    SMALL_ATTR(16, false, 24); SMALL_ATTR(16, false, 36);
    if constexpr (!std::is_same<TW, bf16>::value) {
        SMALL_ATTR(4, true, 24); SMALL_ATTR(4, true, 40); SMALL_ATTR(4, true, 72);
        SMALL_ATTR(8, true, 24); SMALL_ATTR(8, true, 40); SMALL_ATTR(8, true, 72);
        SMALL_ATTR(16, true, 24); SMALL_ATTR(16, true, 36);
    }
#undef SMALL_ATTR
    return 0;
}
template <typename TW, typename TV>
static int window_lds_attrs()
{
    for (const void *f : {(const void *)window_rows_kernel<4, TW, TV>, (const void *)window_rows_kernel<8, TW, TV>,
                          (const void *)window_rows_kernel<16, TW, TV>})
        HIP_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSmallLdsMax));
    return 0;
}
int stencil_device_init()
{
    int rc;
    if ((rc = small_lds_attrs<double, double>())) return rc;
    if ((rc = small_lds_attrs<float, float>())) return rc;
    if ((rc = small_lds_attrs<bf16, float>())) return rc;
    if ((rc = window_lds_attrs<double, double>())) return rc;
    if ((rc = window_lds_attrs<float, float>())) return rc;
    return window_lds_attrs<bf16, float>();
}
int launch_sor_natural(Ell A, int N, const int *iperm, const double *src, double *p, double w,
                       double c, hipStream_t s)
{
    if (A.sorted) MGM_FAIL(-1, "sor_natural: needs the plain slab");
    hipLaunchKernelGGL(sor_natural_kernel, dim3(1), dim3(64), 0, s, A, N, iperm, src, p, w, c);
    HIP_TRY(hipGetLastError());
    return 0;
}
template <typename TW, typename TV>
int launch_residual(EllT<TW> A, int N, int form, const uint8_t *btag, const TV *src, const TV *p,
                    TV *res, hipStream_t s)
{
    static const int modes[4] = {M_RES0, M_RES1, M_RES2, M_RES3};
    if (form < 0 || form > 3) MGM_FAIL(-1, "bad residual form");
    return launch_rows<TW, TV>(A, nullptr, 0, -1, N, modes[form], btag, src, p, res, TV(0), TV(0), 0, s);
}
template <typename TW, typename TV>
int launch_restrict(EllT<TW> R, int N, const uint8_t *btag, int keep_boundary, const TV *in, TV *out,
                    hipStream_t s)
{
    return launch_rows<TW, TV>(R, nullptr, 0, -1, N, M_RESTRICT, btag, nullptr, in, out, TV(0), TV(0),
                               keep_boundary, s);
}
template <typename TW, typename TV>
int launch_prolong(EllT<TW> P, int N, const uint8_t *btag, const TV *pc, TV *p, hipStream_t s)
{
    return launch_rows<TW, TV>(P, nullptr, 0, -1, N, M_PROLONG, btag, nullptr, pc, p, TV(0), TV(0), 0, s);
}
template <typename TV>
int launch_zero(int N, const uint8_t *btag, int interior_only, TV *p, hipStream_t s)
{
    LAUNCH(zero_kernel<TV>, N, N, btag, interior_only, p);
    return 0;
}

#define MGM_INSTANTIATE(TW, TV)                                                                      \
    template int launch_spmv<TW, TV>(EllT<TW>, int, const uint8_t *, const TV *, TV *, hipStream_t); \
    template int launch_sor_slices<TW, TV>(EllT<TW>, const int *, int, int, const uint8_t *,         \
                                           const TV *, TV *, TV, TV, hipStream_t);                   \
    template int launch_jacobi<TW, TV>(EllT<TW>, int, const uint8_t *, const TV *, const TV *, TV *, \
                                       TV, TV, hipStream_t);                                         \
    template int launch_jacobi_zero<TW, TV>(EllT<TW>, int, const uint8_t *, const TV *, TV *, TV, TV, \
                                            hipStream_t);                                            \
    template int launch_residual<TW, TV>(EllT<TW>, int, int, const uint8_t *, const TV *,            \
                                         const TV *, TV *, hipStream_t);                             \
    template int launch_restrict<TW, TV>(EllT<TW>, int, const uint8_t *, int, const TV *, TV *,      \
                                         hipStream_t);                                               \
    template int launch_prolong<TW, TV>(EllT<TW>, int, const uint8_t *, const TV *, TV *, hipStream_t);
MGM_INSTANTIATE(double, double)
template int launch_residual_jacobi<double, double>(Ell, const uint8_t *, const double *, const double *, double *,
                                                    double *, double, double, hipStream_t);
MGM_INSTANTIATE(float, float)
MGM_INSTANTIATE(bf16, float)
#undef MGM_INSTANTIATE
template int launch_zero<double>(int, const uint8_t *, int, double *, hipStream_t);
template int launch_zero<float>(int, const uint8_t *, int, float *, hipStream_t);
template <typename TV>
int launch_pin(int N, const uint8_t *btag, const int *ref0_dev, TV *p, double *scratch, hipStream_t s)
{
    hipLaunchKernelGGL(copy_scalar_kernel<TV>, dim3(1), dim3(64), 0, s, ref0_dev, (const TV *)p, scratch);
    HIP_TRY(hipGetLastError());
    LAUNCH(pin_kernel<TV>, N, N, btag, scratch, p);
    return 0;
}
template int launch_pin<double>(int, const uint8_t *, const int *, double *, double *, hipStream_t);
template int launch_pin<float>(int, const uint8_t *, const int *, float *, double *, hipStream_t);
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
int launch_stream_copy(long n16, const void *in, void *out, hipStream_t s)
{
    hipLaunchKernelGGL(stream_copy_kernel, dim3((unsigned)((n16 + BLK - 1) / BLK)), dim3(BLK), 0, s, n16,
                       (const uint4 *)in, (uint4 *)out);
    HIP_TRY(hipGetLastError());
    return 0;
}
int launch_narrow_f32(long n, const double *in, float *out, hipStream_t s)
{
    LAUNCH(narrow_f32_kernel, n, n, in, out);
    return 0;
}
int launch_narrow_bf16(long n, const double *in, bf16 *out, hipStream_t s)
{
    LAUNCH(narrow_bf16_kernel, n, n, in, out);
    return 0;
}
int launch_add_correction(int N, const uint8_t *btag, const float *e, double *p, hipStream_t s)
{
    LAUNCH(add_correction_kernel, N, N, btag, e, p);
    return 0;
}
int launch_gather_rows(int N, int m, int S, const double *rows, const int *iperm,
                       const int *cloud, double *w_out, int *col_out, hipStream_t s)
{
    LAUNCH(gather_rows_kernel, (long)N * m, N, m, S, rows, iperm, cloud, w_out, col_out);
    return 0;
}
int launch_transfer_slab(int Nt, int m, int S, const double *rows, const int *iperm_t,
                         const int *pts, const int *cloud_src, const int *iperm_src,
                         double *w_out, int *col_out, hipStream_t s)
{
    LAUNCH(transfer_slab_kernel, (long)Nt * m, Nt, m, S, rows, iperm_t, pts, cloud_src, iperm_src,
           w_out, col_out);
    return 0;
}
int launch_lap_poisson(int N, int m, int dim, const double *Dx, const double *Dy, const double *Dz,
                       const double *lap, const double *nx, const double *ny, const double *nz,
                       const uint8_t *btag, double *lapP, hipStream_t s)
{
    LAUNCH(lap_poisson_kernel, (long)N * m, N, m, dim, Dx, Dy, Dz, lap, nx, ny, nz, btag, lapP);
    return 0;
}
int launch_wide_rows(int nbw, int mb, const int *bnodes, const int *cloud_b, const double *w, const int *iperm,
                     const double *x, double *y, hipStream_t s)
{
    if (nbw <= 0) return 0;
    hipLaunchKernelGGL(wide_rows_kernel, dim3(nblk(nbw)), dim3(BLK), 0, s, nbw, mb, bnodes, cloud_b, w, iperm, x, y);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_dirichlet_rows(int N, int m, const uint8_t *btag, double *rows, hipStream_t s)
{
    LAUNCH(dirichlet_rows_kernel, (long)N * m, N, m, btag, rows);
    return 0;
}

}  // namespace mgm
