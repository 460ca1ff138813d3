// device.hpp -- HIP helpers and kernel launcher declarations for libmgm.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "mgm_internal.h"

namespace mgm {

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess) {                                                           \
            ::mgm::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));          \
            return -10;                                                                   \
        }                                                                                 \
    } while (0)

// Owning device buffer.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count)
    {
        if (count == n && p) return hipSuccess;
        release();
        if (count == 0) return hipSuccess;
        hipError_t e = hipMalloc((void **)&p, sizeof(T) * count);
        if (e == hipSuccess) n = count;
        return e;
    }
    hipError_t upload(const T *h, size_t count, hipStream_t s)
    {
        hipError_t e = alloc(count);
        if (e != hipSuccess || count == 0) return e;
        return hipMemcpyAsync(p, h, sizeof(T) * count, hipMemcpyHostToDevice, s);
    }
    hipError_t upload(const std::vector<T> &v, hipStream_t s) { return upload(v.data(), v.size(), s); }
    hipError_t download(T *h, size_t count, hipStream_t s) const
    {
        if (count == 0) return hipSuccess;
        return hipMemcpyAsync(h, p, sizeof(T) * count, hipMemcpyDeviceToHost, s);
    }
};

hipStream_t stream();

// ---- kNN (knn.hip) --------------------------------------------------------
// Candidate set binned into a uniform grid (built on the host).
struct KnnGrid {
    double x0, y0, z0, inv_c, c;
    int nx, ny, nz;
    const int *cell_start;   // ncells+1
    const int *cell_pts;     // candidate ids grouped by cell, ascending id in a cell
    const double *px, *py, *pz;  // coordinates indexed by candidate id
};
// Optional insertion-ordered kd-tree used to order equal distances exactly
// as the reference's range query reports them.
struct KdOrder {
    const int *node_of;      // candidate id -> tree node (-1 if not in tree)
    const int *parent;       // tree node -> parent (-1 root)
    const int *depth;
    const uint8_t *side;     // 0 left child, 1 right child
    const double *pos;       // 3 per tree node
};
// For each query q in [0,nq): the K nearest candidates by sqrt distance.
// out[q*K_out + j], j < K_out, receives ids (K_out <= K).  When `query_ids`
// is non-null only those queries run.  tie_flag[q] is set when equal
// distances occur among the K+1 best (order then needs the kd rule).
// Per-device kNN kernel attribute (dynamic LDS up to the 128-point stencils), set once by mgm_init.
int knn_device_init();
int launch_knn(const KnnGrid &g, const KdOrder *kd, int nq, const double *qx, const double *qy,
               const double *qz, const int *query_ids, int K, int K_out, int slot_offset,
               int *out, int out_stride, uint8_t *tie_flag, hipStream_t s);

// ---- weights (weights.hip) -------------------------------------------------
struct WeightsArgs {
    int nsys;                // systems in this launch
    const int *sys;          // node index of each system (reference order)
    int m, n, dim, phs;
    const short *px, *py, *pz;  // device, n entries
    const double *x, *y, *z;    // level coordinates, reference order
    const int *cloud;           // N*m row-major, reference order
    double *Dx, *Dy, *Dz, *lap; // N*m outputs (nullable)
    // interpolation targets evaluated with each system's factorisation
    const int *rt_off;  const double *rt_xyz; const int *rt_row; double *R_out;  // restriction
    const int *pt_off;  const double *pt_xyz; const int *pt_row; double *P_out;  // prolongation
    double *ws;         // workspace: mpn*mpn*batch
    double *rhs;        // 4*mpn*batch
    int *piv;           // mpn*batch
    int batch;
    int *status;        // nonzero if a zero pivot was met
    int rhs_max;        // most right-hand sides of one system (operators + targets)
};
// Systems with mpn + 4 <= 144 run one workgroup per system in registers
// (weights_tile(mpn) > 0) and need no workspace; wider ones use ws/rhs/piv.
int weights_tile(int mpn);
int launch_weights(const WeightsArgs &a, hipStream_t s);
// A batch of systems from several levels (uneven sizes allowed) for one
// persistent launch: lv[l] describes level l (its nsys / sys fields unused),
// list[t] = (level, node) of system t.  Workgroups take systems from a
// work counter until the list is exhausted.
struct WeightsBatch {
    const WeightsArgs *lv;   // device array
    const int2 *list;        // device array of nsys entries
    int nsys;
    int *next;               // work counter, zero at launch
    int *status;             // nonzero if a zero pivot was met
};
// One launch over every system of the given levels (host descriptors; their
// device arrays must stay valid until the stream reaches the end of the
// launch).  Systems larger than 128 unknowns are not accepted here.
int launch_weights_levels(const std::vector<WeightsArgs> &levels, hipStream_t s);
// Tile grid and register budget the autotuner chose for systems of size mpn
// (-1 before the first build of that size).
int weights_kernel_choice(int mpn, int *tile, int *waves_per_simd);
// The matrix-core form of the same solves (lu32.hip): FP32 LU on MFMA, FP64
// refinement; systems up to 128 unknowns (lu32_tiles(mpn) > 0).
int lu32_tiles(int mpn);
int launch_weights_lu32(const WeightsBatch &b, int tiles, int rmax, hipStream_t s);

// ---- stencil kernels (stencil.hip) ----------------------------------------
// Fixed-width rows in sliced-ELL layout: rows are grouped in slices of S
// consecutive rows (S = 8 or 16), a slice stores its m columns one after the
// other, each column holding the slice's S entries contiguously.  Entry j of
// row r lives at ell_at(m, S, r, j).  A wave covers one slice with 64/S lanes
// per row, so every load instruction of the wave reads 64 consecutive values.
// TW is the stored weight type: double, float, or bf16 (the upper 16 bits of
// an IEEE float, widened exactly to float when used).
struct bf16 {
    uint16_t bits;
};
template <typename TW>
struct EllT {
    const int *col;
    const TW *w;
    int m, S;
    bool stream = false;   // slab larger than the last-level cache: load it non-temporally
    // Gather-sorted form: inside each slice the S*m entries are stored in
    // ascending column order, so the 64 lanes of one gather read nearby x
    // values.  col then packs (slot << shift) | column, slot = the entry's
    // offset j*S + ri in the plain slice (slot_bits(S, m) bits at the top,
    // shift = 32 - those), and w holds the weights in the same order; w0 is
    // the plain slab (entry 0 of row ri at the plain offset, the SOR
    // diagonal).  Products land in the same LDS slots as in the plain form, so
    // the row sums are unchanged.
    const TW *w0 = nullptr;
    bool sorted = false;
    int shift = 0;
    // LDS-window form of the full passes (levels in Morton blocks): block b
    // covers device rows [win_row0[b], win_row0[b+1]); its stencils read those
    // rows and the halo rows win_halo[win_hoff[b] .. win_hoff[b+1]) (sorted).
    // win_col is the plain layout's column slab with every column replaced by
    // its place in the block's window: own rows first (column - row0), then
    // the halo in list order, as 16-bit places (a window fits in LDS, so it
    // holds fewer than 65,536 values).  The weights are the plain slab (w0
    // when sorted, else w).  win_max bounds a window's values.
    const uint16_t *win_col = nullptr;
    const int *win_row0 = nullptr, *win_hoff = nullptr, *win_halo = nullptr;
    int win_blocks = 0, win_max = 0;
    // 0: the window holds only the listed rows (win_halo), no own rows -- a
    // transfer operator whose columns index another level
    int win_self = 1;
};
// Waves per workgroup of the LDS-window full passes (one workgroup per block).
constexpr int kWindowWaves = 16;
// Bits of a slot index of a slice of S*m entries (S*m <= 1024).
inline int slot_bits(int S, int m)
{
    int b = 0;
    while ((1 << b) < S * m) b++;
    return b;
}
// A slab of slice height S and width m whose columns index a vector of ncols
// entries can be stored gather-sorted: its slots and columns fit 32 bits
// (the largest column stays below the all-ones key that pads the sort).
inline bool slab_sortable(int S, int m, long ncols)
{
    return S * m <= 1024 && ncols < (1L << (32 - slot_bits(S, m))) - 1;
}
// Slabs above this size cannot stay in the 256 MB last-level cache between passes.
constexpr size_t kStreamSlabBytes = 256ull << 20;
inline bool stream_slab(size_t rows, int m, size_t wbytes) { return rows * m * (wbytes + 4) > kStreamSlabBytes; }
using Ell = EllT<double>;
// Row stride of the products the slice kernels stage in LDS: a slice's row ri
// holds its m products at [ri * prod_stride(m) + j], padded so a row starts on
// 16 bytes and its sum reads four FP32 or two FP64 products per load.
__host__ __device__ constexpr int prod_stride(int m) { return (m + 3) & ~3; }
__host__ __device__ inline size_t ell_at(int m, int S, int r, int j)
{
    return (size_t)(r / S) * (size_t)S * (size_t)m + (size_t)j * S + (size_t)(r % S);
}
// Rows per slice for stencils of width m: 8 above 36 columns, else 16.  The
// environment variable MGM_SLICE_HEIGHT (4, 8 or 16) overrides it.
int slice_height(int m);
inline size_t ell_size(int N, int m, int S) { return (size_t)((N + S - 1) / S) * S * (size_t)m; }
// Device row tags (dbtag): 0 interior, 1 boundary, 2 padding row.  Row counts
// passed to the full-pass launchers are device row counts (padding included).
enum { ROW_INTERIOR = 0, ROW_BOUNDARY = 1, ROW_PAD = 2 };
// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Renumber them so XCD i takes a contiguous range of blocks of rows: neighbouring
// rows (and the x entries their stencils gather) then stay in one L2.
__device__ inline long xcd_block(long b, long nb)
{
    const long xcd = b % 8, q = nb / 8, rem = nb % 8, idx = b / 8;
    return (xcd < rem) ? xcd * (q + 1) + idx : rem * (q + 1) + (xcd - rem) * q + idx;
}

// The stencil launchers are instantiated for (TW, TV) = (double, double),
// (float, float) and (bf16, float); TV is the vector/accumulation type.
// y = A x over rows [0, nrows); rows tagged ROW_PAD in btag (nullable) are skipped.
template <typename TW, typename TV>
int launch_spmv(EllT<TW> A, int nrows, const uint8_t *btag, const TV *x, TV *y, hipStream_t s);
// One SOR update of the rows of nslices slices (one colour class): the slices
// listed in `slices`, or first, first+1, ... when slices is null.
template <typename TW, typename TV>
int launch_sor_slices(EllT<TW> A, const int *slices, int first, int nslices, const uint8_t *btag,
                      const TV *src, TV *p, TV w, TV c, hipStream_t s);
// One weighted-Jacobi sweep over all N rows: y = the SOR update of x with
// every neighbour from x (x and y distinct); LDS windows on blocked levels.
template <typename TW, typename TV>
int launch_jacobi(EllT<TW> A, int N, const uint8_t *btag, const TV *src, const TV *x, TV *y, TV w, TV c,
                  hipStream_t s);
// launch_jacobi for x = +0 on every row that is not padding (bitwise its
// result): y = w * ((s - 0) / a_rr) + c * 0, no stencil read.
template <typename TW, typename TV>
int launch_jacobi_zero(EllT<TW> A, int N, const uint8_t *btag, const TV *src, TV *y, TV w, TV c, hipStream_t s);
// One pass over a level with LDS windows (has_windows) forming, from the same
// products, the residual form 0 of x into res (bitwise launch_residual's) and
// the Jacobi sweep of x into y (bitwise launch_jacobi's).  FP64.
template <typename TW, typename TV>
int launch_residual_jacobi(EllT<TW> A, const uint8_t *btag, const TV *src, const TV *x, TV *y, TV *res, TV w, TV c,
                           hipStream_t s);
bool has_windows(const EllT<double> &A);
// The whole relaxation of a small level (all sweeps, all colour classes, and
// for the T / F variants the pin and residual after each sweep, A2 being the
// operator of the F residual) in one single-workgroup launch.  Weights TW
// (double, float, or bf16 for the H variant), vectors and sums TV.
template <typename TW, typename TV>
int launch_relax_small(EllT<TW> A, EllT<TW> A2, const int *col_slices, const int *col_soff, int ncol, int nrows,
                       const uint8_t *btag, const int *ref0, const TV *src, TV *p, TV *res, TV w, TV c,
                       int num_relax, int variant, hipStream_t s);
// Whether a level of nrows device rows and stencil widths m, m2 fits the
// single-workgroup relax (its vectors, of vbytes per value, and staging in
// one workgroup's LDS).
bool relax_small_fits(int nrows, int m, int m2, int S, int vbytes = 8);
// Per-device kernel attributes (dynamic LDS limits), set once by mgm_init.
int stencil_device_init();
int launch_sor_natural(Ell A, int N, const int *iperm, const double *src, double *p, double w,
                       double c, hipStream_t s);
// residual forms: 0 = interior, res = s - sum (boundary res = 0)            (H/T)
//                 1 = interior, res = s; res -= ... sequentially (boundary 0) (T fused)
//                 2 = all rows, res = s; res -= ... sequentially             (F fused)
//                 3 = interior, res = s - sum (boundary untouched)           (F)
template <typename TW, typename TV>
int launch_residual(EllT<TW> A, int N, int form, const uint8_t *btag, const TV *src, const TV *p,
                    TV *res, hipStream_t s);
// restrict: out[r] = btag ? (keep_boundary ? out[r] : 0) : sum_j w*in[col]
template <typename TW, typename TV>
int launch_restrict(EllT<TW> R, int N, const uint8_t *btag, int keep_boundary, const TV *in, TV *out,
                    hipStream_t s);
// prolong: p[r] += sum_j w*pc[col] on interior rows
template <typename TW, typename TV>
int launch_prolong(EllT<TW> P, int N, const uint8_t *btag, const TV *pc, TV *p, hipStream_t s);
template <typename TV>
int launch_zero(int N, const uint8_t *btag, int interior_only, TV *p, hipStream_t s);
// p -= p[ref0] on real rows; scratch: one device double
template <typename TV>
int launch_pin(int N, const uint8_t *btag, const int *ref0_dev, TV *p, double *scratch, hipStream_t s);
// sum over interior rows of (a-b)^2 into *out (device scalar); b nullable
int launch_sqdiff_sum(int N, const uint8_t *btag, const double *a, const double *b,
                      double *partial, double *out, hipStream_t s);
// out[i] = in[i] for n16 16-byte elements (streaming copy, HBM reference).
int launch_stream_copy(long n16, const void *in, void *out, hipStream_t s);
// Mixed precision: element-wise narrowing (n values) and p += e on non-padding rows.
int launch_narrow_f32(long n, const double *in, float *out, hipStream_t s);
int launch_narrow_bf16(long n, const double *in, bf16 *out, hipStream_t s);
int launch_add_correction(int N, const uint8_t *btag, const float *e, double *p, hipStream_t s);
// Row-major reference-order rows (N x m) -> slab in device order (iperm:
// reference -> device row); columns mapped through cloud and iperm.
int launch_gather_rows(int N, int m, int S, const double *rows, const int *iperm,
                       const int *cloud, double *w_out, int *col_out, hipStream_t s);
// Gather-sorted columns of a slab (nslices slices of S*m entries): per slice,
// the packed (slot, column) pairs in ascending (column, slot) order.
int launch_gather_sort(int nslices, int m, int S, const int *col, int *gcol, hipStream_t s);
// wg[e] = w[slice base of e + slot of gcol[e]]: weights in gather-sorted order.
int launch_permute_slab(long n, int m, int S, const int *gcol, const double *w, double *wg, hipStream_t s);
// Transfer rows of the target level (Nt x m, reference order) -> slab over the
// target's device order; column j = iperm_src[cloud_src[pts[i]*m + j]].
int launch_transfer_slab(int Nt, int m, int S, const double *rows, const int *iperm_t,
                         const int *pts, const int *cloud_src, const int *iperm_src,
                         double *w_out, int *col_out, hipStream_t s);
int launch_lap_poisson(int N, int m, int dim, const double *Dx, const double *Dy, const double *Dz,
                       const double *lap, const double *nx, const double *ny, const double *nz,
                       const uint8_t *btag, double *lapP, hipStream_t s);
int launch_dirichlet_rows(int N, int m, const uint8_t *btag, double *rows, hipStream_t s);
// y[iperm[bnodes[t]]] = sum_j w[t*mb + j] x[iperm[cloud_b[t*mb + j]]], in order
// from 0.0 (the wide boundary rows over a device-order vector).
int launch_wide_rows(int nbw, int mb, const int *bnodes, const int *cloud_b, const double *w, const int *iperm,
                     const double *x, double *y, hipStream_t s);

}  // namespace mgm
