// problem.hpp -- the multigrid hierarchy held by an mgm_problem.
#pragma once

#include <array>
#include <memory>
#include <vector>

#include "device.hpp"

struct mgm_points {
    mgm::PointSet ps;
};

namespace mgm {

// MGM_WINDOW_PROLONG=0: FP64 prolongations and restrictions gather from memory
// on blocked levels too.
bool prol_window_on();

enum { OP_DX = 0, OP_DY = 1, OP_DZ = 2, OP_LAP = 3, OP_LAPP = 4, NUM_OPS = 5 };

struct Level {
    // ---- input (file order, after corner removal) ----
    PointSet in;
    bool have_input = false;
    // ---- sizes ----
    int N = 0, nb = 0, m = 0, n = 0, poly = 0;
    double d_avg_bbox = 0, d_avg = 0;
    std::vector<short> px, py, pz;
    // ---- reference node order (after the BFS reordering) ----
    std::vector<double> x, y, z, nx, ny, nz;
    std::vector<uint8_t> btag, ctag;   // ctag is NOT permuted, as in the reference
    std::vector<int> cloud;            // N*m
    // Wide boundary stencils (mb > 0): the boundary nodes with a system
    // (ascending, reference order) and their mb-point stencils, nbw x mb.
    int mb = 0;
    std::vector<int> bnodes, cloud_b;
    std::vector<int> rp, pp;           // restriction / prolongation points
    std::vector<int> colour;
    int ncol = 0;
    // Device order: Morton-ranked blocks of rows; inside a block, colour
    // classes; each (block, colour) chunk padded to whole slices with dummy
    // rows (device tag 2, zero weights, never referenced).  Colour c is the
    // slice list col_slices[col_soff[c] .. col_soff[c+1]).
    int Nd = 0;                        // device rows including padding
    std::vector<int> col_slices, col_soff;
    DevBuf<int> d_col_slices, d_col_soff;
    bool col_contig = false;           // every colour's slices are consecutive (no list needed)
    bool blocked = false;              // rows in Morton blocks
    std::vector<int> perm, iperm;      // device -> reference (-1 = padding), reference -> device
    int knn_ties = 0;
    // ---- device, reference order ----
    DevBuf<double> rx, ry, rz, rnx, rny, rnz;
    DevBuf<uint8_t> rbtag;
    DevBuf<int> rcloud, d_iperm, d_rp, d_pp;
    DevBuf<short> d_px, d_py, d_pz;
    std::array<DevBuf<double>, NUM_OPS> rows;  // row-major N*m
    DevBuf<int> rcloud_b, d_bnodes;            // wide boundary stencils (mb > 0)
    std::array<DevBuf<double>, 4> rows_b;      // wide Dx, Dy, Dz, lap rows: bnodes.size() * mb
    DevBuf<double> Rrows;                      // this level as coarse: N * m_{l-1}
    DevBuf<double> Prows;                      // this level as fine:   N * m_{l+1}
    // ---- device, device order ----
    DevBuf<uint8_t> dbtag;
    DevBuf<double> dnx, dny, dnz;              // unit normals, device order
    DevBuf<int> col;                           // m*N
    std::array<DevBuf<double>, NUM_OPS> slab;  // m*N
    DevBuf<int> rcol, pcol;
    DevBuf<double> rw, pw;
    int mr = 0, mp = 0;                        // widths of restriction / prolongation rows
    int S = 8, Sr = 8, Sp = 8;                 // slice heights of the operator / transfer slabs
    std::array<DevBuf<double>, MGM_NUM_FIELDS> f;
    DevBuf<int> ref0;                          // device row of reference node 0
    // ---- mixed-precision copies (built on demand, device order) ----
    bool w16 = false;                          // this level's weights are stored as BF16
    DevBuf<float> lapp32, rw32, pw32;
    DevBuf<bf16> lapp16, rw16, pw16;
    DevBuf<float> p32, src32, res32;
    DevBuf<double> jtmp;                       // Jacobi sweeps: third buffer, so an odd count ends in p
    DevBuf<float> jtmp32;
    DevBuf<float> lap32, srcl32;               // time steps: FP32 lap; level 0's source for lap residuals
    DevBuf<float> glapp32, glap32, gpw32;      // gather-sorted narrowed copies (levels with gs / gsp)
    DevBuf<bf16> glapp16, gpw16;

    // ---- gather-sorted copies for the streaming passes (large levels) ----
    bool gs = false, gsp = false;              // operator / prolongation slabs have them
    DevBuf<int> gcol, gpcol;
    std::array<DevBuf<double>, NUM_OPS> gslab; // OP_LAP and OP_LAPP only
    DevBuf<double> gpw;

    // ---- LDS windows of the full passes (levels in Morton blocks) ----
    std::vector<int> blk_row0;                 // device row where each Morton block starts (+ end)
    DevBuf<uint16_t> win_col;                  // 16-bit places in the block's window
    DevBuf<int> win_row0, win_hoff, win_halo;
    int win_blocks = 0, win_max = 0;           // 0 blocks: no windows
    bool win_tried = false;                    // tables built (or found not to fit) once
    // the prolongation into this level in the window form: per block the
    // sorted coarse rows its prolongation rows reference, and the plain
    // prolongation slab's columns as 16-bit places in that list
    DevBuf<uint16_t> pwin_col;
    DevBuf<int> pwin_hoff, pwin_halo;
    int pwin_max = 0;                          // 0: no prolongation windows
    bool pwin_tried = false;
    // the restriction into this level (its rows read the finer level) likewise
    DevBuf<uint16_t> rwin_col;
    DevBuf<int> rwin_hoff, rwin_halo;
    int rwin_max = 0;
    bool rwin_tried = false;
    template <typename T>
    EllT<T> with_prol_window(EllT<T> A) const
    {
        if (pwin_max && prol_window_on()) {
            A.win_col = pwin_col.p; A.win_row0 = win_row0.p; A.win_hoff = pwin_hoff.p; A.win_halo = pwin_halo.p;
            A.win_blocks = win_blocks; A.win_max = pwin_max; A.win_self = 0;
        }
        return A;
    }
    template <typename T>
    EllT<T> with_window(EllT<T> A) const
    {
        if (win_blocks) {
            A.win_col = win_col.p; A.win_row0 = win_row0.p; A.win_hoff = win_hoff.p; A.win_halo = win_halo.p;
            A.win_blocks = win_blocks; A.win_max = win_max;
        }
        return A;
    }

    // The operator as the stencil passes read it: gather-sorted where built,
    // with the LDS-window tables where built.
    Ell op(int which) const
    {
        if (gs && gslab[which].p)
            return with_window(Ell{gcol.p, gslab[which].p, m, S, stream_slab(Nd, m, 8), slab[which].p, true,
                                   32 - slot_bits(S, m)});
        return with_window(op_plain(which));
    }
    // The operator in plain slice order (entry j of row r at ell_at(m, S, r, j)).
    Ell op_plain(int which) const { return Ell{col.p, slab[which].p, m, S, stream_slab(Nd, m, 8)}; }
    // slices of colour k for launch_sor_slices: a device list, or (null, first)
    void colour_slices(int k, const int **list, int *first, int *n) const
    {
        *n = col_soff[k + 1] - col_soff[k];
        *list = col_contig ? nullptr : d_col_slices.p + col_soff[k];
        *first = (col_contig && *n) ? col_slices[col_soff[k]] : 0;
    }
    Ell restr() const
    {
        Ell A{rcol.p, rw.p, mr, Sr, stream_slab(Nd, mr, 8)};
        if (rwin_max && prol_window_on()) {
            A.win_col = rwin_col.p; A.win_row0 = win_row0.p; A.win_hoff = rwin_hoff.p; A.win_halo = rwin_halo.p;
            A.win_blocks = win_blocks; A.win_max = rwin_max; A.win_self = 0;
        }
        return A;
    }
    Ell prol() const
    {
        if (gsp)
            return with_prol_window(
                Ell{gpcol.p, gpw.p, mp, Sp, stream_slab(Nd, mp, 8), pw.p, true, 32 - slot_bits(Sp, mp)});
        return with_prol_window(Ell{pcol.p, pw.p, mp, Sp, stream_slab(Nd, mp, 8)});
    }
};

}  // namespace mgm

namespace mgm {
// Output buffers of n chained Jacobi sweeps over p / res / a third buffer t,
// ending in p without a copy wherever possible (problem.cpp).
template <typename T>
T *const *jacobi_targets(int n, bool from_p, T *p, T *res, T *t);
// Learned smoother state: FP32 master weights, BF16 inference copies, the
// per-level geometric part of layer 1 and device-order coordinates.
struct LearnedState {
    int in = 0;                       // network inputs: m*dim + m of level 0
    std::vector<float> w;             // FP32 masters (host)
    DevBuf<float> W;                  // FP32 masters (device)
    bool have = false;                // weights set or trained
    bool enabled = false;             // relax uses the network as its first sweep
    bool prepared = false;            // inference copies and G match W
    DevBuf<bf16> W1r, W2, w3;            // W1r: residual columns of W1, [j][k]
    DevBuf<float> b2;
    DevBuf<uint8_t> net;              // the correction's packed network image (mlp_pack_net)
    float b3 = 0;
    std::vector<DevBuf<float>> G;     // per level: Nd x MLP_H
    std::vector<DevBuf<double>> dx, dy, dz;
    std::vector<char> lev_ok;         // level's stencils fit the network
    std::vector<char> lev_use;        // levels the caller lets use it (empty: all)
    int fell_back = 0;                // the last learned solve switched to SOR
};
}  // namespace mgm

struct mgm_problem {
    mgm_config cfg{};
    int L = 0;
    std::vector<mgm::Level> lv;
    bool ready = false;
    double timings[6] = {0, 0, 0, 0, 0, 0};
    mgm::DevBuf<double> partial, scal;
    double *host_scal = nullptr;   // pinned
    mgm_ns_params ns{};
    bool ns_set = false;
    int mp_built = 0;              // precision mode the FP32/BF16 copies hold (0: none / stale)
    bool mp_ns_built = false;      // FP32 lap copies of the time steps' pressure cycle are current
    bool res0_current = false;     // level 0's RES holds s - lap_Poison p (form 0) whenever a cycle starts
    // Inside mgm_solve with weighted Jacobi on level 0: the fine residual pass
    // also forms the first Jacobi sweep of p into jac0, and level 0's next
    // relaxation starts from it (jac0_ready).
    mgm::DevBuf<double> jac0;
    bool jac0_ready = false;
    std::unique_ptr<mgm::LearnedState> ml;
    std::vector<float> omega_lv;   // SOR factor per level; empty: cfg.omega on every level
    std::vector<int> smoother_lv;  // MGM_SMOOTH_* per level; empty: SOR on every level
    float omega(int l) const { return l < (int)omega_lv.size() ? omega_lv[l] : cfg.omega; }
    // the H variant may relax a level by weighted Jacobi; T and F keep SOR
    bool jacobi(int l, int variant) const
    {
        return variant == MGM_VAR_H && l < (int)smoother_lv.size() && smoother_lv[l] == MGM_SMOOTH_JACOBI;
    }
    ~mgm_problem();
};
