// problem.cpp -- setup pipeline, V-cycle operators, resident solve and the
// extern "C" entry points declared in include/mgm.h.
//
// Setup (reference: read_complete_mesh_data, hf/mesh_functions.h:914-928,
// and create_derivative_matrices, hf/rbf.h:401-421), per level:
//   1. stencils: GPU kNN over a uniform grid (knn.hip), file order
//   2. BFS reordering from node 0 (hf/mesh_functions.h:774-872), host
//   3. normal orientation (:478-527) and mean spacing (:895-910), host
//   4. restriction / prolongation points: GPU nearest neighbour
//   5. greedy colouring -> device order (colour, then Morton code)
//   6. weights + transfer rows: batched local solves (weights.hip)
//   7. lap_Poison rows, then column-major device-order slabs
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <thread>

#include "mlp.hpp"
#include "problem.hpp"

using namespace mgm;

namespace mgm {

static hipStream_t g_stream = nullptr;
static int g_device = -1;
hipStream_t stream() { return g_stream; }

static double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int noalloc_init(int device);
int spmv_rows_static(int device, int n_rows, int n_cols, const double *D, const int32_t *cloud, int ncol,
                     const double *f, double *out, hipStream_t s);

static int ensure_init()
{
    if (g_stream) return 0;
    MGM_FAIL(-9, "mgm_init(device) has not been called (or no GPU is visible)");
}

// ---------------------------------------------------------------------------
// Uniform grid over a candidate subset of a point set.
struct HostGrid {
    KnnGrid g{};
    std::vector<int> start, pts;
    DevBuf<int> d_start, d_pts;
};

static int build_grid(const double *x, const double *y, const double *z, int N,
                      const std::vector<int> &cand, int K, int dim, HostGrid &G)
{
    if (cand.empty()) MGM_FAIL(-5, "kNN: empty candidate set");
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    const double *P[3] = {x, y, z};
    for (int id : cand)
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], P[d][id]);
            hi[d] = std::max(hi[d], P[d][id]);
        }
    (void)N;
    double ext[3], vol = 1;
    int active = 0;
    for (int d = 0; d < 3; d++) {
        ext[d] = hi[d] - lo[d];
        if (ext[d] > 0 && (d < 2 || dim == 3)) { vol *= ext[d]; active++; }
    }
    double c;
    if (active == 0) c = 1;
    else c = 0.8 * pow((double)K * vol / (double)cand.size(), 1.0 / active);
    if (!(c > 0)) c = 1;
    int dims[3];
    long ncell = 1;
    for (int d = 0; d < 3; d++) {
        if (ext[d] > 0 && (d < 2 || dim == 3)) dims[d] = std::max(1, (int)ceil(ext[d] / c));
        else dims[d] = 1;
        ncell *= dims[d];
    }
    while (ncell > 8L * (long)cand.size() + 64) {  // guard degenerate extents
        c *= 1.5;
        ncell = 1;
        for (int d = 0; d < 3; d++) {
            dims[d] = (dims[d] == 1) ? 1 : std::max(1, (int)ceil(ext[d] / c));
            ncell *= dims[d];
        }
    }
    G.g.x0 = lo[0]; G.g.y0 = lo[1]; G.g.z0 = lo[2];
    G.g.c = c; G.g.inv_c = 1.0 / c;
    G.g.nx = dims[0]; G.g.ny = dims[1]; G.g.nz = dims[2];
    G.start.assign(ncell + 1, 0);
    std::vector<int> cellof(cand.size());
    for (size_t e = 0; e < cand.size(); e++) {
        int id = cand[e];
        int ci[3];
        for (int d = 0; d < 3; d++) {
            double f = floor((P[d][id] - lo[d]) * G.g.inv_c);
            ci[d] = (f < 0) ? 0 : (f > dims[d] - 1 ? dims[d] - 1 : (int)f);
        }
        int cell = (ci[2] * dims[1] + ci[1]) * dims[0] + ci[0];
        cellof[e] = cell;
        G.start[cell + 1]++;
    }
    for (long k = 0; k < ncell; k++) G.start[k + 1] += G.start[k];
    G.pts.resize(cand.size());
    std::vector<int> fill(G.start.begin(), G.start.end() - 1);
    for (size_t e = 0; e < cand.size(); e++) G.pts[fill[cellof[e]]++] = cand[e];  // ascending id
    HIP_TRY(G.d_start.upload(G.start, g_stream));
    HIP_TRY(G.d_pts.upload(G.pts, g_stream));
    G.g.cell_start = G.d_start.p;
    G.g.cell_pts = G.d_pts.p;
    return 0;
}

// Insertion-ordered kd-tree over a candidate subset (hf/kdtree.c:167-194),
// used only to order equal distances as the reference's range query does.
struct HostKd {
    std::vector<int> node_of, parent, depth;
    std::vector<uint8_t> side;
    std::vector<double> pos;
    DevBuf<int> d_node_of, d_parent, d_depth;
    DevBuf<uint8_t> d_side;
    DevBuf<double> d_pos;
    KdOrder kd{};
};

static int build_kd(const double *x, const double *y, const double *z, int N,
                    const std::vector<int> &cand, HostKd &T)
{
    T.node_of.assign(N, -1);
    size_t nn = cand.size();
    T.parent.assign(nn, -1); T.depth.assign(nn, 0); T.side.assign(nn, 0); T.pos.resize(3 * nn);
    std::vector<int> left(nn, -1), right(nn, -1);
    for (size_t e = 0; e < nn; e++) {
        int id = cand[e];
        int nd = (int)e;
        T.node_of[id] = nd;
        T.pos[3 * e] = x[id]; T.pos[3 * e + 1] = y[id]; T.pos[3 * e + 2] = z[id];
        if (e == 0) continue;
        const double p[3] = {x[id], y[id], z[id]};
        int cur = 0;
        for (;;) {
            int d = T.depth[cur] % 3;
            if (p[d] < T.pos[3 * cur + d]) {
                if (left[cur] < 0) { left[cur] = nd; T.side[nd] = 0; break; }
                cur = left[cur];
            } else {
                if (right[cur] < 0) { right[cur] = nd; T.side[nd] = 1; break; }
                cur = right[cur];
            }
        }
        T.parent[nd] = cur;
        T.depth[nd] = T.depth[cur] + 1;
    }
    HIP_TRY(T.d_node_of.upload(T.node_of, g_stream));
    HIP_TRY(T.d_parent.upload(T.parent, g_stream));
    HIP_TRY(T.d_depth.upload(T.depth, g_stream));
    HIP_TRY(T.d_side.upload(T.side, g_stream));
    HIP_TRY(T.d_pos.upload(T.pos, g_stream));
    T.kd = KdOrder{T.d_node_of.p, T.d_parent.p, T.d_depth.p, T.d_side.p, T.d_pos.p};
    return 0;
}

// Runs one kNN pass; queries with tied distances are rerun with the kd order.
// Returns the number of queries that needed the tie-break.
static int knn_pass(const double *hx, const double *hy, const double *hz, int Ncand_set,
                    const double *d_cx, const double *d_cy, const double *d_cz,
                    const std::vector<int> &cand, int dim,
                    const double *d_qx, const double *d_qy, const double *d_qz, int Nq,
                    const std::vector<int> &queries, int K, int K_out, int slot, int *d_out,
                    int stride, int *ties_out)
{
    if (queries.empty()) { *ties_out = 0; return 0; }
    HostGrid G;
    int rc = build_grid(hx, hy, hz, Ncand_set, cand, K, dim, G);
    if (rc) return rc;
    G.g.px = d_cx; G.g.py = d_cy; G.g.pz = d_cz;
    DevBuf<int> d_q;
    DevBuf<uint8_t> d_tie;
    HIP_TRY(d_q.upload(queries, g_stream));
    HIP_TRY(d_tie.alloc(Nq));
    HIP_TRY(hipMemsetAsync(d_tie.p, 0, Nq, g_stream));
    rc = launch_knn(G.g, nullptr, (int)queries.size(), d_qx, d_qy, d_qz, d_q.p, K, K_out, slot, d_out,
                    stride, d_tie.p, g_stream);
    if (rc) return rc;
    std::vector<uint8_t> tie(Nq);
    HIP_TRY(d_tie.download(tie.data(), Nq, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    std::vector<int> again;
    for (int q : queries) {
        if (tie[q] == 2) MGM_FAIL(-6, "kNN: fewer candidates than stencil points");
        if (tie[q] == 1) again.push_back(q);
    }
    *ties_out = (int)again.size();
    if (again.empty()) return 0;
    HostKd T;
    rc = build_kd(hx, hy, hz, Ncand_set, cand, T);
    if (rc) return rc;
    DevBuf<int> d_again;
    HIP_TRY(d_again.upload(again, g_stream));
    rc = launch_knn(G.g, &T.kd, (int)again.size(), d_qx, d_qy, d_qz, d_again.p, K, K_out, slot,
                    d_out, stride, nullptr, g_stream);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(g_stream));
    return 0;
}

// find_cloud_index (hf/kdtree_functions.h:113-151) on the GPU, file order.
static int level_knn(Level &L, int dim)
{
    const int N = L.N, m = L.m, nb = L.nb;
    const PointSet &P = L.in;
    DevBuf<double> dx, dy, dz;
    DevBuf<int> dcloud;
    HIP_TRY(dx.upload(P.x, g_stream));
    HIP_TRY(dy.upload(P.y, g_stream));
    HIP_TRY(dz.upload(P.z, g_stream));
    HIP_TRY(dcloud.alloc((size_t)N * m));
    HIP_TRY(hipMemsetAsync(dcloud.p, 0, sizeof(int) * (size_t)N * m, g_stream));
    std::vector<int> candA, candB, qA, qB;
    for (int i = 0; i < N; i++) {
        if (!P.ctag[i]) candA.push_back(i);
        if (!P.btag[i]) candB.push_back(i);
    }
    for (int i = nb; i < N; i++)
        if (!P.ctag[i]) qA.push_back(i);
    for (int i = 0; i < nb; i++) qB.push_back(i);
    int tA = 0, tB = 0, rc;
    rc = knn_pass(P.x.data(), P.y.data(), P.z.data(), N, dx.p, dy.p, dz.p, candA, dim, dx.p, dy.p,
                  dz.p, N, qA, m + 1, m, 0, dcloud.p, m, &tA);
    if (rc) return rc;
    rc = knn_pass(P.x.data(), P.y.data(), P.z.data(), N, dx.p, dy.p, dz.p, candB, dim, dx.p, dy.p,
                  dz.p, N, qB, m, m - 1, 1, dcloud.p, m, &tB);
    if (rc) return rc;
    L.knn_ties = tA + tB;
    L.cloud.resize((size_t)N * m);
    HIP_TRY(dcloud.download(L.cloud.data(), (size_t)N * m, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    for (int i = 0; i < nb; i++) L.cloud[(size_t)i * m] = i;
    // wide boundary stencils: the node itself, then its mb - 1 nearest
    // interior nodes (file order here; level_bfs maps them)
    L.cloud_b.clear();
    if (L.mb) {
        const int mb = L.mb;
        DevBuf<int> dwide;
        HIP_TRY(dwide.alloc((size_t)nb * mb));
        HIP_TRY(hipMemsetAsync(dwide.p, 0, sizeof(int) * (size_t)nb * mb, g_stream));
        int tW = 0;
        rc = knn_pass(P.x.data(), P.y.data(), P.z.data(), N, dx.p, dy.p, dz.p, candB, dim, dx.p, dy.p,
                      dz.p, N, qB, mb, mb - 1, 1, dwide.p, mb, &tW);
        if (rc) return rc;
        L.knn_ties += tW;
        L.cloud_b.resize((size_t)nb * mb);
        HIP_TRY(dwide.download(L.cloud_b.data(), (size_t)nb * mb, g_stream));
        HIP_TRY(hipStreamSynchronize(g_stream));
        for (int i = 0; i < nb; i++) L.cloud_b[(size_t)i * mb] = i;
    }
    return 0;
}

// rcm_reordering_with_boundarynodes (hf/mesh_functions.h:774-872).
// f(i0, i1) over [0, N) in contiguous chunks on up to 16 host threads
// (one thread below 65536 rows).
template <typename F>
static void parallel_rows(int N, F &&f)
{
    const int nt = N < 65536 ? 1 : (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (nt == 1) { f(0, N); return; }
    std::vector<std::thread> th;
    const int chunk = (N + nt - 1) / nt;
    for (int t = 0; t < nt; t++) {
        const int a = std::min(N, t * chunk), b = std::min(N, a + chunk);
        th.emplace_back([&f, a, b] { f(a, b); });
    }
    for (auto &t : th) t.join();
}

static int level_bfs(Level &L)
{
    const int N = L.N, m = L.m;
    std::vector<int> q1(N), q2(N);
    std::vector<uint8_t> vis(N, 0);
    int count = 1, head = 0;
    q1[0] = 0; vis[0] = 1;
    while (head < count) {
        int cur = q1[head++];
        const int *cr = &L.cloud[(size_t)cur * m];
        for (int i = 1; i < m; i++) {
            int nbr = cr[i];
            if (!vis[nbr]) { q1[count++] = nbr; vis[nbr] = 1; }
        }
    }
    if (count != N) MGM_FAIL(-7, "BFS reordering did not reach every node (disconnected stencil graph)");
    for (int i = 0; i < N; i++) q2[q1[i]] = i;
    const PointSet &P = L.in;
    L.x.resize(N); L.y.resize(N); L.z.resize(N); L.nx.resize(N); L.ny.resize(N); L.nz.resize(N);
    L.btag.resize(N);
    for (int i = 0; i < N; i++) {
        int s = q1[i];
        L.x[i] = P.x[s]; L.y[i] = P.y[s]; L.z[i] = P.z[s];
        L.nx[i] = P.nx[s]; L.ny[i] = P.ny[s]; L.nz[i] = P.nz[s];
        L.btag[i] = P.btag[s];
    }
    L.ctag = P.ctag;
    std::vector<int> nc((size_t)N * m);
    parallel_rows(N, [&](int i0, int i1) {
        for (int i = i0; i < i1; i++) {
            const int *src = &L.cloud[(size_t)q1[i] * m];
            int *dst = &nc[(size_t)i * m];
            for (int j = 0; j < m; j++) dst[j] = q2[src[j]];
        }
    });
    L.cloud.swap(nc);
    // wide boundary stencils in the new order: one row per boundary node with a
    // system (the reference skips nodes whose corner tag, in file order, is set)
    L.bnodes.clear();
    std::vector<int> wide;
    if (L.mb) {
        const int mb = L.mb;
        for (int i = 0; i < N; i++)
            if (L.btag[i] && !L.ctag[i]) L.bnodes.push_back(i);
        wide.resize(L.bnodes.size() * (size_t)mb);
        for (size_t t = 0; t < L.bnodes.size(); t++) {
            const int *src = &L.cloud_b[(size_t)q1[L.bnodes[t]] * mb];
            for (int j = 0; j < mb; j++) wide[t * mb + j] = q2[src[j]];
        }
    }
    L.cloud_b.swap(wide);
    return 0;
}

// correct_normal_directions (hf/mesh_functions.h:478-527) and
// calculate_avg_dx (:895-910).
static void level_normals_spacing(Level &L, int dim)
{
    const int N = L.N, m = L.m;
    for (int i = 0; i < N; i++) {
        if (!L.btag[i]) continue;
        double xa = 0, ya = 0, za = 0;
        short count = 0;
        for (int j = m - 1; j > 0; j--) {
            int c = L.cloud[(size_t)i * m + j];
            if (!L.btag[c]) {
                xa += L.x[c]; ya += L.y[c]; za += L.z[c];
                count++;
                if (count == 1) break;
            }
        }
        xa = xa / count; ya = ya / count; za = za / count;
        double ddx = xa - L.x[i], ddy = ya - L.y[i], ddz = za - L.z[i];
        double dot = ddx * L.nx[i] + ddy * L.ny[i] + ddz * L.nz[i];
        if (dot > 0) { L.nx[i] = -L.nx[i]; L.ny[i] = -L.ny[i]; L.nz[i] = -L.nz[i]; }
        double mag = sqrt(L.nx[i] * L.nx[i] + L.ny[i] * L.ny[i] + L.nz[i] * L.nz[i]);
        L.nx[i] = L.nx[i] / mag; L.ny[i] = L.ny[i] / mag; L.nz[i] = L.nz[i] / mag;
    }
    double s = 0;
    for (int i = 0; i < N; i++) {
        if (L.btag[i]) continue;
        int c = L.cloud[(size_t)i * m + 1];
        double ddx = L.x[i] - L.x[c], ddy = L.y[i] - L.y[c];
        double ddz = (dim == 3) ? L.z[i] - L.z[c] : 0;
        s += sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
    }
    L.d_avg = s / N;
}

// find_nearest_point (hf/kdtree_functions.h:153-175): for every node of A the
// nearest non-corner node of B (reference order on both sides).
static int nearest_points(const Level &A, const Level &B, int dim, std::vector<int> &out)
{
    DevBuf<double> ax, ay, az, bx, by, bz;
    HIP_TRY(ax.upload(A.x, g_stream)); HIP_TRY(ay.upload(A.y, g_stream)); HIP_TRY(az.upload(A.z, g_stream));
    HIP_TRY(bx.upload(B.x, g_stream)); HIP_TRY(by.upload(B.y, g_stream)); HIP_TRY(bz.upload(B.z, g_stream));
    std::vector<int> cand, q;
    for (int i = 0; i < B.N; i++)
        if (!B.ctag[i]) cand.push_back(i);
    for (int i = 0; i < A.N; i++)
        if (!A.ctag[i]) q.push_back(i);
    DevBuf<int> d_out;
    HIP_TRY(d_out.alloc(A.N));
    HIP_TRY(hipMemsetAsync(d_out.p, 0, sizeof(int) * A.N, g_stream));
    int ties = 0;
    int rc = knn_pass(B.x.data(), B.y.data(), B.z.data(), B.N, bx.p, by.p, bz.p, cand, dim, ax.p,
                      ay.p, az.p, A.N, q, 2, 1, 0, d_out.p, 1, &ties);
    if (rc) return rc;
    out.resize(A.N);
    HIP_TRY(d_out.download(out.data(), A.N, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    return 0;
}

static uint64_t spread_bits(uint64_t v, int stride)
{
    uint64_t r = 0;
    for (int b = 0; b < 21; b++)
        if (v & (1ull << b)) r |= 1ull << (b * stride);
    return r;
}

// Levels above this many rows (MGM_BLOCKED_ROWS, default 262,144: vectors
// that no longer fit an XCD's L2 many times over) are ordered in Morton
// blocks of MGM_BLOCK_ROWS (4096) rows and read gather-sorted slabs.  Read
// whenever a level is ordered, so tests can reproduce the large-level layout
// on small clouds.
constexpr int kLargeRows = 262144;
static long blocked_rows()
{
    const char *e = getenv("MGM_BLOCKED_ROWS");
    return e ? atol(e) : kLargeRows;
}
// Mid-sized levels (above MGM_MID_BLOCKED_ROWS, default 65,536 rows, up to
// the large-level threshold) are cut into smaller Morton blocks of at most
// 512 rows -- one per CU on the 1M cube's 125k-row level 1 -- so that their
// full passes read x from LDS windows too.
static long mid_blocked_rows()
{
    const char *e = getenv("MGM_MID_BLOCKED_ROWS");
    return e ? atol(e) : 65536;
}


int small_level_rows();

// Slice height of a level's operator slabs.  Stencils of at most 36 entries:
// 16.  Wider ones: 8 on levels small enough for the single-workgroup relax
// and on large levels (bandwidth-bound passes over gather-sorted slabs); 4 on
// the levels between, whose colour steps are latency-bound -- half as many
// entries per wave shortens each wave's load chain (level-2 relaxation of the
// 1M cube 441 instead of 520 us).  A large level whose padded rows exceed the
// gather-sorted column range at 8 (above ~4.2M rows) takes 4, which keeps it
// sortable.  A mid-sized level in Morton blocks takes 8 as well: its full
// passes read LDS windows and gain from 8-row slices (1M cube, against the
// unblocked level 1 at 4: tuned Jacobi solve 0.0493 -> 0.0467 s; with SOR on
// every level its colour steps lose a little, 0.127 -> 0.129 s).
// MGM_SLICE_HEIGHT overrides all of this.
static long mid_blocked_rows();
static int level_slice_height(int m, int N, int ncol, int nblocks)
{
    const int base = slice_height(m);
    if (getenv("MGM_SLICE_HEIGHT") || base != 8) return base;
    if (N <= small_level_rows()) return 8;
    if (N <= blocked_rows()) return N > mid_blocked_rows() ? 8 : 4;
    const auto padded = [&](int S) { return (long)N + (long)S * ncol * nblocks; };
    if (!slab_sortable(8, m, padded(8)) && slab_sortable(4, m, padded(4))) return 4;
    return 8;
}

static void order_by_colour(Level &L, int dim);

// Greedy colouring on the symmetrised stencil graph, in reference order, then
// the device order: colour classes contiguous, Morton order inside a class.
static void level_colour_order(Level &L, int dim)
{
    const int N = L.N, m = L.m;
    // Nodes are coloured in index order, each with the smallest colour that no
    // earlier node adjacent to it in the symmetrised graph has: neither an
    // earlier node of its own stencil nor an earlier node whose stencil holds
    // it.  The second set is pushed forward as a bit mask when that node is
    // coloured, which needs no transposed graph; with more than 64 colours the
    // transposed-graph form below takes over (same colours).
    L.colour.assign(N, -1);
    {
        std::vector<uint64_t> pushed(N, 0);
        int maxc = 0, i = 0;
        for (; i < N; i++) {
            const int *cr = &L.cloud[(size_t)i * m];
            uint64_t forb = pushed[i];
            for (int j = 0; j < m; j++)
                if (cr[j] < i) forb |= 1ull << L.colour[cr[j]];
            if (~forb == 0) break;
            const int col = __builtin_ctzll(~forb);
            L.colour[i] = col;
            maxc = std::max(maxc, col + 1);
            const uint64_t bit = 1ull << col;
            for (int j = 0; j < m; j++)
                if (cr[j] > i) pushed[cr[j]] |= bit;
        }
        if (i == N) {
            L.ncol = maxc;
            order_by_colour(L, dim);
            return;
        }
    }
    std::vector<int> deg(N + 1, 0);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < m; j++) {
            int c = L.cloud[(size_t)i * m + j];
            if (c != i) deg[c + 1]++;
        }
    for (int i = 0; i < N; i++) deg[i + 1] += deg[i];
    std::vector<int> radj(deg[N]), fill(deg.begin(), deg.end() - 1);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < m; j++) {
            int c = L.cloud[(size_t)i * m + j];
            if (c != i) radj[fill[c]++] = i;
        }
    L.colour.assign(N, -1);
    std::vector<int> mark(4096, 0);
    int maxc = 0;
    for (int i = 0; i < N; i++) {
        int stamp = i + 1;
        for (int j = 0; j < m; j++) {
            int c = L.cloud[(size_t)i * m + j];
            if (c != i && L.colour[c] >= 0) mark[L.colour[c]] = stamp;
        }
        for (int k = deg[i]; k < deg[i + 1]; k++) {
            int c = radj[k];
            if (L.colour[c] >= 0) mark[L.colour[c]] = stamp;
        }
        int col = 0;
        while (col < 4095 && mark[col] == stamp) col++;
        L.colour[i] = col;
        maxc = std::max(maxc, col + 1);
    }
    L.ncol = maxc;
    order_by_colour(L, dim);
}

// Device order of a coloured level: Morton-ranked blocks, colour classes
// inside a block, slices padded per (block, colour) chunk.
static void order_by_colour(Level &L, int dim)
{
    const int N = L.N;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    const std::vector<double> *P[3] = {&L.x, &L.y, &L.z};
    for (int d = 0; d < 3; d++)
        for (int i = 0; i < N; i++) { lo[d] = std::min(lo[d], (*P[d])[i]); hi[d] = std::max(hi[d], (*P[d])[i]); }
    std::vector<uint64_t> key(N);
    for (int i = 0; i < N; i++) {
        uint64_t k = 0;
        for (int d = 0; d < dim; d++) {
            double e = hi[d] - lo[d];
            double t = e > 0 ? ((*P[d])[i] - lo[d]) / e : 0;
            uint64_t q = (uint64_t)std::min(2097151.0, std::max(0.0, t * 2097151.0));
            k |= spread_bits(q, dim) << d;
        }
        key[i] = k;
    }
    // Morton rank -> block of kBlock rows; inside a block, colour classes.
    // Levels whose vectors fit one L2 many times over stay one block, so each
    // colour class is one contiguous run of slices.
    std::vector<int> byk(N);
    std::iota(byk.begin(), byk.end(), 0);
    std::sort(byk.begin(), byk.end(), [&](int a, int b) { return key[a] != key[b] ? key[a] < key[b] : a < b; });
    // A blocked level's full passes run one workgroup per block and one block
    // per CU (the window fills the LDS), so the number of blocks is a multiple
    // of the CU count: blocks of at most the nominal size (4,096 rows, or 512
    // on mid-sized levels), as many as the smallest such multiple needs.  At
    // 1M rows that is 256 blocks of 3,904 rows: the level-0 apply takes 148 us
    // instead of 154 us with 244 blocks of 4,096 (12 idle CUs), and 264 us with
    // 261 blocks (a second round of workgroups).  MGM_BLOCK_ROWS and
    // MGM_MID_BLOCK_ROWS set the sizes directly.
    const auto fit_blocks = [&](int nominal) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g_device) != hipSuccess || cus < 1)
            cus = 256;
        const long per = (long)nominal * cus;
        if (2L * N < per) return nominal;   // too few rows to fill the CUs with blocks of half the size
        const long nb = (long)cus * ((N + per - 1) / per);
        const long rows = ((N + nb - 1) / nb + 7) / 8 * 8;
        return (int)std::max(8L, rows);
    };
    const int block_rows = [&] {
        const char *e = getenv("MGM_BLOCK_ROWS");
        const int v = e ? atoi(e) : 0;
        return v > 0 ? v : fit_blocks(4096);
    }();
    const int mid_rows = [&] {
        const char *e = getenv("MGM_MID_BLOCK_ROWS");
        const int v = e ? atoi(e) : 0;
        return v > 0 ? v : fit_blocks(512);
    }();
    const int kBlock = N > blocked_rows() ? block_rows : N > mid_blocked_rows() ? mid_rows : N;
    L.blocked = kBlock < N;
    std::vector<int> block(N);
    for (int r = 0; r < N; r++) block[byk[r]] = r / kBlock;
    std::vector<int> ord(byk);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
        if (block[a] != block[b]) return block[a] < block[b];
        return L.colour[a] < L.colour[b];
    });
    L.S = level_slice_height(L.m, N, L.ncol, (N + kBlock - 1) / kBlock);
    const int S = L.S;
    std::vector<std::vector<int>> slices(L.ncol);
    L.perm.clear();
    L.perm.reserve((size_t)N + (size_t)S * L.ncol * (N / kBlock + 1));
    L.iperm.assign(N, -1);
    const int nmb = (N + kBlock - 1) / kBlock;
    std::vector<int> blk_first(nmb + 1, -1);
    for (size_t q = 0; q < ord.size();) {
        size_t e = q;
        while (e < ord.size() && block[ord[e]] == block[ord[q]] && L.colour[ord[e]] == L.colour[ord[q]]) e++;
        const int c = L.colour[ord[q]];
        const int first = (int)L.perm.size();
        if (blk_first[block[ord[q]]] < 0) blk_first[block[ord[q]]] = first;
        for (size_t t = q; t < e; t++) {
            L.iperm[ord[t]] = (int)L.perm.size();
            L.perm.push_back(ord[t]);
        }
        while (L.perm.size() % S) L.perm.push_back(-1);
        for (int s = first / S; s < (int)(L.perm.size() / S); s++) slices[c].push_back(s);
        q = e;
    }
    L.Nd = (int)L.perm.size();
    blk_first[nmb] = L.Nd;
    for (int b = nmb - 1; b >= 0; b--)
        if (blk_first[b] < 0) blk_first[b] = blk_first[b + 1];
    L.blk_row0.clear();
    if (L.blocked) L.blk_row0 = blk_first;
    L.col_soff.assign(L.ncol + 1, 0);
    L.col_slices.clear();
    L.col_contig = true;
    for (int c = 0; c < L.ncol; c++) {
        for (size_t q = 1; q < slices[c].size(); q++)
            if (slices[c][q] != slices[c][q - 1] + 1) L.col_contig = false;
        L.col_slices.insert(L.col_slices.end(), slices[c].begin(), slices[c].end());
        L.col_soff[c + 1] = (int)L.col_slices.size();
    }
}

static int upload_level(Level &L)
{
    HIP_TRY(L.rx.upload(L.x, g_stream)); HIP_TRY(L.ry.upload(L.y, g_stream)); HIP_TRY(L.rz.upload(L.z, g_stream));
    HIP_TRY(L.rnx.upload(L.nx, g_stream)); HIP_TRY(L.rny.upload(L.ny, g_stream)); HIP_TRY(L.rnz.upload(L.nz, g_stream));
    HIP_TRY(L.rbtag.upload(L.btag, g_stream));
    HIP_TRY(L.rcloud.upload(L.cloud, g_stream));
    if (L.mb) {
        HIP_TRY(L.rcloud_b.upload(L.cloud_b, g_stream));
        HIP_TRY(L.d_bnodes.upload(L.bnodes, g_stream));
    }
    HIP_TRY(L.d_iperm.upload(L.iperm, g_stream));
    HIP_TRY(L.d_rp.upload(L.rp, g_stream));
    HIP_TRY(L.d_pp.upload(L.pp, g_stream));
    HIP_TRY(L.d_px.upload(L.px, g_stream)); HIP_TRY(L.d_py.upload(L.py, g_stream)); HIP_TRY(L.d_pz.upload(L.pz, g_stream));
    HIP_TRY(L.d_col_slices.upload(L.col_slices, g_stream));
    HIP_TRY(L.d_col_soff.upload(L.col_soff, g_stream));
    const int Nd = L.Nd;
    std::vector<uint8_t> db(Nd);
    for (int d = 0; d < Nd; d++) db[d] = L.perm[d] < 0 ? (uint8_t)ROW_PAD : (L.btag[L.perm[d]] ? 1 : 0);
    HIP_TRY(L.dbtag.upload(db, g_stream));
    std::vector<double> t[3];
    const std::vector<double> *src[3] = {&L.nx, &L.ny, &L.nz};
    DevBuf<double> *dst[3] = {&L.dnx, &L.dny, &L.dnz};
    for (int k = 0; k < 3; k++) {
        t[k].assign(Nd, 0.0);
        for (int d = 0; d < Nd; d++)
            if (L.perm[d] >= 0) t[k][d] = (*src[k])[L.perm[d]];
        HIP_TRY(dst[k]->upload(t[k], g_stream));
    }
    int r0 = L.iperm[0];
    HIP_TRY(L.ref0.upload(&r0, 1, g_stream));
    for (auto &f : L.f) {
        HIP_TRY(f.alloc(Nd));
        HIP_TRY(hipMemsetAsync(f.p, 0, sizeof(double) * Nd, g_stream));
    }
    HIP_TRY(L.jtmp.alloc(Nd));   // allocated here: the sweeps may run under graph capture
    HIP_TRY(hipStreamSynchronize(g_stream));  // host staging vectors die here
    return 0;
}

// The level relaxes in one single-workgroup launch (op_relax).
static bool relax_small_here(const mgm_problem *pb, const Level &L)
{
    return pb->cfg.sweep_order != MGM_ORDER_NATURAL && L.Nd <= small_level_rows() &&
           relax_small_fits(L.Nd, L.m, L.m, L.S);
}

static bool row_valid(const Level &L, int i, int mode)
{
    return mode == MGM_ROWS_REFERENCE ? (i >= L.nb) : (L.btag[i] == 0);
}

// Inputs of the batched weight + transfer-row solves of one level: its
// systems, the interpolation targets mapped onto them, and the output rows
// (zeroed).  The buffers stay alive until the launch has finished.
struct WeightsPrep {
    DevBuf<int> d_sys, d_roff, d_poff, d_rrow, d_prow;
    DevBuf<double> d_rxyz, d_pxyz, ws, rhs;
    DevBuf<int> piv;
    WeightsArgs a{};
    // wide boundary systems (mb > 0): rows 0 .. nbw-1 of the wide stencils,
    // the four derivative operators only
    DevBuf<int> d_rows;
    WeightsArgs wide{};
    bool has_wide = false;
};

static int prepare_weights(mgm_problem *pb, int l, WeightsPrep &w, int *d_status)
{
    Level &L = pb->lv[l];
    const int dim = pb->cfg.dimension, mode = pb->cfg.transfer_rows;
    const int N = L.N, m = L.m, mpn = L.m + L.n;
    for (int o = 0; o < NUM_OPS; o++) {
        if (o == OP_DZ && dim != 3) continue;
        HIP_TRY(L.rows[o].alloc((size_t)N * m));
        HIP_TRY(hipMemsetAsync(L.rows[o].p, 0, sizeof(double) * N * m, g_stream));
    }
    std::vector<int> sys;
    for (int i = 0; i < N; i++)
        if (!L.ctag[i]) sys.push_back(i);
    // interpolation targets mapped onto this level's systems
    std::vector<int> roff(N + 1, 0), poff(N + 1, 0), rrow, prow;
    std::vector<double> rxyz, pxyz;
    if (l + 1 < pb->L) {
        Level &C = pb->lv[l + 1];
        for (int i = 0; i < C.N; i++)
            if (row_valid(C, i, mode)) roff[C.rp[i] + 1]++;
        for (int s = 0; s < N; s++) roff[s + 1] += roff[s];
        rrow.resize(roff[N]); rxyz.resize(3 * (size_t)roff[N]);
        std::vector<int> fill(roff.begin(), roff.end() - 1);
        for (int i = 0; i < C.N; i++)
            if (row_valid(C, i, mode)) {
                int k = fill[C.rp[i]]++;
                rrow[k] = i; rxyz[3 * k] = C.x[i]; rxyz[3 * k + 1] = C.y[i]; rxyz[3 * k + 2] = C.z[i];
            }
        HIP_TRY(C.Rrows.alloc((size_t)C.N * m));
        HIP_TRY(hipMemsetAsync(C.Rrows.p, 0, sizeof(double) * C.N * m, g_stream));
    }
    if (l > 0) {
        Level &F = pb->lv[l - 1];
        for (int i = 0; i < F.N; i++)
            if (row_valid(F, i, mode)) poff[F.pp[i] + 1]++;
        for (int s = 0; s < N; s++) poff[s + 1] += poff[s];
        prow.resize(poff[N]); pxyz.resize(3 * (size_t)poff[N]);
        std::vector<int> fill(poff.begin(), poff.end() - 1);
        for (int i = 0; i < F.N; i++)
            if (row_valid(F, i, mode)) {
                int k = fill[F.pp[i]]++;
                prow[k] = i; pxyz[3 * k] = F.x[i]; pxyz[3 * k + 1] = F.y[i]; pxyz[3 * k + 2] = F.z[i];
            }
        HIP_TRY(F.Prows.alloc((size_t)F.N * m));
        HIP_TRY(hipMemsetAsync(F.Prows.p, 0, sizeof(double) * F.N * m, g_stream));
    }
    HIP_TRY(w.d_sys.upload(sys, g_stream));
    HIP_TRY(w.d_roff.upload(roff, g_stream)); HIP_TRY(w.d_poff.upload(poff, g_stream));
    if (!rrow.empty()) { HIP_TRY(w.d_rrow.upload(rrow, g_stream)); HIP_TRY(w.d_rxyz.upload(rxyz, g_stream)); }
    if (!prow.empty()) { HIP_TRY(w.d_prow.upload(prow, g_stream)); HIP_TRY(w.d_pxyz.upload(pxyz, g_stream)); }
    // batch sized for ~4 GiB of factor workspace (only the wide-system kernel uses one)
    long batch = (long)((4.0 * 1024 * 1024 * 1024) / (8.0 * mpn * mpn));
    batch = std::max(64L, std::min(batch, (long)sys.size()));
    batch = (batch + 63) / 64 * 64;
    if (!weights_tile(mpn) && !lu32_tiles(mpn)) {
        HIP_TRY(w.ws.alloc((size_t)mpn * mpn * batch));
        HIP_TRY(w.rhs.alloc((size_t)4 * mpn * batch));
        HIP_TRY(w.piv.alloc((size_t)mpn * batch));
    }
    WeightsArgs &a = w.a;
    a.nsys = (int)sys.size(); a.sys = w.d_sys.p; a.m = m; a.n = L.n; a.dim = dim; a.phs = pb->cfg.phs_degree;
    a.px = L.d_px.p; a.py = L.d_py.p; a.pz = L.d_pz.p;
    a.x = L.rx.p; a.y = L.ry.p; a.z = L.rz.p; a.cloud = L.rcloud.p;
    a.Dx = L.rows[OP_DX].p; a.Dy = L.rows[OP_DY].p; a.Dz = (dim == 3) ? L.rows[OP_DZ].p : nullptr;
    a.lap = L.rows[OP_LAP].p;
    if (l + 1 < pb->L) { a.rt_off = w.d_roff.p; a.rt_xyz = w.d_rxyz.p; a.rt_row = w.d_rrow.p; a.R_out = pb->lv[l + 1].Rrows.p; }
    if (l > 0) { a.pt_off = w.d_poff.p; a.pt_xyz = w.d_pxyz.p; a.pt_row = w.d_prow.p; a.P_out = pb->lv[l - 1].Prows.p; }
    a.ws = w.ws.p; a.rhs = w.rhs.p; a.piv = w.piv.p; a.batch = (int)batch; a.status = d_status;
    int most = 0;
    for (int i = 0; i < N; i++) most = std::max(most, (roff[i + 1] - roff[i]) + (poff[i + 1] - poff[i]));
    a.rhs_max = 3 + (dim == 3 ? 1 : 0) + most;
    w.has_wide = L.mb > 0 && !L.bnodes.empty();
    if (w.has_wide) {
        const int nbw = (int)L.bnodes.size(), mb = L.mb;
        std::vector<int> rows(nbw);
        for (int t = 0; t < nbw; t++) rows[t] = t;
        HIP_TRY(w.d_rows.upload(rows, g_stream));
        for (int o = 0; o < 4; o++) {
            if (o == OP_DZ && dim != 3) continue;
            HIP_TRY(L.rows_b[o].alloc((size_t)nbw * mb));
            HIP_TRY(hipMemsetAsync(L.rows_b[o].p, 0, sizeof(double) * nbw * mb, g_stream));
        }
        WeightsArgs &b = w.wide;
        b = a;
        b.nsys = nbw; b.sys = w.d_rows.p; b.m = mb; b.cloud = L.rcloud_b.p;
        b.Dx = L.rows_b[OP_DX].p; b.Dy = L.rows_b[OP_DY].p; b.Dz = (dim == 3) ? L.rows_b[OP_DZ].p : nullptr;
        b.lap = L.rows_b[OP_LAP].p;
        b.rt_off = nullptr; b.rt_xyz = nullptr; b.rt_row = nullptr; b.R_out = nullptr;
        b.pt_off = nullptr; b.pt_xyz = nullptr; b.pt_row = nullptr; b.P_out = nullptr;
        b.ws = nullptr; b.rhs = nullptr; b.piv = nullptr;
        b.rhs_max = 3 + (dim == 3 ? 1 : 0);
    }
    return 0;
}

// Weights and transfer rows of levels l0 .. l1-1: one persistent launch over
// all their systems (uneven sizes included) when every system fits the
// register kernels (at most 128 unknowns), else one launch per level.
static int build_weights(mgm_problem *pb, int l0, int l1, double *kernel_seconds)
{
    std::vector<WeightsPrep> prep(l1 - l0);
    DevBuf<int> d_status;
    HIP_TRY(d_status.alloc(1));
    HIP_TRY(hipMemsetAsync(d_status.p, 0, sizeof(int), g_stream));
    bool batch = true;
    int rc;
    for (int l = l0; l < l1; l++) {
        if ((rc = prepare_weights(pb, l, prep[l - l0], d_status.p))) return rc;
        const int mpn = pb->lv[l].m + pb->lv[l].n, mpnb = pb->lv[l].mb + pb->lv[l].n;
        batch = batch && (weights_tile(mpn) || lu32_tiles(mpn)) && (!pb->lv[l].mb || weights_tile(mpnb) || lu32_tiles(mpnb));
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0)); HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, g_stream));
    if (batch) {
        // one launch per register-tile class: every level's systems (and the
        // wide boundary systems where their tile shape is the same) together
        std::map<int, std::vector<WeightsArgs>> cls;
        auto klass = [](const WeightsArgs &x) {
            const int mpn = x.m + x.n, t = weights_tile(mpn);
            return t ? t : 100 + lu32_tiles(mpn);
        };
        for (auto &w : prep) {
            cls[klass(w.a)].push_back(w.a);
            if (w.has_wide) cls[klass(w.wide)].push_back(w.wide);
        }
        rc = 0;
        for (auto &c : cls)
            if (!rc) rc = launch_weights_levels(c.second, g_stream);
    } else {
        rc = 0;
        for (auto &w : prep) {
            if (!rc) rc = launch_weights(w.a, g_stream);
            if (!rc && w.has_wide) rc = launch_weights(w.wide, g_stream);
        }
    }
    if (rc) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return rc; }
    HIP_TRY(hipEventRecord(e1, g_stream));
    int status = 0;
    HIP_TRY(hipMemcpyAsync(&status, d_status.p, sizeof(int), hipMemcpyDeviceToHost, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    if (kernel_seconds) *kernel_seconds = ms * 1e-3;
    if (status) MGM_FAIL(-8, "weights: singular local system (zero pivot)");
    return 0;
}

static int level_lap_poisson(mgm_problem *pb, int l)
{
    Level &L = pb->lv[l];
    int dim = pb->cfg.dimension;
    return launch_lap_poisson(L.N, L.m, dim, L.rows[OP_DX].p, L.rows[OP_DY].p,
                              dim == 3 ? L.rows[OP_DZ].p : nullptr, L.rows[OP_LAP].p, L.rnx.p,
                              L.rny.p, L.rnz.p, L.rbtag.p, L.rows[OP_LAPP].p, g_stream);
}

// Streaming passes over slabs of more than MGM_GATHER_SORT_ROWS rows
// (default 262,144) read them gather-sorted (EllT::sorted): such a pass
// streams its slab from HBM, and the sorted order cuts the cache lines one
// gather instruction touches.  MGM_GATHER_SORT=0 keeps every slab plain.
// Results are identical either way.  Both are read whenever a level's slabs
// are built.
static bool gather_sort_on(int rows, int S, int m, int ncols)
{
    const char *e = getenv("MGM_GATHER_SORT");
    if (e && atoi(e) == 0) return false;
    const char *t = getenv("MGM_GATHER_SORT_ROWS");
    const long min_rows = t ? atol(t) : kLargeRows;
    return rows > min_rows && slab_sortable(S, m, ncols);
}

static int level_slab(mgm_problem *pb, int l, int which)
{
    Level &L = pb->lv[l];
    pb->mp_built = 0;  // narrowed copies no longer match
    pb->mp_ns_built = false;
    if (!L.rows[which].p) return 0;
    size_t sz = ell_size(L.Nd, L.m, L.S);
    if (L.slab[which].n != sz) {
        HIP_TRY(L.slab[which].alloc(sz));
        HIP_TRY(hipMemsetAsync(L.slab[which].p, 0, sizeof(double) * sz, g_stream));
    }
    int rc = launch_gather_rows(L.N, L.m, L.S, L.rows[which].p, L.d_iperm.p, L.rcloud.p,
                                L.slab[which].p, nullptr, g_stream);
    if (rc || !L.gs || (which != OP_LAP && which != OP_LAPP)) return rc;
    if (L.gslab[which].n != sz) HIP_TRY(L.gslab[which].alloc(sz));
    return launch_permute_slab((long)sz, L.m, L.S, L.gcol.p, L.slab[which].p, L.gslab[which].p, g_stream);
}

// Full passes over a level in Morton blocks read x from an LDS window per
// block (launch_rows' window form) unless MGM_WINDOW=0, when the block's
// window -- its own rows and the halo rows its stencils reach -- fits next to
// the staging of kWindowWaves waves (FP64 values).  1M-point cube, level 0:
// 244 blocks, at most 9,871 window values, 151 KB of LDS with the staging;
// the apply takes 167 us instead of 203 us gathering from memory.
constexpr size_t kWindowLds = 159 * 1024;
static_assert(kWindowLds / sizeof(float) <= 65536, "window places must fit 16 bits");
static bool window_on()
{
    const char *e = getenv("MGM_WINDOW");
    return !(e && e[0] == '0');
}

// The window tables of level l: per Morton block the sorted halo rows, and
// the plain slab's columns mapped into the block's window.
static int level_window(Level &L)
{
    L.win_blocks = 0;
    L.win_max = 0;
    if (!L.blocked || !window_on() || L.blk_row0.size() < 2) return 0;
    const int nb = (int)L.blk_row0.size() - 1, m = L.m, S = L.S;
    std::vector<std::vector<int>> halo(nb);
    std::vector<uint16_t> wcol(ell_size(L.Nd, m, S), 0);
    std::vector<int> wsize(nb, 0);
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            for (int b = t; b < nb; b += nt) {
                const int r0 = L.blk_row0[b], r1 = L.blk_row0[b + 1];
                std::vector<int> &h = halo[b];
                for (int d = r0; d < r1; d++) {
                    const int i = L.perm[d];
                    if (i < 0) continue;
                    for (int j = 0; j < m; j++) {
                        const int c = L.iperm[L.cloud[(size_t)i * m + j]];
                        if (c < r0 || c >= r1) h.push_back(c);
                    }
                }
                std::sort(h.begin(), h.end());
                h.erase(std::unique(h.begin(), h.end()), h.end());
                for (int d = r0; d < r1; d++) {
                    const int i = L.perm[d];
                    for (int j = 0; j < m; j++) {
                        int local = 0;   // padding rows: weight 0, any valid place
                        if (i >= 0) {
                            const int c = L.iperm[L.cloud[(size_t)i * m + j]];
                            local = (c >= r0 && c < r1)
                                        ? c - r0
                                        : (r1 - r0) + (int)(std::lower_bound(h.begin(), h.end(), c) - h.begin());
                        }
                        wcol[ell_at(m, S, d, j)] = (uint16_t)local;
                    }
                }
                wsize[b] = (r1 - r0) + (int)h.size();
            }
        });
    for (auto &x : th) x.join();
    const int wmax = *std::max_element(wsize.begin(), wsize.end());
    if ((size_t)kWindowWaves * prod_stride(m) * S * 8 + (size_t)wmax * 8 > kWindowLds) return 0;
    std::vector<int> hoff(nb + 1, 0), hall;
    for (int b = 0; b < nb; b++) {
        hoff[b + 1] = hoff[b] + (int)halo[b].size();
        hall.insert(hall.end(), halo[b].begin(), halo[b].end());
    }
    if (hall.empty()) hall.push_back(0);
    HIP_TRY(L.win_col.upload(wcol, g_stream));
    HIP_TRY(L.win_row0.upload(L.blk_row0, g_stream));
    HIP_TRY(L.win_hoff.upload(hoff, g_stream));
    HIP_TRY(L.win_halo.upload(hall, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));   // host tables die here
    L.win_blocks = nb;
    L.win_max = wmax;
    return 0;
}

bool prol_window_on()
{
    const char *e = getenv("MGM_WINDOW_PROLONG");
    return window_on() && !(e && e[0] == '0');
}

// List windows of a transfer operator whose rows are level L's device rows
// (Morton blocks with windows) and whose columns index another level: per
// block the sorted, distinct columns its rows reference (col(i, j) for the
// block's real rows, i = L.perm[d]) and the plain slab's columns (width mt,
// slice height St) as 16-bit places in that list (padding rows: place 0).
// Returns the largest list, 0 when the windows do not fit or blocks are not
// whole slices of St.
template <typename ColFn>
static int level_list_window(const Level &L, int mt, int St, ColFn col, DevBuf<uint16_t> &wcol_d,
                             DevBuf<int> &hoff_d, DevBuf<int> &list_d, int *wmax_out)
{
    *wmax_out = 0;
    if (!L.win_blocks || !window_on()) return 0;
    const int nb = L.win_blocks;
    // the window pass walks slices of St rows from block start to block end
    for (int b = 0; b <= nb; b++)
        if (L.blk_row0[b] % St) return 0;
    std::vector<std::vector<int>> lists(nb);
    std::vector<uint16_t> wcol(ell_size(L.Nd, mt, St), 0);
    std::vector<int> wsize(nb, 0);
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            for (int b = t; b < nb; b += nt) {
                const int r0 = L.blk_row0[b], r1 = L.blk_row0[b + 1];
                std::vector<int> &h = lists[b];
                for (int d = r0; d < r1; d++) {
                    const int i = L.perm[d];
                    if (i < 0) continue;
                    for (int j = 0; j < mt; j++) h.push_back(col(i, j));
                }
                std::sort(h.begin(), h.end());
                h.erase(std::unique(h.begin(), h.end()), h.end());
                wsize[b] = (int)h.size();
                if (h.size() > 65535) continue;
                for (int d = r0; d < r1; d++) {
                    const int i = L.perm[d];
                    if (i < 0) continue;
                    for (int j = 0; j < mt; j++)
                        wcol[ell_at(mt, St, d, j)] =
                            (uint16_t)(std::lower_bound(h.begin(), h.end(), col(i, j)) - h.begin());
                }
            }
        });
    for (auto &x : th) x.join();
    const int wmax = *std::max_element(wsize.begin(), wsize.end());
    if (wmax > 65535 || (size_t)kWindowWaves * prod_stride(mt) * St * 8 + (size_t)wmax * 8 > kWindowLds) return 0;
    std::vector<int> hoff(nb + 1, 0), hall;
    for (int b = 0; b < nb; b++) {
        hoff[b + 1] = hoff[b] + (int)lists[b].size();
        hall.insert(hall.end(), lists[b].begin(), lists[b].end());
    }
    if (hall.empty()) hall.push_back(0);
    HIP_TRY(wcol_d.upload(wcol, g_stream));
    HIP_TRY(hoff_d.upload(hoff, g_stream));
    HIP_TRY(list_d.upload(hall, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));   // host tables die here
    *wmax_out = std::max(wmax, 1);
    return 0;
}

// The prolongation into fine level L from coarse level C: row d of L reads
// C.iperm[C.cloud[L.pp[i] * mc + j]] (as the transfer slab composes it).
static int level_prol_window(Level &L, const Level &C)
{
    L.pwin_max = 0;
    if (!prol_window_on()) return 0;
    const int mc = L.mp;
    return level_list_window(L, mc, L.Sp, [&](int i, int j) { return C.iperm[C.cloud[(size_t)L.pp[i] * mc + j]]; },
                             L.pwin_col, L.pwin_hoff, L.pwin_halo, &L.pwin_max);
}

// The restriction into coarse level C from fine level F: row d of C reads
// F.iperm[F.cloud[C.rp[i] * mf + j]].
static int level_restr_window(Level &C, const Level &F)
{
    C.rwin_max = 0;
    if (!prol_window_on()) return 0;
    const int mf = C.mr;
    return level_list_window(C, mf, C.Sr, [&](int i, int j) { return F.iperm[F.cloud[(size_t)C.rp[i] * mf + j]]; },
                             C.rwin_col, C.rwin_hoff, C.rwin_halo, &C.rwin_max);
}

// (Re)build every slab of level l: stencil columns, the five operators and
// the transfer rows living on this level.  Padding rows hold weight 0 and
// column 0.
static int level_slabs(mgm_problem *pb, int l)
{
    Level &L = pb->lv[l];
    pb->mp_built = 0;
    pb->mp_ns_built = false;
    size_t sz = ell_size(L.Nd, L.m, L.S);
    if (L.col.n != sz) {
        HIP_TRY(L.col.alloc(sz));
        HIP_TRY(hipMemsetAsync(L.col.p, 0, sizeof(int) * sz, g_stream));
    }
    int rc = launch_gather_rows(L.N, L.m, L.S, L.rows[OP_LAP].p, L.d_iperm.p, L.rcloud.p,
                                nullptr, L.col.p, g_stream);
    if (rc) return rc;
    if (!L.win_tried) {
        L.win_tried = true;
        if ((rc = level_window(L))) return rc;
    }
    L.gs = gather_sort_on(L.Nd, L.S, L.m, L.Nd);
    if (L.gs) {
        if (L.gcol.n != sz) HIP_TRY(L.gcol.alloc(sz));
        if ((rc = launch_gather_sort((int)(sz / ((size_t)L.S * L.m)), L.m, L.S, L.col.p, L.gcol.p, g_stream)))
            return rc;
    }
    for (int o = 0; o < NUM_OPS; o++)
        if ((rc = level_slab(pb, l, o))) return rc;
    if (l > 0) {
        Level &F = pb->lv[l - 1];
        L.mr = F.m;
        L.Sr = slice_height(L.mr);
        size_t n = ell_size(L.Nd, L.mr, L.Sr);
        if (L.rcol.n != n) {
            HIP_TRY(L.rcol.alloc(n)); HIP_TRY(L.rw.alloc(n));
            HIP_TRY(hipMemsetAsync(L.rcol.p, 0, sizeof(int) * n, g_stream));
            HIP_TRY(hipMemsetAsync(L.rw.p, 0, sizeof(double) * n, g_stream));
        }
        rc = launch_transfer_slab(L.N, L.mr, L.Sr, L.Rrows.p, L.d_iperm.p, L.d_rp.p, F.rcloud.p, F.d_iperm.p,
                                  L.rw.p, L.rcol.p, g_stream);
        if (rc) return rc;
        if (!L.rwin_tried) {
            L.rwin_tried = true;
            if ((rc = level_restr_window(L, F))) return rc;
        }
    }
    if (l + 1 < pb->L) {
        Level &C = pb->lv[l + 1];
        L.mp = C.m;
        L.Sp = slice_height(L.mp);
        size_t n = ell_size(L.Nd, L.mp, L.Sp);
        if (L.pcol.n != n) {
            HIP_TRY(L.pcol.alloc(n)); HIP_TRY(L.pw.alloc(n));
            HIP_TRY(hipMemsetAsync(L.pcol.p, 0, sizeof(int) * n, g_stream));
            HIP_TRY(hipMemsetAsync(L.pw.p, 0, sizeof(double) * n, g_stream));
        }
        rc = launch_transfer_slab(L.N, L.mp, L.Sp, L.Prows.p, L.d_iperm.p, L.d_pp.p, C.rcloud.p, C.d_iperm.p,
                                  L.pw.p, L.pcol.p, g_stream);
        if (rc) return rc;
        L.gsp = gather_sort_on(L.Nd, L.Sp, L.mp, C.Nd);
        if (L.gsp) {
            if (L.gpcol.n != n) { HIP_TRY(L.gpcol.alloc(n)); HIP_TRY(L.gpw.alloc(n)); }
            if ((rc = launch_gather_sort((int)(n / ((size_t)L.Sp * L.mp)), L.mp, L.Sp, L.pcol.p, L.gpcol.p,
                                         g_stream)))
                return rc;
            if ((rc = launch_permute_slab((long)n, L.mp, L.Sp, L.gpcol.p, L.pw.p, L.gpw.p, g_stream))) return rc;
        }
        if (!L.pwin_tried) {
            L.pwin_tried = true;
            if ((rc = level_prol_window(L, C))) return rc;
        }
    }
    return 0;
}

}  // namespace mgm

mgm_problem::~mgm_problem()
{
    if (host_scal) (void)hipHostFree(host_scal);
}

// ===========================================================================
// extern "C" API
// ===========================================================================
extern "C" {

const char *mgm_last_error(void) { return mgm::last_error(); }

int mgm_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int mgm_init(int device)
{
    if (g_stream && device == g_device) return 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) MGM_FAIL(-9, "no GPU visible to HIP");
    if (device < 0 || device >= n) MGM_FAIL(-1, "mgm_init: bad device index");
    HIP_TRY(hipSetDevice(device));
    int rc = noalloc_init(device);   // the allocation-free entry point's scratch address
    if (rc) return rc;
    if ((rc = stencil_device_init())) return rc;
    if ((rc = mlp_device_init())) return rc;
    if ((rc = knn_device_init())) return rc;
    if (g_stream) (void)hipStreamDestroy(g_stream);
    HIP_TRY(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    g_device = device;
    return 0;
}

void mgm_config_default(mgm_config *c)
{
    c->dimension = 2;
    c->poly_degree = 3;
    c->coarse_poly_degree = 3;
    c->phs_degree = 3;
    c->cloud_size_multiplier = 2.0;
    c->omega = 1.0f;
    c->num_relax = 3;
    c->transfer_rows = MGM_ROWS_REFERENCE;
    c->sweep_order = MGM_ORDER_COLOUR;
    c->boundary_cloud_size_multiplier = 0.0;
}

int mgm_poly_terms(int dim, int deg, short *px, short *py, short *pz)
{
    return mgm::poly_terms(dim, deg, px, py, pz);
}

uint64_t mgm_splitmix64(uint64_t *state) { return mgm::splitmix64(*state); }

int mgm_points_read_gmsh(const char *path, int dim, mgm_points **out)
{
    if (!path || !out) MGM_FAIL(-1, "null argument");
    if (dim != 2 && dim != 3) MGM_FAIL(-1, "dimension must be 2 or 3");
    auto *p = new mgm_points();
    int rc = mgm::read_gmsh(path, dim, p->ps);
    if (rc) { delete p; return rc; }
    *out = p;
    return 0;
}

int mgm_points_generate(int dim, int n, uint64_t seed, double jitter, int with_volume_elements,
                        const char *gmsh_path, mgm_points **out)
{
    if (!out) MGM_FAIL(-1, "null argument");
    auto *p = new mgm_points();
    int rc = mgm::generate_cloud(dim, n, seed, jitter, with_volume_elements, p->ps, gmsh_path);
    if (rc) { delete p; return rc; }
    *out = p;
    return 0;
}

int mgm_points_from_arrays(int N, int nb, int num_elem, double d_avg, const double *x,
                           const double *y, const double *z, const double *nx, const double *ny,
                           const double *nz, const uint8_t *btag, const uint8_t *ctag,
                           mgm_points **out)
{
    if (!out || !x || !y || !z || !nx || !ny || !nz || !btag || !ctag) MGM_FAIL(-1, "null argument");
    if (N <= 0 || nb < 0 || nb > N) MGM_FAIL(-1, "bad sizes");
    auto *p = new mgm_points();
    auto &s = p->ps;
    s.N = N; s.nb = nb; s.num_elem = num_elem; s.d_avg = d_avg;
    s.x.assign(x, x + N); s.y.assign(y, y + N); s.z.assign(z, z + N);
    s.nx.assign(nx, nx + N); s.ny.assign(ny, ny + N); s.nz.assign(nz, nz + N);
    s.btag.assign(btag, btag + N); s.ctag.assign(ctag, ctag + N);
    *out = p;
    return 0;
}

int mgm_points_info(const mgm_points *p, int *N, int *nb, int *num_elem, double *d_avg)
{
    if (!p) MGM_FAIL(-1, "null points");
    if (N) *N = p->ps.N;
    if (nb) *nb = p->ps.nb;
    if (num_elem) *num_elem = p->ps.num_elem;
    if (d_avg) *d_avg = p->ps.d_avg;
    return 0;
}

int mgm_points_copy(const mgm_points *p, double *x, double *y, double *z, double *nx, double *ny,
                    double *nz, uint8_t *btag, uint8_t *ctag)
{
    if (!p) MGM_FAIL(-1, "null points");
    const auto &s = p->ps;
    size_t N = s.N;
    if (x) memcpy(x, s.x.data(), 8 * N);
    if (y) memcpy(y, s.y.data(), 8 * N);
    if (z) memcpy(z, s.z.data(), 8 * N);
    if (nx) memcpy(nx, s.nx.data(), 8 * N);
    if (ny) memcpy(ny, s.ny.data(), 8 * N);
    if (nz) memcpy(nz, s.nz.data(), 8 * N);
    if (btag) memcpy(btag, s.btag.data(), N);
    if (ctag) memcpy(ctag, s.ctag.data(), N);
    return 0;
}

void mgm_points_free(mgm_points *p) { delete p; }

int mgm_problem_create(const mgm_config *cfg, int num_levels, mgm_problem **out)
{
    if (!cfg || !out) MGM_FAIL(-1, "null argument");
    if (num_levels < 1 || num_levels > 16) MGM_FAIL(-1, "num_levels must be in [1,16]");
    if (cfg->dimension != 2 && cfg->dimension != 3) MGM_FAIL(-1, "dimension must be 2 or 3");
    if (cfg->phs_degree < 1 || cfg->phs_degree > 15) MGM_FAIL(-1, "phs_degree out of range");
    if (cfg->poly_degree < 1 || cfg->coarse_poly_degree < 1) MGM_FAIL(-1, "poly degree must be >= 1");
    auto *p = new mgm_problem();
    p->cfg = *cfg;
    p->L = num_levels;
    p->lv.resize(num_levels);
    *out = p;
    return 0;
}

void mgm_problem_destroy(mgm_problem *p) { delete p; }

int mgm_problem_set_level(mgm_problem *p, int lev, const mgm_points *pts)
{
    if (!p || !pts) MGM_FAIL(-1, "null argument");
    if (lev < 0 || lev >= p->L) MGM_FAIL(-1, "level out of range");
    p->lv[lev].in = pts->ps;
    p->lv[lev].have_input = true;
    p->ready = false;
    return 0;
}

int mgm_problem_setup(mgm_problem *pb)
{
    if (!pb) MGM_FAIL(-1, "null problem");
    int rc = ensure_init();
    if (rc) return rc;
    const int dim = pb->cfg.dimension;
    double t0 = now(), tk = 0, tr = 0, tp = 0, tw = 0, ts = 0;
    for (int l = 0; l < pb->L; l++) {
        Level &L = pb->lv[l];
        if (!L.have_input) MGM_FAIL(-1, "level without points");
        L.poly = (l == 0) ? pb->cfg.poly_degree : pb->cfg.coarse_poly_degree;
        L.n = poly_terms(dim, L.poly, nullptr, nullptr, nullptr);
        L.px.resize(L.n); L.py.resize(L.n); L.pz.resize(L.n);
        poly_terms(dim, L.poly, L.px.data(), L.py.data(), L.pz.data());
        L.m = (int)floor(pb->cfg.cloud_size_multiplier * L.n + 1e-9);
        L.N = L.in.N; L.nb = L.in.nb; L.d_avg_bbox = L.in.d_avg;
        if (L.m < 2 || L.m > 127) MGM_FAIL(-1, "stencil size must be in [2,127]");
        if (L.m + L.n > 256) MGM_FAIL(-1, "local system larger than 256");
        L.mb = 0;
        if (pb->cfg.boundary_cloud_size_multiplier > 0) {
            const int mb = (int)floor(pb->cfg.boundary_cloud_size_multiplier * L.n + 1e-9);
            if (mb > L.m) {
                if (mb + L.n > 128) MGM_FAIL(-1, "wide boundary stencils: m_b + n must not exceed 128");
                L.mb = mb;
            }
        }
        if (L.N < L.m + 1) MGM_FAIL(-1, "level has fewer nodes than the stencil size");
        double a = now();
        if ((rc = level_knn(L, dim))) return rc;
        double b = now();
        if ((rc = level_bfs(L))) return rc;
        level_normals_spacing(L, dim);
        tk += b - a;
        tr += now() - b;
    }
    double a = now();
    for (int l = 0; l < pb->L; l++) {
        Level &L = pb->lv[l];
        if ((rc = nearest_points(L, (l == pb->L - 1) ? L : pb->lv[l + 1], dim, L.pp))) return rc;
        if ((rc = nearest_points(L, (l == 0) ? L : pb->lv[l - 1], dim, L.rp))) return rc;
    }
    tp = now() - a;
    a = now();
    for (int l = 0; l < pb->L; l++) {
        level_colour_order(pb->lv[l], dim);
        if ((rc = upload_level(pb->lv[l]))) return rc;
    }
    tr += now() - a;
    a = now();
    if ((rc = build_weights(pb, 0, pb->L, nullptr))) return rc;
    for (int l = 0; l < pb->L; l++)
        if ((rc = level_lap_poisson(pb, l))) return rc;
    HIP_TRY(hipStreamSynchronize(g_stream));
    tw = now() - a;
    a = now();
    for (int l = 0; l < pb->L; l++)
        if ((rc = level_slabs(pb, l))) return rc;
    HIP_TRY(pb->partial.alloc(1024));
    HIP_TRY(pb->scal.alloc(16));
    if (!pb->host_scal) HIP_TRY(hipHostMalloc((void **)&pb->host_scal, 16 * sizeof(double), 0));
    HIP_TRY(hipStreamSynchronize(g_stream));
    ts = now() - a;
    pb->timings[0] = tk; pb->timings[1] = tr; pb->timings[2] = tp; pb->timings[3] = tw;
    pb->timings[4] = ts; pb->timings[5] = now() - t0;
    pb->ready = true;
    return 0;
}

#define CHECK_READY(pb, lev)                                                      \
    do {                                                                          \
        if (!(pb)) MGM_FAIL(-1, "null problem");                                  \
        if (!(pb)->ready) MGM_FAIL(-1, "problem is not set up");                  \
        if ((lev) < 0 || (lev) >= (pb)->L) MGM_FAIL(-1, "level out of range");    \
    } while (0)

int mgm_level_info_get(const mgm_problem *pb, int lev, mgm_level_info *o)
{
    CHECK_READY(pb, lev);
    const Level &L = pb->lv[lev];
    o->N = L.N; o->nb = L.nb; o->m = L.m; o->n = L.n; o->poly_degree = L.poly;
    o->num_colours = L.ncol; o->d_avg = L.d_avg; o->d_avg_bbox = L.d_avg_bbox; o->knn_ties = L.knn_ties;
    o->device_rows = L.Nd; o->slice_height = L.S; o->blocked = L.blocked; o->gather_sorted = L.gs;
    o->small_relax = relax_small_here(pb, L);
    o->window_blocks = L.win_blocks;
    o->window_max = L.win_max;
    o->prolong_window_max = L.pwin_max;
    return 0;
}

int mgm_setup_timings(const mgm_problem *pb, double *t6)
{
    if (!pb) MGM_FAIL(-1, "null problem");
    for (int i = 0; i < 6; i++) t6[i] = pb->timings[i];
    return 0;
}

int mgm_rebuild_weights(mgm_problem *pb, int lev, double *seconds, long *systems)
{
    CHECK_READY(pb, lev);
    int rc = build_weights(pb, lev, lev + 1, seconds);
    if (rc) return rc;
    if ((rc = level_lap_poisson(pb, lev))) return rc;
    for (int l = std::max(0, lev - 1); l <= std::min(pb->L - 1, lev + 1); l++)
        if ((rc = level_slabs(pb, l))) return rc;
    HIP_TRY(hipStreamSynchronize(g_stream));
    if (systems) {
        long s = 0;
        for (int i = 0; i < pb->lv[lev].N; i++) s += !pb->lv[lev].ctag[i];
        if (pb->lv[lev].mb) s += (long)pb->lv[lev].bnodes.size();   // the wide boundary systems
        *systems = s;
    }
    return 0;
}

int mgm_weight_kernel_choice(int mpn, int *tile, int *waves_per_simd)
{
    if (!tile || !waves_per_simd) MGM_FAIL(-1, "null argument");
    return weights_kernel_choice(mpn, tile, waves_per_simd);
}

int mgm_get_geometry(const mgm_problem *pb, int lev, double *x, double *y, double *z, double *nx,
                     double *ny, double *nz, uint8_t *btag)
{
    CHECK_READY(pb, lev);
    const Level &L = pb->lv[lev];
    size_t N = L.N;
    if (x) memcpy(x, L.x.data(), 8 * N);
    if (y) memcpy(y, L.y.data(), 8 * N);
    if (z) memcpy(z, L.z.data(), 8 * N);
    if (nx) memcpy(nx, L.nx.data(), 8 * N);
    if (ny) memcpy(ny, L.ny.data(), 8 * N);
    if (nz) memcpy(nz, L.nz.data(), 8 * N);
    if (btag) memcpy(btag, L.btag.data(), N);
    return 0;
}

int mgm_get_cloud(const mgm_problem *pb, int lev, int32_t *cloud)
{
    CHECK_READY(pb, lev);
    memcpy(cloud, pb->lv[lev].cloud.data(), sizeof(int) * pb->lv[lev].cloud.size());
    return 0;
}

int mgm_get_points(const mgm_problem *pb, int lev, int which, int32_t *out)
{
    CHECK_READY(pb, lev);
    const auto &v = (which == MGM_PTS_RESTRICTION) ? pb->lv[lev].rp : pb->lv[lev].pp;
    memcpy(out, v.data(), sizeof(int) * v.size());
    return 0;
}

int mgm_get_colours(const mgm_problem *pb, int lev, int32_t *colour)
{
    CHECK_READY(pb, lev);
    memcpy(colour, pb->lv[lev].colour.data(), sizeof(int) * pb->lv[lev].N);
    return 0;
}

static int matrix_buf(const mgm_problem *pb, int lev, int which, double **buf, size_t *count, int *slab_op)
{
    Level &L = const_cast<Level &>(pb->lv[lev]);
    *slab_op = -1;
    if (which >= MGM_MAT_DX && which <= MGM_MAT_LAPP) {
        if (which == MGM_MAT_DZ && pb->cfg.dimension != 3) MGM_FAIL(-1, "Dz exists only in 3D");
        *buf = L.rows[which].p;
        *count = (size_t)L.N * L.m;
        *slab_op = which;
    } else if (which == MGM_MAT_RESTR) {
        if (lev == 0) MGM_FAIL(-1, "level 0 has no restriction rows");
        *buf = L.Rrows.p;
        *count = (size_t)L.N * pb->lv[lev - 1].m;
    } else if (which == MGM_MAT_PROL) {
        if (lev == pb->L - 1) MGM_FAIL(-1, "the coarsest level has no prolongation rows");
        *buf = L.Prows.p;
        *count = (size_t)L.N * pb->lv[lev + 1].m;
    } else {
        MGM_FAIL(-1, "unknown matrix id");
    }
    return 0;
}

int mgm_get_matrix(const mgm_problem *pb, int lev, int which, double *out)
{
    CHECK_READY(pb, lev);
    double *buf; size_t cnt; int op;
    int rc = matrix_buf(pb, lev, which, &buf, &cnt, &op);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(out, buf, 8 * cnt, hipMemcpyDeviceToHost, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    return 0;
}

int mgm_boundary_info(const mgm_problem *pb, int lev, int *nbw, int *mb)
{
    CHECK_READY(pb, lev);
    if (!nbw || !mb) MGM_FAIL(-1, "null argument");
    const Level &L = pb->lv[lev];
    *mb = L.mb;
    *nbw = L.mb ? (int)L.bnodes.size() : 0;
    return 0;
}

int mgm_get_boundary_rows(const mgm_problem *pb, int lev, int which, int *nodes, int *cloud, double *w)
{
    CHECK_READY(pb, lev);
    const Level &L = pb->lv[lev];
    if (!L.mb) MGM_FAIL(-1, "boundary_rows: this level has no wide boundary stencils");
    if (which < MGM_MAT_DX || which > MGM_MAT_LAP || (which == MGM_MAT_DZ && pb->cfg.dimension != 3))
        MGM_FAIL(-1, "boundary_rows: Dx, Dy, Dz (3D) or lap only");
    const size_t nbw = L.bnodes.size();
    if (nodes) std::copy(L.bnodes.begin(), L.bnodes.end(), nodes);
    if (cloud) std::copy(L.cloud_b.begin(), L.cloud_b.end(), cloud);
    if (w) {
        HIP_TRY(hipMemcpyAsync(w, L.rows_b[which].p, sizeof(double) * nbw * L.mb, hipMemcpyDeviceToHost, g_stream));
        HIP_TRY(hipStreamSynchronize(g_stream));
    }
    return 0;
}

int mgm_set_matrix(mgm_problem *pb, int lev, int which, const double *in)
{
    CHECK_READY(pb, lev);
    double *buf; size_t cnt; int op;
    int rc = matrix_buf(pb, lev, which, &buf, &cnt, &op);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(buf, in, 8 * cnt, hipMemcpyHostToDevice, g_stream));
    if (op >= 0) rc = level_slab(pb, lev, op);
    else rc = level_slabs(pb, lev);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(g_stream));
    return 0;
}

int mgm_dirichlet_rows(mgm_problem *pb)
{
    CHECK_READY(pb, 0);
    for (int l = 0; l < pb->L; l++) {
        Level &L = pb->lv[l];
        int rc = launch_dirichlet_rows(L.N, L.m, L.rbtag.p, L.rows[OP_LAPP].p, g_stream);
        if (rc) return rc;
        if ((rc = level_slab(pb, l, OP_LAPP))) return rc;
    }
    HIP_TRY(hipStreamSynchronize(g_stream));
    return 0;
}

int mgm_get_field(const mgm_problem *pb, int lev, int which, double *out)
{
    CHECK_READY(pb, lev);
    if (which < 0 || which >= MGM_NUM_FIELDS) MGM_FAIL(-1, "unknown field");
    const Level &L = pb->lv[lev];
    std::vector<double> tmp(L.Nd);
    HIP_TRY(hipMemcpyAsync(tmp.data(), L.f[which].p, 8 * (size_t)L.Nd, hipMemcpyDeviceToHost, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    for (int i = 0; i < L.N; i++) out[i] = tmp[L.iperm[i]];
    return 0;
}

int mgm_set_field(mgm_problem *pb, int lev, int which, const double *in)
{
    CHECK_READY(pb, lev);
    if (which < 0 || which >= MGM_NUM_FIELDS) MGM_FAIL(-1, "unknown field");
    Level &L = pb->lv[lev];
    std::vector<double> tmp(L.Nd);
    for (int d = 0; d < L.Nd; d++) tmp[d] = L.perm[d] < 0 ? 0.0 : in[L.perm[d]];
    HIP_TRY(hipMemcpyAsync(L.f[which].p, tmp.data(), 8 * (size_t)L.Nd, hipMemcpyHostToDevice, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    return 0;
}

int mgm_set_smoother(mgm_problem *pb, float omega, int num_relax)
{
    if (!pb) MGM_FAIL(-1, "null problem");
    if (!(omega > 0.0f && omega < 2.0f)) MGM_FAIL(-1, "omega must be in (0, 2)");
    if (num_relax < 0 || num_relax > 64) MGM_FAIL(-1, "num_relax must be in [0, 64]");
    pb->cfg.omega = omega;
    pb->cfg.num_relax = num_relax;
    pb->omega_lv.clear();
    return 0;
}

int mgm_set_smoother_levels(mgm_problem *pb, const float *omega, int n)
{
    if (!pb) MGM_FAIL(-1, "null problem");
    if (n != 0 && n != pb->L) MGM_FAIL(-1, "set_smoother_levels: n must be 0 or the number of levels");
    if (n && !omega) MGM_FAIL(-1, "set_smoother_levels: null omega");
    for (int l = 0; l < n; l++)
        if (!(omega[l] > 0.0f && omega[l] < 2.0f)) MGM_FAIL(-1, "omega must be in (0, 2)");
    pb->omega_lv.assign(omega, omega + n);
    return 0;
}

int mgm_get_smoother_levels(const mgm_problem *pb, float *omega)
{
    if (!pb || !omega) MGM_FAIL(-1, "null argument");
    for (int l = 0; l < pb->L; l++) omega[l] = pb->omega(l);
    return 0;
}

int mgm_set_smoother_kinds(mgm_problem *pb, const int *kind, int n)
{
    if (!pb) MGM_FAIL(-1, "null problem");
    if (n != 0 && n != pb->L) MGM_FAIL(-1, "set_smoother_kinds: n must be 0 or the number of levels");
    if (n && !kind) MGM_FAIL(-1, "set_smoother_kinds: null kind");
    for (int l = 0; l < n; l++)
        if (kind[l] != MGM_SMOOTH_SOR && kind[l] != MGM_SMOOTH_JACOBI) MGM_FAIL(-1, "unknown smoother kind");
    pb->smoother_lv.assign(kind, kind + n);
    return 0;
}

int mgm_get_smoother_kinds(const mgm_problem *pb, int *kind)
{
    if (!pb || !kind) MGM_FAIL(-1, "null argument");
    for (int l = 0; l < pb->L; l++) kind[l] = l < (int)pb->smoother_lv.size() ? pb->smoother_lv[l] : MGM_SMOOTH_SOR;
    return 0;
}

int mgm_zero_fields(mgm_problem *pb)
{
    CHECK_READY(pb, 0);
    for (auto &L : pb->lv)
        for (auto &f : L.f) HIP_TRY(hipMemsetAsync(f.p, 0, 8 * (size_t)L.Nd, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Operators, stream-ordered (no host synchronisation) so a V-cycle can be
// enqueued back to back or captured into a graph.
namespace mgm {

// Levels with at most this many device rows (and whose p vector fits in LDS)
// relax in one single-workgroup launch; MGM_SMALL_ROWS overrides, 0 disables.
int small_level_rows()
{
    const char *e = getenv("MGM_SMALL_ROWS");
    return e ? atoi(e) : 12000;
}

int op_learned_sweep(mgm_problem *pb, int l);
bool learned_active(const mgm_problem *pb, int l, int variant);

static int pfield(int variant) { return variant == MGM_VAR_T ? MGM_F_PPRIME : MGM_F_P; }

// Output buffers of n Jacobi sweeps (each reads the previous one's output):
// p, res (free until the residual that follows the relaxation) and a third
// buffer t.  The last sweep writes p where that takes no copy: from p (from_p)
// an even count alternates res / p; an odd count alternates, then ends res ->
// t -> p; a start outside p alternates so that the last lands in p.  Only
// from p with n = 1 does the result end outside p (in res; the caller copies).
template <typename T>
T *const *jacobi_targets(int n, bool from_p, T *p, T *res, T *t)
{
    static thread_local T *seq[64];
    for (int k = 0; k < n && k < 64; k++) {
        const int left = n - k;   // sweeps still to write, this one included
        if (from_p && n % 2 == 1 && n >= 3 && left <= 2)
            seq[k] = left == 2 ? t : p;
        else if (from_p)
            seq[k] = (k % 2 == 0) ? res : p;
        else
            seq[k] = (left % 2 == 1) ? p : res;
    }
    return seq;
}
template double *const *jacobi_targets<double>(int, bool, double *, double *, double *);
template float *const *jacobi_targets<float>(int, bool, float *, float *, float *);

int op_relax(mgm_problem *pb, int l, int variant)
{
    Level &L = pb->lv[l];
    const float wl = pb->omega(l);
    const double w = wl;
    const double c = (float)(1 - wl);  // (1-omega) evaluated in float, as the reference
    double *p = L.f[pfield(variant)].p, *src = L.f[MGM_F_SOURCE].p, *res = L.f[MGM_F_RES].p;
    Ell A = L.op(OP_LAPP);
    int nrelax = pb->cfg.num_relax;
    if (learned_active(pb, l, variant) && nrelax > 0) {
        // the learned correction takes the place of the first sweep
        int rc = op_learned_sweep(pb, l);
        if (rc) return rc;
        nrelax--;
    }
    if (pb->jacobi(l, variant)) {
        // weighted Jacobi: sweeps alternate between p and res (free until
        // the residual that follows the relaxation); an odd count ends in
        // res and is copied back.  When the solve's fine residual already
        // formed the first sweep (jac0), the rest alternate so that the last
        // lands in p.
        const double *in = p;
        int it = 0;
        if (l == 0 && pb->jac0_ready && nrelax > 0) {
            in = pb->jac0.p;
            it = 1;
        }
        double *const *seq = jacobi_targets(nrelax - it, in == p, p, res, L.jtmp.p);
        for (int k = 0; it < nrelax; it++, k++) {
            int rc = launch_jacobi(A, L.Nd, L.dbtag.p, src, in, seq[k], w, c, g_stream);
            if (rc) return rc;
            in = seq[k];
        }
        return in == p ? 0 : launch_stream_copy((long)L.Nd * 8 / 16, in, p, g_stream);
    }
    if (relax_small_here(pb, L))
        return nrelax ? launch_relax_small(L.op_plain(OP_LAPP), L.op_plain(OP_LAP), L.d_col_slices.p,
                                           L.d_col_soff.p, L.ncol, L.Nd,
                                           L.dbtag.p, L.ref0.p, src, p, res, w, c, nrelax, variant, g_stream)
                      : 0;
    for (int it = 0; it < nrelax; it++) {
        int rc;
        if (pb->cfg.sweep_order == MGM_ORDER_NATURAL) {
            rc = launch_sor_natural(L.op_plain(OP_LAPP), L.N, L.d_iperm.p, src, p, w, c, g_stream);
            if (rc) return rc;
        } else {
            for (int k = 0; k < L.ncol; k++) {
                const int *list;
                int first, n;
                L.colour_slices(k, &list, &first, &n);
                if ((rc = launch_sor_slices(A, list, first, n, L.dbtag.p, src, p, w, c, g_stream))) return rc;
            }
        }
        if (variant == MGM_VAR_H) continue;
        if ((rc = launch_pin(L.Nd, L.dbtag.p, L.ref0.p, p, pb->scal.p + 8, g_stream))) return rc;
        if (variant == MGM_VAR_T) rc = launch_residual(A, L.Nd, 1, L.dbtag.p, src, p, res, g_stream);
        else rc = launch_residual(L.op(OP_LAP), L.Nd, 2, L.dbtag.p, src, p, res, g_stream);
        if (rc) return rc;
    }
    return 0;
}

int op_residual(mgm_problem *pb, int l, int variant)
{
    Level &L = pb->lv[l];
    if (variant == MGM_VAR_F)
        return launch_residual(L.op(OP_LAP), L.Nd, 3, L.dbtag.p, L.f[MGM_F_SOURCE].p, L.f[MGM_F_P].p,
                               L.f[MGM_F_RES].p, g_stream);
    return launch_residual(L.op(OP_LAPP), L.Nd, 0, L.dbtag.p, L.f[MGM_F_SOURCE].p,
                           L.f[pfield(variant)].p, L.f[MGM_F_RES].p, g_stream);
}

int op_restrict(mgm_problem *pb, int l, int variant)
{
    Level &F = pb->lv[l], &C = pb->lv[l + 1];
    return launch_restrict(C.restr(), C.Nd, C.dbtag.p, variant == MGM_VAR_F, F.f[MGM_F_RES].p,
                           C.f[MGM_F_SOURCE].p, g_stream);
}

int op_prolong(mgm_problem *pb, int l, int variant)
{
    Level &C = pb->lv[l], &F = pb->lv[l - 1];
    int pf = pfield(variant);
    int rc = launch_prolong(F.prol(), F.Nd, F.dbtag.p, C.f[pf].p, F.f[pf].p, g_stream);
    if (rc) return rc;
    return launch_zero(C.Nd, C.dbtag.p, variant != MGM_VAR_T, C.f[pf].p, g_stream);
}

int op_vcycle(mgm_problem *pb, int variant)
{
    int rc;
    const int L = pb->L;
    for (int l = 0; l < L; l++) {
        if ((rc = op_relax(pb, l, variant))) return rc;
        if ((rc = op_residual(pb, l, variant))) return rc;
        if (l != L - 1 && (rc = op_restrict(pb, l, variant))) return rc;
    }
    for (int l = L - 1; l > 0; l--) {
        if ((rc = op_prolong(pb, l, variant))) return rc;
        if (l != 1 && (rc = op_relax(pb, l - 1, variant))) return rc;
    }
    return 0;
}

}  // namespace mgm

extern "C" {

#define OP_SYNC(expr)                                     \
    do {                                                  \
        int _rc = (expr);                                 \
        if (_rc) return _rc;                              \
        HIP_TRY(hipStreamSynchronize(g_stream));          \
        return 0;                                         \
    } while (0)

int mgm_relax(mgm_problem *pb, int lev, int variant)
{
    CHECK_READY(pb, lev);
    OP_SYNC(op_relax(pb, lev, variant));
}

int mgm_residual(mgm_problem *pb, int lev, int variant)
{
    CHECK_READY(pb, lev);
    OP_SYNC(op_residual(pb, lev, variant));
}

int mgm_restrict(mgm_problem *pb, int lev, int variant)
{
    CHECK_READY(pb, lev);
    if (lev >= pb->L - 1) MGM_FAIL(-1, "restrict: no coarser level");
    OP_SYNC(op_restrict(pb, lev, variant));
}

int mgm_prolong(mgm_problem *pb, int lev, int variant)
{
    CHECK_READY(pb, lev);
    if (lev < 1) MGM_FAIL(-1, "prolong: no finer level");
    OP_SYNC(op_prolong(pb, lev, variant));
}

int mgm_vcycle(mgm_problem *pb, int variant, int ncycles)
{
    CHECK_READY(pb, 0);
    for (int k = 0; k < ncycles; k++) {
        int rc = op_vcycle(pb, variant);
        if (rc) return rc;
    }
    HIP_TRY(hipStreamSynchronize(g_stream));
    return 0;
}

int mgm_spmv(mgm_problem *pb, int lev, int which_matrix, int in_field, int out_field)
{
    CHECK_READY(pb, lev);
    if (which_matrix < MGM_MAT_DX || which_matrix > MGM_MAT_LAPP) MGM_FAIL(-1, "spmv: unknown matrix");
    if (in_field < 0 || in_field >= MGM_NUM_FIELDS || out_field < 0 || out_field >= MGM_NUM_FIELDS ||
        in_field == out_field)
        MGM_FAIL(-1, "spmv: bad field ids");
    Level &L = pb->lv[lev];
    if (!L.slab[which_matrix].p) MGM_FAIL(-1, "spmv: matrix not built");
    OP_SYNC(launch_spmv(L.op(which_matrix), L.Nd, L.dbtag.p, L.f[in_field].p, L.f[out_field].p, g_stream));
}

int mgm_spmv_wide(mgm_problem *pb, int lev, int which_matrix, int in_field, int out_field)
{
    CHECK_READY(pb, lev);
    if (which_matrix < MGM_MAT_DX || which_matrix > MGM_MAT_LAP ||
        (which_matrix == MGM_MAT_DZ && pb->cfg.dimension != 3))
        MGM_FAIL(-1, "spmv_wide: Dx, Dy, Dz (3D) or lap only");
    if (in_field < 0 || in_field >= MGM_NUM_FIELDS || out_field < 0 || out_field >= MGM_NUM_FIELDS ||
        in_field == out_field)
        MGM_FAIL(-1, "spmv_wide: bad field ids");
    Level &L = pb->lv[lev];
    if (!L.mb) MGM_FAIL(-1, "spmv_wide: this level has no wide boundary stencils");
    int rc = launch_spmv(L.op(which_matrix), L.Nd, L.dbtag.p, L.f[in_field].p, L.f[out_field].p, g_stream);
    if (rc) return rc;
    rc = launch_wide_rows((int)L.bnodes.size(), L.mb, L.d_bnodes.p, L.rcloud_b.p, L.rows_b[which_matrix].p,
                          L.d_iperm.p, L.f[in_field].p, L.f[out_field].p, g_stream);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(g_stream));
    return 0;
}

int mgm_spmv_rows(int n_rows, int n_cols, const double *D, const int32_t *cloud, int ncol,
                  const double *f, double *out)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (n_rows <= 0 || n_cols <= 0 || ncol <= 0 || !D || !cloud || !f || !out) MGM_FAIL(-1, "spmv_rows: bad arguments");
    const int S = slice_height(n_cols);
    size_t sz = ell_size(n_rows, n_cols, S);
    std::vector<double> w(sz, 0.0);
    std::vector<int> c(sz, 0);
    for (int i = 0; i < n_rows; i++)
        for (int j = 0; j < n_cols; j++) {
            int v = cloud[(size_t)i * n_cols + j];
            if (v < 0 || v >= ncol) MGM_FAIL(-1, "spmv_rows: column index out of range");
            w[ell_at(n_cols, S, i, j)] = D[(size_t)i * n_cols + j];
            c[ell_at(n_cols, S, i, j)] = v;
        }
    DevBuf<double> dw, dx, dy;
    DevBuf<int> dc;
    HIP_TRY(dw.upload(w, g_stream));
    HIP_TRY(dc.upload(c, g_stream));
    HIP_TRY(dx.upload(f, ncol, g_stream));
    HIP_TRY(dy.alloc(n_rows));
    rc = launch_spmv(Ell{dc.p, dw.p, n_cols, S}, n_rows, nullptr, dx.p, dy.p, g_stream);
    if (rc) return rc;
    HIP_TRY(dy.download(out, n_rows, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    return 0;
}

int mgm_spmv_rows_static(int n_rows, int n_cols, const double *D, const int32_t *cloud, int ncol,
                         const double *f, double *out)
{
    int rc = ensure_init();
    if (rc) return rc;
    return spmv_rows_static(g_device, n_rows, n_cols, D, cloud, ncol, f, out, g_stream);
}

int mgm_l2_norm_gen(const mgm_problem *pbc, int lev, int fa, int fb, double *out)
{
    CHECK_READY(pbc, lev);
    mgm_problem *pb = const_cast<mgm_problem *>(pbc);
    Level &L = pb->lv[lev];
    int rc = launch_sqdiff_sum(L.Nd, L.dbtag.p, L.f[fa].p, fb >= 0 ? L.f[fb].p : nullptr, pb->partial.p,
                               pb->scal.p, g_stream);
    if (rc) return rc;
    double s = 0;
    HIP_TRY(hipMemcpyAsync(&s, pb->scal.p, 8, hipMemcpyDeviceToHost, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    *out = sqrt(s / L.N);
    return 0;
}

}  // extern "C"
