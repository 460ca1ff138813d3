// knn.hip -- stencil (cloud) search on the GPU.
//
// Semantics restated from the reference (hf/kdtree_functions.h:61-151,
// hf/kdtree.c:262-293, 812-830): a stencil is the K candidates with the
// smallest sqrt(dx*dx + dy*dy + dz*dz) (no FMA, correctly rounded sqrt, so the
// values equal the host's bit for bit); equal distances keep the order in
// which the reference's insertion-ordered kd-tree range query reports them,
// i.e. the node visited later in its preorder (nearer child first) comes
// first.  That order is only consulted when distances tie, which the kernel
// detects and reports.
//
// Design: one lane per query, uniform grid over the candidates, shells of
// cells visited outward until the K-th best distance is closer than any
// unvisited cell; each lane keeps its sorted top-K in LDS (lane-interleaved so
// the 64 lanes hit distinct banks).
#pragma clang fp contract(off)

#include "device.hpp"

namespace mgm {

namespace {

constexpr int KNN_BLOCK = 64;

// true when candidate a is reported before b by the reference range query
// (a visited after b in the preorder of an insertion-ordered kd-tree).
__device__ bool kd_reported_first(const KdOrder &kd, const double p[3], int a, int b)
{
    int na = kd.node_of[a], nb = kd.node_of[b];
    int da = kd.depth[na], db = kd.depth[nb];
    while (da > db) {
        int pa = kd.parent[na];
        if (pa == nb) return true;  // b is an ancestor of a: b visited first
        na = pa;
        da--;
    }
    while (db > da) {
        int pb = kd.parent[nb];
        if (pb == na) return false;  // a is an ancestor of b
        nb = pb;
        db--;
    }
    while (kd.parent[na] != kd.parent[nb]) {
        na = kd.parent[na];
        nb = kd.parent[nb];
    }
    int c = kd.parent[na];
    int dir = kd.depth[c] % 3;
    double dx = p[dir] - kd.pos[3 * c + dir];
    int far_side = (dx <= 0.0) ? 1 : 0;  // near child is left when dx <= 0
    return kd.side[na] == far_side;      // the far subtree is visited later
}

__global__ void __launch_bounds__(KNN_BLOCK)
knn_kernel(KnnGrid g, KdOrder kd, int use_kd, int nq, const double *__restrict__ qx,
           const double *__restrict__ qy, const double *__restrict__ qz,
           const int *__restrict__ query_ids, int K, int K_out, int slot_offset,
           int *__restrict__ out, int out_stride, uint8_t *__restrict__ tie_flag)
{
    extern __shared__ unsigned char smem[];
    double *lk = (double *)smem;                   // [K][KNN_BLOCK]
    int *li = (int *)(lk + (size_t)K * KNN_BLOCK);  // [K][KNN_BLOCK]
    const int lane = threadIdx.x;
    int t = blockIdx.x * KNN_BLOCK + lane;
    if (t >= nq) return;
    const int q = query_ids ? query_ids[t] : t;
    const double p[3] = {qx[q], qy[q], qz[q]};
    auto KEY = [&](int k) -> double & { return lk[(size_t)k * KNN_BLOCK + lane]; };
    auto ID = [&](int k) -> int & { return li[(size_t)k * KNN_BLOCK + lane]; };
    auto before = [&](double ka, int ia, double kb, int ib) -> bool {
        if (ka < kb) return true;
        if (ka > kb) return false;
        if (use_kd) return kd_reported_first(kd, p, ia, ib);
        return ia < ib;
    };

    int cnt = 0;
    int cc[3];
    const double org[3] = {g.x0, g.y0, g.z0};
    const int dims[3] = {g.nx, g.ny, g.nz};
    for (int d = 0; d < 3; d++) {
        double f = floor((p[d] - org[d]) * g.inv_c);
        int ci = (f < 0) ? 0 : (f > dims[d] - 1 ? dims[d] - 1 : (int)f);
        cc[d] = ci;
    }
    for (int r = 0;; r++) {
        const int rz = (g.nz > 1) ? r : 0;
        for (int dz = -rz; dz <= rz; dz++) {
            int kz = cc[2] + dz;
            if (kz < 0 || kz >= g.nz) continue;
            for (int dy = -r; dy <= r; dy++) {
                int ky = cc[1] + dy;
                if (ky < 0 || ky >= g.ny) continue;
                int ady = dy < 0 ? -dy : dy, adz = dz < 0 ? -dz : dz;
                for (int dx = -r; dx <= r; dx++) {
                    int kx = cc[0] + dx;
                    if (kx < 0 || kx >= g.nx) continue;
                    int adx = dx < 0 ? -dx : dx;
                    int cheb = adx > ady ? adx : ady;
                    cheb = cheb > adz ? cheb : adz;
                    if (cheb != r) continue;  // only the shell at Chebyshev distance r
                    int cell = (kz * g.ny + ky) * g.nx + kx;
                    for (int e = g.cell_start[cell]; e < g.cell_start[cell + 1]; e++) {
                        int id = g.cell_pts[e];
                        double s = 0, df;
                        df = p[0] - g.px[id]; s += df * df;
                        df = p[1] - g.py[id]; s += df * df;
                        df = p[2] - g.pz[id]; s += df * df;
                        double key = sqrt(s);
                        if (cnt == K && !before(key, id, KEY(K - 1), ID(K - 1))) continue;
                        int j = (cnt < K) ? cnt++ : K - 1;
                        while (j > 0 && before(key, id, KEY(j - 1), ID(j - 1))) {
                            KEY(j) = KEY(j - 1);
                            ID(j) = ID(j - 1);
                            j--;
                        }
                        KEY(j) = key;
                        ID(j) = id;
                    }
                }
            }
        }
        // distance from p to the nearest face of the visited block that still
        // has unvisited cells beyond it
        double reach = 1e300;
        bool all = true;
        for (int d = 0; d < 3; d++) {
            if (dims[d] == 1) continue;
            int rr = r;
            if (cc[d] - rr > 0) { all = false; reach = fmin(reach, p[d] - (org[d] + (cc[d] - rr) * g.c)); }
            if (cc[d] + rr < dims[d] - 1) { all = false; reach = fmin(reach, (org[d] + (cc[d] + rr + 1) * g.c) - p[d]); }
        }
        if (all) break;
        if (cnt == K && KEY(K - 1) < reach) break;
    }
    uint8_t tie = 0;
    for (int k = 1; k < cnt; k++)
        if (KEY(k) == KEY(k - 1)) tie = 1;
    if (cnt < K_out) tie = 2;  // not enough candidates
    if (tie_flag) tie_flag[q] = tie;
    int *o = out + (size_t)q * out_stride + slot_offset;
    for (int k = 0; k < K_out; k++) o[k] = (k < cnt) ? ID(k) : -1;
}

}  // namespace

// Each query keeps its K best in LDS: 12 bytes x K x 64 lanes, 96 KiB at the
// widest stencils (128 points), above the 64 KiB default.
constexpr int kKnnMaxK = 128;

int knn_device_init()
{
    HIP_TRY(hipFuncSetAttribute((const void *)knn_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)((size_t)kKnnMaxK * KNN_BLOCK * (sizeof(double) + sizeof(int)))));
    return 0;
}

int launch_knn(const KnnGrid &g, const KdOrder *kd, int nq, const double *qx, const double *qy,
               const double *qz, const int *query_ids, int K, int K_out, int slot_offset, int *out,
               int out_stride, uint8_t *tie_flag, hipStream_t s)
{
    if (nq <= 0) return 0;
    size_t shm = (size_t)K * KNN_BLOCK * (sizeof(double) + sizeof(int));
    if (K > kKnnMaxK) MGM_FAIL(-1, "knn: stencils of more than 128 points are not supported");
    KdOrder k0{};
    int blocks = (nq + KNN_BLOCK - 1) / KNN_BLOCK;
    hipLaunchKernelGGL(knn_kernel, dim3(blocks), dim3(KNN_BLOCK), shm, s, g, kd ? *kd : k0,
                       kd ? 1 : 0, nq, qx, qy, qz, query_ids, K, K_out, slot_offset, out,
                       out_stride, tie_flag);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace mgm
