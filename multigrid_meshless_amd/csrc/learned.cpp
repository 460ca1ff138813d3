// learned.cpp -- the learned smoother: training on (error, residual) pairs the
// plain V-cycle produces, inference as the first sweep of each level's
// relaxation, and its C API.  No reference counterpart (the reference's
// relaxation is SOR, hf/heat_conduction.h:25-56).
//
// Training data (level 0): for each snapshot a random interior source f,
// the converged FP64 solution p* (mgm_solve to 1e-12), then from the start
// p = p* + noise the states after 0, 1 and 2 SOR sweeps; each state gives
// r = f - A p and e = p* - p on every node, of which random interior rows are
// sampled.  Targets t = e a_ii / sigma are clipped to |t| <= 8: smooth error
// components are not determined by a local residual and would otherwise
// dominate the squared loss.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mlp.hpp"
#include "problem.hpp"

namespace mgm {
int op_relax(mgm_problem *pb, int l, int variant);
}

using namespace mgm;

namespace {

constexpr float kTargetClip = 8.0f;

uint64_t mix(uint64_t &s) { return splitmix64(s); }
double unif(uint64_t &s) { return (double)(mix(s) >> 11) * 0x1.0p-53; }
double normal(uint64_t &s)
{
    const double u1 = std::max(unif(s), 1e-300), u2 = unif(s);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

LearnedState &state(mgm_problem *pb)
{
    if (!pb->ml) pb->ml.reset(new LearnedState());
    LearnedState &S = *pb->ml;
    const Level &L0 = pb->lv[0];
    S.in = L0.m * pb->cfg.dimension + L0.m;
    return S;
}

// device-order coordinates of every level (padding rows 0)
int ensure_coords(mgm_problem *pb, LearnedState &S)
{
    if ((int)S.dx.size() == pb->L) return 0;
    S.dx.clear(); S.dy.clear(); S.dz.clear();
    S.dx.resize(pb->L); S.dy.resize(pb->L); S.dz.resize(pb->L);
    for (int l = 0; l < pb->L; l++) {
        const Level &L = pb->lv[l];
        const std::vector<double> *src[3] = {&L.x, &L.y, &L.z};
        DevBuf<double> *dst[3] = {&S.dx[l], &S.dy[l], &S.dz[l]};
        for (int d = 0; d < 3; d++) {
            std::vector<double> t(L.Nd, 0.0);
            for (int i = 0; i < L.Nd; i++)
                if (L.perm[i] >= 0) t[i] = (*src[d])[L.perm[i]];
            HIP_TRY(dst[d]->upload(t, stream()));
            HIP_TRY(hipStreamSynchronize(stream()));
        }
    }
    return 0;
}

// BF16 inference copies and the per-level geometric part of layer 1.
int prepare(mgm_problem *pb, LearnedState &S)
{
    if (S.prepared) return 0;
    if (!S.have) MGM_FAIL(-1, "learned smoother: no weights (train or set them first)");
    int rc = ensure_coords(pb, S);
    if (rc) return rc;
    const int dim = pb->cfg.dimension, m0 = pb->lv[0].m, in = S.in;
    const float *w = S.w.data();
    const float *W1 = w, *W2 = w + MLP_H * in + MLP_H, *b2 = W2 + MLP_H * MLP_H, *w3 = b2 + MLP_H;
    std::vector<float> w1r((size_t)MLP_H * m0);   // residual columns of W1, transposed: [j][k]
    for (int k = 0; k < MLP_H; k++)
        for (int j = 0; j < m0; j++) w1r[(size_t)j * MLP_H + k] = W1[(size_t)k * in + m0 * dim + j];
    DevBuf<float> tmp;
    auto to16 = [&](const float *h, size_t n, DevBuf<bf16> &out) -> int {
        HIP_TRY(tmp.upload(h, n, stream()));
        HIP_TRY(out.alloc(n));
        int r = mlp_to_bf16((int)n, tmp.p, out.p, stream());
        if (r) return r;
        HIP_TRY(hipStreamSynchronize(stream()));
        return 0;
    };
    if ((rc = to16(w1r.data(), w1r.size(), S.W1r))) return rc;
    if ((rc = to16(W2, (size_t)MLP_H * MLP_H, S.W2))) return rc;
    if ((rc = to16(w3, MLP_H, S.w3))) return rc;
    HIP_TRY(S.b2.upload(b2, MLP_H, stream()));
    S.b3 = w3[MLP_H];
    HIP_TRY(S.net.alloc(mlp_net_bytes()));
    if ((rc = mlp_pack_net(m0, S.W1r.p, S.W2.p, S.b2.p, S.w3.p, S.net.p, stream()))) return rc;
    HIP_TRY(S.W.upload(S.w, stream()));
    S.G.clear();
    S.G.resize(pb->L);
    S.lev_ok.assign(pb->L, 0);
    for (int l = 0; l < pb->L; l++) {
        Level &L = pb->lv[l];
        if (L.m != m0 || L.m > MLP_MAX_M) continue;   // network input width differs: SOR only
        HIP_TRY(S.G[l].alloc(mlp_geometry_size(L.Nd)));
        if ((rc = mlp_geometry(L.Nd, L.op_plain(OP_LAPP), dim, L.d_avg, S.dx[l].p, S.dy[l].p, S.dz[l].p, L.dbtag.p,
                               S.W.p, S.G[l].p, stream())))
            return rc;
        S.lev_ok[l] = 1;
    }
    HIP_TRY(hipStreamSynchronize(stream()));
    S.prepared = true;
    return 0;
}

int set_weights(mgm_problem *pb, LearnedState &S, const float *w)
{
    S.w.assign(w, w + mlp_weight_count(S.in));
    for (float v : S.w)
        if (!std::isfinite(v)) MGM_FAIL(-1, "learned smoother: non-finite weight");
    S.have = true;
    S.prepared = false;
    HIP_TRY(S.W.upload(S.w, stream()));
    HIP_TRY(hipStreamSynchronize(stream()));
    return 0;
}

// Adam on the current weights over device rows X (n x in), targets t:
// `epochs` passes, each over a fresh shuffle (Fisher-Yates from `seed`) cut
// into floor(n / batch) batches; loss over all rows before and after.
int fit(mgm_problem *pb, LearnedState &S, int n, const float *X, const float *t, int epochs, int batch, float lr,
        uint64_t &seed, double *loss_first, double *loss_last)
{
    const int in = S.in;
    const size_t nw = mlp_weight_count(in);
    DevBuf<float> M, V, G, H1, H2, y, dy, dZ1, dZ2, Xb, tb, part;
    HIP_TRY(M.alloc(nw)); HIP_TRY(V.alloc(nw)); HIP_TRY(G.alloc(nw));
    HIP_TRY(part.alloc(mlp_gradient_workspace(in)));
    HIP_TRY(hipMemsetAsync(M.p, 0, sizeof(float) * nw, stream()));
    HIP_TRY(hipMemsetAsync(V.p, 0, sizeof(float) * nw, stream()));
    const int B = std::min(batch, n);
    HIP_TRY(H1.alloc((size_t)n * MLP_H)); HIP_TRY(H2.alloc((size_t)n * MLP_H)); HIP_TRY(y.alloc(n));
    HIP_TRY(dy.alloc(B)); HIP_TRY(dZ1.alloc((size_t)B * MLP_H)); HIP_TRY(dZ2.alloc((size_t)B * MLP_H));
    HIP_TRY(Xb.alloc((size_t)B * in)); HIP_TRY(tb.alloc(B));
    DevBuf<int> didx;
    auto full_loss = [&](double *out) -> int {
        int r = mlp_forward(n, in, X, S.W.p, H1.p, H2.p, y.p, stream());
        if (!r) r = mlp_sqerr(n, y.p, t, pb->partial.p, pb->scal.p, stream());
        if (r) return r;
        double s = 0;
        HIP_TRY(hipMemcpyAsync(&s, pb->scal.p, sizeof(double), hipMemcpyDeviceToHost, stream()));
        HIP_TRY(hipStreamSynchronize(stream()));
        *out = s / n;
        return 0;
    };
    double l0 = 0, l1 = 0;
    int rc = full_loss(&l0);
    if (rc) return rc;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    long step = 0;
    for (int e = 0; e < epochs; e++) {
        for (int i = n - 1; i > 0; i--) std::swap(order[i], order[mix(seed) % (uint64_t)(i + 1)]);
        HIP_TRY(didx.upload(order, stream()));
        for (int b0 = 0; b0 + B <= n; b0 += B) {
            step++;
            // bias-corrected step size: lr sqrt(1 - b2^step) / (1 - b1^step)
            const float lr_t = lr * (float)(std::sqrt(1.0 - std::pow(b2, step)) / (1.0 - std::pow(b1, step)));
            rc = mlp_gather_batch(B, in, didx.p + b0, X, t, Xb.p, tb.p, stream());
            if (!rc) rc = mlp_forward(B, in, Xb.p, S.W.p, H1.p, H2.p, y.p, stream());
            if (!rc) rc = mlp_gradient(B, in, Xb.p, tb.p, S.W.p, H1.p, H2.p, y.p, dy.p, dZ2.p, dZ1.p, part.p, G.p,
                                       stream());
            if (!rc) rc = mlp_adam((int)nw, S.W.p, G.p, M.p, V.p, lr_t, b1, b2, eps, stream());
            if (rc) return rc;
        }
        HIP_TRY(hipStreamSynchronize(stream()));
    }
    if ((rc = full_loss(&l1))) return rc;
    HIP_TRY(S.W.download(S.w.data(), nw, stream()));
    HIP_TRY(hipStreamSynchronize(stream()));
    for (float v : S.w)
        if (!std::isfinite(v)) MGM_FAIL(-11, "train_smoother: training diverged");
    S.prepared = false;
    if (loss_first) *loss_first = l0;
    if (loss_last) *loss_last = l1;
    return 0;
}

// Level-0 reference-order field as a device-order vector (padding rows 0).
int upload_ref_field(const Level &L, const double *v, DevBuf<double> &out)
{
    std::vector<double> t(L.Nd, 0.0);
    for (int d = 0; d < L.Nd; d++)
        if (L.perm[d] >= 0) t[d] = v[L.perm[d]];
    HIP_TRY(out.upload(t, stream()));
    return 0;
}

}  // namespace

namespace mgm {

// The learned first sweep of level l: RES = s - A p (residual form 0), then
// p += network correction on interior rows.  Inside mgm_solve level 0's RES
// already holds that residual when a cycle starts (the loop's own fine
// residual, same form, p unchanged since), so it is not formed again.
int op_learned_sweep(mgm_problem *pb, int l)
{
    LearnedState &S = *pb->ml;
    Level &L = pb->lv[l];
    if (!(l == 0 && pb->res0_current)) {
        int rc = launch_residual(L.op(OP_LAPP), L.Nd, 0, L.dbtag.p, (const double *)L.f[MGM_F_SOURCE].p,
                                 (const double *)L.f[MGM_F_P].p, L.f[MGM_F_RES].p, stream());
        if (rc) return rc;
    }
    return mlp_correct(L.Nd, L.with_window(L.op_plain(OP_LAPP)), L.dbtag.p, S.G[l].p, S.net.p, S.b3,
                       L.f[MGM_F_RES].p, L.f[MGM_F_P].p, stream());
}

// The same sweep inside the reduced-precision V-cycle: FP32 residual (res32
// = src32 - A p32 with the level's narrowed weights) and FP32 correction.
// Level 0 is relaxed once per correction, first, with p32 = 0: its residual
// is src32 itself (the row sums of zero products are +0, and src32 is 0 on
// boundary rows as residual form 0 leaves them), so src32 is used directly.
int op_learned_sweep_mp(mgm_problem *pb, int l, int (*residual_mp)(Level &))
{
    LearnedState &S = *pb->ml;
    Level &L = pb->lv[l];
    const float *r = L.src32.p;
    if (l != 0) {
        int rc = residual_mp(L);
        if (rc) return rc;
        r = L.res32.p;
    }
    return mlp_correct(L.Nd, L.with_window(L.op_plain(OP_LAPP)), L.dbtag.p, S.G[l].p, S.net.p, S.b3,
                       r, L.p32.p, stream());
}

bool learned_active(const mgm_problem *pb, int l, int variant)
{
    return pb->ml && pb->ml->enabled && pb->ml->prepared && variant == MGM_VAR_H &&
           pb->cfg.sweep_order == MGM_ORDER_COLOUR && l < (int)pb->ml->lev_ok.size() && pb->ml->lev_ok[l] &&
           (pb->ml->lev_use.empty() || (l < (int)pb->ml->lev_use.size() && pb->ml->lev_use[l]));
}

}  // namespace mgm

extern "C" {

int mgm_mlp_inputs(const mgm_problem *pb)
{
    if (!pb || !pb->ready) MGM_FAIL(-1, "problem is not set up");
    return pb->lv[0].m * pb->cfg.dimension + pb->lv[0].m;
}

size_t mgm_mlp_weight_count(int inputs) { return inputs > 0 ? mlp_weight_count(inputs) : 0; }

int mgm_mlp_set_weights(mgm_problem *pb, const float *w)
{
    if (!pb || !pb->ready) MGM_FAIL(-1, "problem is not set up");
    if (!w) MGM_FAIL(-1, "null weights");
    return set_weights(pb, state(pb), w);
}

int mgm_mlp_get_weights(const mgm_problem *pbc, float *w)
{
    mgm_problem *pb = const_cast<mgm_problem *>(pbc);
    if (!pb || !pb->ready) MGM_FAIL(-1, "problem is not set up");
    LearnedState &S = state(pb);
    if (!S.have) MGM_FAIL(-1, "learned smoother: no weights");
    memcpy(w, S.w.data(), sizeof(float) * S.w.size());
    return 0;
}

int mgm_mlp_forward(mgm_problem *pb, int n, const float *X, float *y)
{
    if (!pb || !pb->ready) MGM_FAIL(-1, "problem is not set up");
    LearnedState &S = state(pb);
    if (!S.have) MGM_FAIL(-1, "learned smoother: no weights");
    if (n <= 0 || !X || !y) MGM_FAIL(-1, "mlp_forward: bad arguments");
    DevBuf<float> dX, dH1, dH2, dy;
    HIP_TRY(dX.upload(X, (size_t)n * S.in, stream()));
    HIP_TRY(dH1.alloc((size_t)n * MLP_H)); HIP_TRY(dH2.alloc((size_t)n * MLP_H)); HIP_TRY(dy.alloc(n));
    int rc = mlp_forward(n, S.in, dX.p, S.W.p, dH1.p, dH2.p, dy.p, stream());
    if (rc) return rc;
    HIP_TRY(dy.download(y, n, stream()));
    HIP_TRY(hipStreamSynchronize(stream()));
    return 0;
}

int mgm_mlp_gradient(mgm_problem *pb, int n, const float *X, const float *t, double *loss, float *grad)
{
    if (!pb || !pb->ready) MGM_FAIL(-1, "problem is not set up");
    LearnedState &S = state(pb);
    if (!S.have) MGM_FAIL(-1, "learned smoother: no weights");
    if (n <= 0 || !X || !t) MGM_FAIL(-1, "mlp_gradient: bad arguments");
    const size_t nw = mlp_weight_count(S.in);
    DevBuf<float> dX, dt, dH1, dH2, dy, ddy, dZ1, dZ2, dG, dpart;
    HIP_TRY(dX.upload(X, (size_t)n * S.in, stream()));
    HIP_TRY(dt.upload(t, n, stream()));
    for (DevBuf<float> *b : {&dH1, &dH2, &dZ1, &dZ2}) HIP_TRY(b->alloc((size_t)n * MLP_H));
    HIP_TRY(dy.alloc(n)); HIP_TRY(ddy.alloc(n)); HIP_TRY(dG.alloc(nw));
    HIP_TRY(dpart.alloc(mlp_gradient_workspace(S.in)));
    int rc = mlp_forward(n, S.in, dX.p, S.W.p, dH1.p, dH2.p, dy.p, stream());
    if (!rc) rc = mlp_gradient(n, S.in, dX.p, dt.p, S.W.p, dH1.p, dH2.p, dy.p, ddy.p, dZ2.p, dZ1.p, dpart.p, dG.p,
                               stream());
    if (!rc) rc = mlp_sqerr(n, dy.p, dt.p, pb->partial.p, pb->scal.p, stream());
    if (rc) return rc;
    double s = 0;
    HIP_TRY(hipMemcpyAsync(&s, pb->scal.p, sizeof(double), hipMemcpyDeviceToHost, stream()));
    if (grad) HIP_TRY(dG.download(grad, nw, stream()));
    HIP_TRY(hipStreamSynchronize(stream()));
    if (loss) *loss = s / n;
    return 0;
}

int mgm_train_smoother(mgm_problem *pb, const mgm_train_config *cfg, double *loss_first, double *loss_last)
{
    if (!pb || !pb->ready) MGM_FAIL(-1, "problem is not set up");
    if (!cfg || cfg->snapshots < 1 || cfg->samples_per_snapshot < 1 || cfg->epochs < 1 || cfg->batch < 1 ||
        !(cfg->lr > 0))
        MGM_FAIL(-1, "train_smoother: bad configuration");
    if (pb->cfg.sweep_order != MGM_ORDER_COLOUR) MGM_FAIL(-1, "train_smoother: needs the multicolour order");
    LearnedState &S = state(pb);
    Level &L0 = pb->lv[0];
    if (L0.m > MLP_MAX_M) MGM_FAIL(-1, "learned smoother: stencil wider than supported");
    int rc = ensure_coords(pb, S);
    if (rc) return rc;
    const bool was_enabled = S.enabled;
    S.enabled = false;                       // data come from the plain V-cycle
    const int num_relax = pb->cfg.num_relax;
    const int dim = pb->cfg.dimension, in = S.in, Nd = L0.Nd, N = L0.N;
    uint64_t seed = cfg->seed;

    std::vector<int> interior;                // device rows of interior nodes
    for (int d = 0; d < Nd; d++)
        if (L0.perm[d] >= 0 && !L0.btag[L0.perm[d]]) interior.push_back(d);
    if (interior.empty()) MGM_FAIL(-1, "train_smoother: no interior nodes");

    // ---- (error, residual) snapshots from the plain V-cycle --------------------
    constexpr int kStates = 3;               // after 0, 1, 2 SOR sweeps
    const int nsnap = cfg->snapshots * kStates;
    DevBuf<double> res_s, err_s, pstar;
    HIP_TRY(res_s.alloc((size_t)nsnap * Nd));
    HIP_TRY(err_s.alloc((size_t)nsnap * Nd));
    HIP_TRY(pstar.alloc(Nd));
    std::vector<double> f(N), p0(N);
    std::vector<int> rows, snaps;
    for (int s = 0; s < cfg->snapshots; s++) {
        for (int i = 0; i < N; i++) f[i] = L0.btag[i] ? 0.0 : normal(seed);
        if ((rc = mgm_zero_fields(pb))) return rc;
        if ((rc = mgm_set_field(pb, 0, MGM_F_SOURCE, f.data()))) return rc;
        int cycles = 0;
        if ((rc = mgm_solve(pb, MGM_VAR_H, 0, 1e-12, 400, &cycles, nullptr, nullptr))) return rc;
        std::vector<double> ps(N);
        if ((rc = mgm_get_field(pb, 0, MGM_F_P, ps.data()))) return rc;
        double rms = 0;
        for (int i = 0; i < N; i++) rms += ps[i] * ps[i];
        rms = std::sqrt(rms / N);
        HIP_TRY(hipMemcpyAsync(pstar.p, L0.f[MGM_F_P].p, sizeof(double) * Nd, hipMemcpyDeviceToDevice, stream()));
        for (int i = 0; i < N; i++) p0[i] = ps[i] + (L0.btag[i] ? 0.0 : rms * normal(seed));
        if ((rc = mgm_set_field(pb, 0, MGM_F_P, p0.data()))) return rc;
        pb->cfg.num_relax = 1;
        for (int k = 0; k < kStates; k++) {
            if (k > 0 && (rc = op_relax(pb, 0, MGM_VAR_H))) { pb->cfg.num_relax = num_relax; return rc; }
            const int idx = s * kStates + k;
            rc = launch_residual(L0.op(OP_LAPP), Nd, 0, L0.dbtag.p, (const double *)L0.f[MGM_F_SOURCE].p,
                                 (const double *)L0.f[MGM_F_P].p, res_s.p + (size_t)idx * Nd, stream());
            if (!rc) rc = mlp_diff(Nd, pstar.p, L0.f[MGM_F_P].p, err_s.p + (size_t)idx * Nd, stream());
            if (rc) { pb->cfg.num_relax = num_relax; return rc; }
            for (int q = 0; q < cfg->samples_per_snapshot; q++) {
                rows.push_back(interior[mix(seed) % interior.size()]);
                snaps.push_back(idx);
            }
        }
        pb->cfg.num_relax = num_relax;
        HIP_TRY(hipStreamSynchronize(stream()));
    }
    mgm_zero_fields(pb);

    // ---- feature rows ------------------------------------------------------------
    const int n = (int)rows.size();
    DevBuf<int> drows, dsnap;
    DevBuf<float> X, t;
    HIP_TRY(drows.upload(rows, stream()));
    HIP_TRY(dsnap.upload(snaps, stream()));
    HIP_TRY(X.alloc((size_t)n * in));
    HIP_TRY(t.alloc(n));
    if ((rc = mlp_features(n, L0.op_plain(OP_LAPP), dim, L0.d_avg, S.dx[0].p, S.dy[0].p, S.dz[0].p, drows.p, dsnap.p,
                           Nd, res_s.p, err_s.p, X.p, t.p, stream())))
        return rc;
    std::vector<float> th(n);
    HIP_TRY(t.download(th.data(), n, stream()));
    HIP_TRY(hipStreamSynchronize(stream()));
    for (float &v : th) v = std::max(-kTargetClip, std::min(kTargetClip, v));
    HIP_TRY(t.upload(th, stream()));
    res_s.release(); err_s.release();

    // ---- initial weights (He normal), Adam -------------------------------------------
    const size_t nw = mlp_weight_count(in);
    std::vector<float> w(nw, 0.0f);
    {
        float *W1 = w.data(), *W2 = W1 + MLP_H * in + MLP_H, *w3 = W2 + MLP_H * MLP_H + MLP_H;
        for (int i = 0; i < MLP_H * in; i++) W1[i] = (float)(normal(seed) * std::sqrt(2.0 / in));
        for (int i = 0; i < MLP_H * MLP_H; i++) W2[i] = (float)(normal(seed) * std::sqrt(2.0 / MLP_H));
        for (int i = 0; i < MLP_H; i++) w3[i] = (float)(normal(seed) * std::sqrt(1.0 / MLP_H));
    }
    if ((rc = set_weights(pb, S, w.data()))) return rc;
    rc = fit(pb, S, n, X.p, t.p, cfg->epochs, cfg->batch, cfg->lr, seed, loss_first, loss_last);
    if (rc) return rc;
    S.enabled = was_enabled;
    if (S.enabled && (rc = prepare(pb, S))) return rc;
    return 0;
}

int mgm_mlp_features(mgm_problem *pb, int n, const int32_t *nodes, const double *res, const double *err, float *X,
                     float *t)
{
    if (!pb || !pb->ready) MGM_FAIL(-1, "problem is not set up");
    if (n <= 0 || !nodes || !res || !err || !X || !t) MGM_FAIL(-1, "mlp_features: bad arguments");
    LearnedState &S = state(pb);
    Level &L0 = pb->lv[0];
    if (L0.m > MLP_MAX_M) MGM_FAIL(-1, "learned smoother: stencil wider than supported");
    int rc = ensure_coords(pb, S);
    if (rc) return rc;
    std::vector<int> rows(n);
    for (int b = 0; b < n; b++) {
        if (nodes[b] < 0 || nodes[b] >= L0.N) MGM_FAIL(-1, "mlp_features: node index out of range");
        rows[b] = L0.iperm[nodes[b]];
    }
    DevBuf<int> drows;
    DevBuf<double> dres, derr;
    DevBuf<float> dX, dt;
    HIP_TRY(drows.upload(rows, stream()));
    if ((rc = upload_ref_field(L0, res, dres)) || (rc = upload_ref_field(L0, err, derr))) return rc;
    HIP_TRY(dX.alloc((size_t)n * S.in));
    HIP_TRY(dt.alloc(n));
    if ((rc = mlp_features(n, L0.op_plain(OP_LAPP), pb->cfg.dimension, L0.d_avg, S.dx[0].p, S.dy[0].p, S.dz[0].p,
                           drows.p, nullptr, L0.Nd, dres.p, derr.p, dX.p, dt.p, stream())))
        return rc;
    HIP_TRY(dX.download(X, (size_t)n * S.in, stream()));
    HIP_TRY(dt.download(t, n, stream()));
    HIP_TRY(hipStreamSynchronize(stream()));
    return 0;
}

int mgm_mlp_fit(mgm_problem *pb, int n, const float *X, const float *t, int epochs, int batch, float lr,
                uint64_t seed, double *loss_first, double *loss_last)
{
    if (!pb || !pb->ready) MGM_FAIL(-1, "problem is not set up");
    if (n <= 0 || !X || !t || epochs < 1 || batch < 1 || !(lr > 0)) MGM_FAIL(-1, "mlp_fit: bad arguments");
    LearnedState &S = state(pb);
    if (!S.have) MGM_FAIL(-1, "learned smoother: no weights (set them first)");
    DevBuf<float> dX, dt;
    HIP_TRY(dX.upload(X, (size_t)n * S.in, stream()));
    HIP_TRY(dt.upload(t, n, stream()));
    int rc = fit(pb, S, n, dX.p, dt.p, epochs, batch, lr, seed, loss_first, loss_last);
    if (rc) return rc;
    if (S.enabled && (rc = prepare(pb, S))) return rc;
    return 0;
}

int mgm_use_learned_smoother(mgm_problem *pb, int enable)
{
    if (!pb || !pb->ready) MGM_FAIL(-1, "problem is not set up");
    LearnedState &S = state(pb);
    if (!enable) {
        S.enabled = false;
        return 0;
    }
    if (pb->cfg.sweep_order != MGM_ORDER_COLOUR) MGM_FAIL(-1, "learned smoother: needs the multicolour order");
    int rc = prepare(pb, S);
    if (rc) return rc;
    S.enabled = true;
    return 0;
}

int mgm_learned_fell_back(const mgm_problem *pb)
{
    return (pb && pb->ml) ? pb->ml->fell_back : 0;
}

int mgm_set_learned_levels(mgm_problem *pb, const int *use, int n)
{
    if (!pb || !pb->ready) MGM_FAIL(-1, "problem is not set up");
    if (n != 0 && n != pb->L) MGM_FAIL(-1, "set_learned_levels: n must be 0 or the number of levels");
    if (n && !use) MGM_FAIL(-1, "set_learned_levels: null use");
    if (!pb->ml) pb->ml.reset(new LearnedState());
    pb->ml->lev_use.assign(n, 0);
    for (int l = 0; l < n; l++) pb->ml->lev_use[l] = use[l] != 0;
    return 0;
}

}  // extern "C"
