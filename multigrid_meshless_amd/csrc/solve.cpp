// solve.cpp -- resident solve-to-tolerance and stencil-apply timing.
//
// The solve runs the reference V-cycle (mg_heat_conduction.c:64-75) with the
// operators resident in HBM; after every cycle the finest-level residual
// r = source - lap_Poison * p on interior nodes is formed and its
// l2_norm_gen norm (hf/general_functions.h:181-192) is compared with the
// source's.  The per-cycle launch sequence is captured once into a hipGraph
// and replayed.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "problem.hpp"

namespace mgm {
int op_vcycle(mgm_problem *pb, int variant);
int mp_prepare(mgm_problem *pb, int precision);
int op_correction_mp(mgm_problem *pb);
int op_relax(mgm_problem *pb, int l, int variant);
}  // namespace mgm

using namespace mgm;

namespace {

double wall()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// residual of level 0 into RES and its squared interior norm into scal[0];
// with fuse, the same pass forms the first Jacobi sweep of p into jac0
int enqueue_fine_residual(mgm_problem *pb, int variant, bool fuse = false)
{
    Level &L = pb->lv[0];
    int pf = variant == MGM_VAR_T ? MGM_F_PPRIME : MGM_F_P;
    int rc;
    if (fuse) {
        const float wl = pb->omega(0);   // the factors op_relax passes to its sweeps
        rc = launch_residual_jacobi(L.op(OP_LAPP), L.dbtag.p, (const double *)L.f[MGM_F_SOURCE].p,
                                    (const double *)L.f[pf].p, pb->jac0.p, L.f[MGM_F_RES].p, (double)wl,
                                    (double)(float)(1 - wl), stream());
    } else {
        rc = launch_residual(L.op(OP_LAPP), L.Nd, 0, L.dbtag.p, L.f[MGM_F_SOURCE].p, L.f[pf].p, L.f[MGM_F_RES].p,
                             stream());
    }
    if (rc) return rc;
    return launch_sqdiff_sum(L.Nd, L.dbtag.p, L.f[MGM_F_RES].p, nullptr, pb->partial.p, pb->scal.p, stream());
}

int read_scalar(mgm_problem *pb, double *v)
{
    HIP_TRY(hipMemcpyAsync(pb->host_scal, pb->scal.p, sizeof(double), hipMemcpyDeviceToHost, stream()));
    HIP_TRY(hipStreamSynchronize(stream()));
    *v = pb->host_scal[0];
    return 0;
}

}  // namespace

extern "C" {

int mgm_solve(mgm_problem *pb, int variant, int precision, double tol, int max_cycles,
              int *cycles_done, double *history, double *seconds)
{
    if (!pb || !pb->ready) MGM_FAIL(-1, "problem is not set up");
    if (precision < 0 || precision > 2) MGM_FAIL(-1, "mgm_solve: precision must be 0, 1 or 2");
    if (variant != MGM_VAR_H && variant != MGM_VAR_T && variant != MGM_VAR_F) MGM_FAIL(-1, "bad variant");
    if (precision && variant != MGM_VAR_H) MGM_FAIL(-1, "mgm_solve: reduced precision supports the H variant");
    if (precision && pb->cfg.sweep_order != MGM_ORDER_COLOUR)
        MGM_FAIL(-1, "mgm_solve: reduced precision needs the multicolour sweep order");
    if (max_cycles < 0) MGM_FAIL(-1, "max_cycles < 0");
    Level &L0 = pb->lv[0];
    int rc;
    if (precision && (rc = mp_prepare(pb, precision))) return rc;
    // one cycle: FP64 V-cycle, or reduced-precision correction of the FP64 iterate
    auto cycle = [&]() { return precision ? op_correction_mp(pb) : op_vcycle(pb, variant); };
    // |f| over interior nodes
    rc = launch_sqdiff_sum(L0.Nd, L0.dbtag.p, L0.f[MGM_F_SOURCE].p, nullptr, pb->partial.p, pb->scal.p, stream());
    if (rc) return rc;
    double fs;
    if ((rc = read_scalar(pb, &fs))) return rc;
    double fnorm = sqrt(fs / L0.N);
    if (!(fnorm > 0)) fnorm = 1.0;

    // The learned smoother's cycles are checked against one plain cycle: if a
    // learned cycle reduces the residual less than the plain one did, the rest
    // of the solve falls back to plain SOR.
    LearnedState *ml = pb->ml.get();
    const bool learned = ml && ml->enabled && ml->prepared && variant == MGM_VAR_H;
    // Weighted Jacobi on level 0 (FP64 H cycles, LDS windows, no learned
    // sweep): the fine residual after each cycle and the first sweep of the
    // next cycle's level-0 relaxation read the same iterate, so one pass forms
    // both (the residual and the sweep are bitwise those of separate passes).
    // MGM_FUSE_RESIDUAL=0 keeps them apart.
    const char *fe = getenv("MGM_FUSE_RESIDUAL");
    const bool fuse = !(fe && fe[0] == '0') && precision == 0 && variant == MGM_VAR_H && !learned &&
                      pb->jacobi(0, variant) && pb->cfg.num_relax >= 1 && has_windows(L0.op(OP_LAPP));
    struct Jac0Flag {
        mgm_problem *pb;
        ~Jac0Flag() { pb->jac0_ready = false; }
    } jac0_flag{pb};
    if (fuse && pb->jac0.n != (size_t)L0.Nd) HIP_TRY(pb->jac0.alloc(L0.Nd));
    if ((rc = enqueue_fine_residual(pb, variant, fuse))) return rc;
    pb->jac0_ready = fuse;
    double rs;
    if ((rc = read_scalar(pb, &rs))) return rc;
    if (history) history[0] = sqrt(rs / L0.N) / fnorm;
    if (ml) ml->fell_back = 0;
    auto set_learned = [&](bool on) { if (ml) ml->enabled = on; };
    const bool ml_enabled = ml && ml->enabled;

    // capture one cycle (V-cycle + fine residual + norm) into a graph -- one
    // plain and, with the learned smoother, one learned -- unless MGM_GRAPH=0
    // asks for plain stream launches (profilers that do not follow graphs)
    const char *ge = getenv("MGM_GRAPH");
    const bool use_graph = !(ge && ge[0] == '0');
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    auto cleanup = [&]() {
        for (int g = 0; g < 2; g++) {
            if (exec[g]) (void)hipGraphExecDestroy(exec[g]);
            if (graph[g]) (void)hipGraphDestroy(graph[g]);
        }
        set_learned(ml_enabled);
    };
    // Every cycle below starts right after a fine residual of the FP64 iterate;
    // the flag holds until this call returns, on every path.
    struct Res0Flag {
        mgm_problem *pb;
        ~Res0Flag() { pb->res0_current = false; }
    } res0_flag{pb};
    pb->res0_current = precision == 0 && variant == MGM_VAR_H;
    if (use_graph) {
        for (int g = 0; g < (learned ? 2 : 1); g++) {
            set_learned(g == 1);
            HIP_TRY(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
            rc = cycle();
            if (!rc) rc = enqueue_fine_residual(pb, variant, fuse);
            hipError_t ce = hipStreamEndCapture(stream(), &graph[g]);
            if (rc) { cleanup(); return rc; }
            if (ce != hipSuccess) { cleanup(); MGM_FAIL(-10, std::string("graph capture failed: ") + hipGetErrorString(ce)); }
            hipError_t ie = hipGraphInstantiate(&exec[g], graph[g], nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                cleanup();
                MGM_FAIL(-10, std::string("graph instantiate failed: ") + hipGetErrorString(ie));
            }
        }
    }
    HIP_TRY(hipStreamSynchronize(stream()));
    double t0 = wall();
    int k = 0;
    double ratio = history ? history[0] : 1.0;
    double prev = sqrt(rs / L0.N) / fnorm, rho_plain = 0;
    bool use_learned = false;
    for (k = 1; k <= max_cycles; k++) {
        const int g = use_learned ? 1 : 0;
        if (use_graph) {
            hipError_t le = hipGraphLaunch(exec[g], stream());
            if (le != hipSuccess) { rc = -10; set_error(hipGetErrorString(le)); break; }
        } else {
            set_learned(use_learned);
            if ((rc = cycle())) break;
            if ((rc = enqueue_fine_residual(pb, variant, fuse))) break;
        }
        if ((rc = read_scalar(pb, &rs))) break;
        ratio = sqrt(rs / L0.N) / fnorm;
        if (history) history[k] = ratio;
        if (!(ratio > tol)) break;  // also stops on NaN
        if (learned) {
            const double rho = prev > 0 ? ratio / prev : 1.0;
            if (k == 1) {
                rho_plain = rho;
                use_learned = true;
            } else if (use_learned && rho > rho_plain) {
                use_learned = false;
                ml->fell_back = 1;
            }
        }
        prev = ratio;
    }
    double t1 = wall();
    cleanup();
    if (rc) return rc;
    if (cycles_done) *cycles_done = (k > max_cycles) ? max_cycles : k;
    if (seconds) *seconds = t1 - t0;
    if (std::isnan(ratio)) MGM_FAIL(-11, "solve diverged (NaN residual)");
    return 0;
}

}  // extern "C"

namespace {

// Search state of the smoother autotuners: every candidate solves from the
// caller's level-0 iterate and source with every other field zero (a diverged
// candidate leaves non-finite values behind); the same state is restored at
// the end.  Candidates that differ only in their factors cost the same per
// cycle and are ranked by cycles (a solve is capped at the best count so
// far); candidates that differ in their smoother kinds are ranked by solve
// time (capped at three times the best count).
struct OmegaSearch {
    mgm_problem *pb;
    int variant, precision, max_cycles;
    double tol;
    int pf;
    size_t nb;
    DevBuf<double> saved;
    std::vector<double> hist;
    std::vector<float> best;
    std::vector<int> best_kind;
    int best_cycles;
    double best_ratio = 0.0, best_s = 1e30;
    bool by_time = false;
    const bool trace = getenv("MGM_TUNE_TRACE") != nullptr;   // one stderr line per candidate

    OmegaSearch(mgm_problem *p, int var, int prec, double t, int maxc)
        : pb(p), variant(var), precision(prec), max_cycles(maxc), tol(t),
          pf(var == MGM_VAR_T ? MGM_F_PPRIME : MGM_F_P), nb(8 * (size_t)p->lv[0].Nd), hist(maxc + 1),
          best_cycles(maxc + 1) {}

    int save()
    {
        Level &L0 = pb->lv[0];
        HIP_TRY(saved.alloc(2 * (size_t)L0.Nd));
        HIP_TRY(hipMemcpyAsync(saved.p, L0.f[pf].p, nb, hipMemcpyDeviceToDevice, stream()));
        HIP_TRY(hipMemcpyAsync(saved.p + L0.Nd, L0.f[MGM_F_SOURCE].p, nb, hipMemcpyDeviceToDevice, stream()));
        return 0;
    }
    int restore()
    {
        Level &L0 = pb->lv[0];
        for (auto &L : pb->lv)
            for (auto &f : L.f) HIP_TRY(hipMemsetAsync(f.p, 0, 8 * (size_t)L.Nd, stream()));
        HIP_TRY(hipMemcpyAsync(L0.f[pf].p, saved.p, nb, hipMemcpyDeviceToDevice, stream()));
        HIP_TRY(hipMemcpyAsync(L0.f[MGM_F_SOURCE].p, saved.p + L0.Nd, nb, hipMemcpyDeviceToDevice, stream()));
        return 0;
    }
    // one candidate (a factor per level); kept if it reaches tol in fewer
    // cycles than the best so far, or in as many with a smaller final ratio;
    // a diverged solve counts as never converging
    int trial(const std::vector<float> &w) { return trial(w, pb->smoother_lv); }
    int trial(const std::vector<float> &w, const std::vector<int> &kind)
    {
        pb->omega_lv = w;
        pb->smoother_lv = kind;
        int rc = restore();
        if (rc) return rc;
        int cap = best_cycles <= max_cycles ? best_cycles : max_cycles;
        if (by_time && best_cycles <= max_cycles && 3 * best_cycles < max_cycles) cap = 3 * best_cycles;
        if (by_time && best_cycles > max_cycles) cap = max_cycles;
        int done = 0;
        double sec = 0.0;
        rc = mgm_solve(pb, variant, precision, tol, cap, &done, hist.data(), &sec);
        if (rc == -11) return 0;
        if (rc) return rc;
        const double ratio = hist[done];
        if (trace) {
            fprintf(stderr, "[tune] kinds");
            for (int l = 0; l < pb->L; l++) fprintf(stderr, " %d", l < (int)kind.size() ? kind[l] : 0);
            fprintf(stderr, " omega");
            for (int l = 0; l < pb->L; l++) fprintf(stderr, " %.3f", pb->omega(l));
            fprintf(stderr, ": %d cycles (cap %d) ratio %.3e %.4f s\n", done, cap, ratio, sec);
        }
        if (!(ratio <= tol)) return 0;
        // by time: a candidate must be 0.5% faster (timing noise)
        const bool better = by_time ? sec < 0.995 * best_s
                                    : (done < best_cycles || (done == best_cycles && ratio < best_ratio));
        if (better) {
            best_s = sec;
            best = w;
            best_kind = kind;
            best_cycles = done;
            best_ratio = ratio;
        }
        return 0;
    }
};

int check_omega_range(float lo, float hi, float step)
{
    if (!(lo > 0.0f && hi < 2.0f && lo <= hi && step > 0.0f)) MGM_FAIL(-1, "tune_omega: need 0 < lo <= hi < 2, step > 0");
    if ((hi - lo) / step > 64.0f) MGM_FAIL(-1, "tune_omega: more than 65 candidates");
    return 0;
}

}  // namespace

extern "C" {

int mgm_tune_omega(mgm_problem *pb, int variant, int precision, double tol, int max_cycles, float lo, float hi,
                   float step, float *omega, int *cycles)
{
    if (!pb || !pb->ready) MGM_FAIL(-1, "problem is not set up");
    if (int rc = check_omega_range(lo, hi, step)) return rc;
    if (max_cycles < 1) MGM_FAIL(-1, "tune_omega: max_cycles < 1");
    const std::vector<float> omega_lv0 = pb->omega_lv;
    const std::vector<int> kind0 = pb->smoother_lv;
    OmegaSearch S(pb, variant, precision, tol, max_cycles);
    int rc = S.save();
    if (rc) return rc;
    auto uniform = [&](float w) { return std::vector<float>(pb->L, w); };
    // a grid, then the midpoints on either side of the best point
    const int n = (int)((hi - lo) / step + 0.5f);
    for (int i = 0; i <= n && !rc; i++) rc = S.trial(uniform(lo + i * step));
    if (!rc && S.best_cycles <= max_cycles) {
        const float b = S.best[0];
        if (b - 0.5f * step > 0.0f) rc = S.trial(uniform(b - 0.5f * step));
        if (!rc && b + 0.5f * step < 2.0f) rc = S.trial(uniform(b + 0.5f * step));
    }
    const int rc2 = S.restore();
    HIP_TRY(hipStreamSynchronize(stream()));
    if (!rc) rc = rc2;
    pb->omega_lv = omega_lv0;
    pb->smoother_lv = kind0;
    if (rc) return rc;
    if (S.best_cycles > max_cycles) MGM_FAIL(-12, "tune_omega: no candidate reached the tolerance within max_cycles");
    pb->cfg.omega = S.best[0];
    pb->omega_lv.clear();
    if (omega) *omega = S.best[0];
    if (cycles) *cycles = S.best_cycles;
    return 0;
}

}  // extern "C"

namespace {

// Coordinate descent over the levels' (kind, factor): from the current
// choice, each level in turn over every allowed kind (kinds_allowed 0: its
// current kind) and the factor grid, the other levels held; up to `sweeps`
// passes, stopping after one that changes nothing.
int tune_levels(mgm_problem *pb, int variant, int precision, double tol, int max_cycles, float lo, float hi,
                float step, int sweeps, int kinds_allowed, float *omega, int *kind, int *cycles)
{
    if (!pb || !pb->ready) MGM_FAIL(-1, "problem is not set up");
    if (int rc = check_omega_range(lo, hi, step)) return rc;
    if (max_cycles < 1 || sweeps < 1 || sweeps > 8) MGM_FAIL(-1, "tune: max_cycles >= 1, sweeps 1..8");
    if (kinds_allowed && !(kinds_allowed & 3)) MGM_FAIL(-1, "tune: no smoother kind allowed");
    const std::vector<float> omega_lv0 = pb->omega_lv;
    const std::vector<int> kind0 = pb->smoother_lv;
    OmegaSearch S(pb, variant, precision, tol, max_cycles);
    S.by_time = kinds_allowed != 0 && (kinds_allowed & 3) == 3;   // kinds differ in their cost per cycle
    int rc = S.save();
    if (rc) return rc;
    std::vector<float> w0(pb->L);
    std::vector<int> k0(pb->L);
    for (int l = 0; l < pb->L; l++) {
        w0[l] = pb->omega(l);
        k0[l] = l < (int)kind0.size() ? kind0[l] : MGM_SMOOTH_SOR;
    }
    rc = S.trial(w0, k0);
    const int n = (int)((hi - lo) / step + 0.5f);
    for (int sw = 0; sw < sweeps && !rc; sw++) {
        const std::vector<float> before = S.best;
        const std::vector<int> before_k = S.best_kind;
        for (int l = 0; l < pb->L && !rc; l++)
            for (int kd = 0; kd < 2 && !rc; kd++) {
                const int cur = (S.best_kind.empty() ? k0 : S.best_kind)[l];
                if (kinds_allowed ? !(kinds_allowed & (1 << kd)) : kd != cur) continue;
                for (int i = 0; i <= n && !rc; i++) {
                    std::vector<float> w = S.best.empty() ? w0 : S.best;
                    std::vector<int> k = S.best_kind.empty() ? k0 : S.best_kind;
                    w[l] = lo + i * step;
                    k[l] = kd;
                    rc = S.trial(w, k);
                }
            }
        if (S.best == before && S.best_kind == before_k) break;
    }
    const int rc2 = S.restore();
    HIP_TRY(hipStreamSynchronize(stream()));
    if (!rc) rc = rc2;
    if (rc || S.best_cycles > max_cycles) {
        pb->omega_lv = omega_lv0;
        pb->smoother_lv = kind0;
        if (rc) return rc;
        MGM_FAIL(-12, "tune: no candidate reached the tolerance within max_cycles");
    }
    pb->omega_lv = S.best;
    pb->smoother_lv = S.best_kind;
    if (omega) std::copy(S.best.begin(), S.best.end(), omega);
    if (kind) std::copy(S.best_kind.begin(), S.best_kind.end(), kind);
    if (cycles) *cycles = S.best_cycles;
    return 0;
}

}  // namespace

extern "C" {

int mgm_tune_omega_levels(mgm_problem *pb, int variant, int precision, double tol, int max_cycles, float lo,
                          float hi, float step, int sweeps, float *omega, int *cycles)
{
    // the factors only: each level keeps its smoother kind
    return tune_levels(pb, variant, precision, tol, max_cycles, lo, hi, step, sweeps, 0, omega, nullptr, cycles);
}

int mgm_tune_smoother(mgm_problem *pb, int variant, int precision, double tol, int max_cycles, float lo, float hi,
                      float step, int sweeps, int kinds_allowed, float *omega, int *kind, int *cycles)
{
    return tune_levels(pb, variant, precision, tol, max_cycles, lo, hi, step, sweeps, kinds_allowed, omega, kind,
                       cycles);
}

int mgm_stream_copy(size_t bytes, int reps, double *seconds)
{
    if (bytes < 16 || reps < 1 || !seconds) MGM_FAIL(-1, "stream_copy: bad arguments");
    const long n16 = (long)(bytes / 16);
    DevBuf<uint8_t> a, b;
    HIP_TRY(a.alloc((size_t)n16 * 16));
    HIP_TRY(b.alloc((size_t)n16 * 16));
    HIP_TRY(hipMemsetAsync(a.p, 1, (size_t)n16 * 16, stream()));
    int rc = launch_stream_copy(n16, a.p, b.p, stream());
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, stream()));
    for (int r = 0; r < reps; r++)
        if ((rc = launch_stream_copy(n16, a.p, b.p, stream()))) return rc;
    HIP_TRY(hipEventRecord(e1, stream()));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *seconds = ms * 1e-3 / reps;
    return 0;
}

int mgm_time_spmv(mgm_problem *pb, int lev, int reps, double *seconds)
{
    if (!pb || !pb->ready) MGM_FAIL(-1, "problem is not set up");
    if (lev < 0 || lev >= pb->L || reps < 1) MGM_FAIL(-1, "bad arguments");
    Level &L = pb->lv[lev];
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int rc = launch_spmv(L.op(OP_LAPP), L.Nd, L.dbtag.p, L.f[MGM_F_P].p, L.f[MGM_F_DPDN].p, stream());
    if (rc) return rc;
    HIP_TRY(hipEventRecord(e0, stream()));
    for (int r = 0; r < reps; r++)
        if ((rc = launch_spmv(L.op(OP_LAPP), L.Nd, L.dbtag.p, L.f[MGM_F_P].p, L.f[MGM_F_DPDN].p, stream()))) return rc;
    HIP_TRY(hipEventRecord(e1, stream()));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *seconds = ms * 1e-3 / reps;
    return 0;
}

}  // extern "C"
