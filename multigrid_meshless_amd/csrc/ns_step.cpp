// ns_step.cpp -- the TIMPLE and explicit fractional time steps on the
// resident hierarchy (reference: hf/timple_solver.h:40-410,
// hf/fractional_step_solver.h:47-303).  Every loop of those functions becomes
// a kernel on the library stream; the only host synchronisation is reading
// back the steady-state error at the end of the step.  mgm_ns_march runs the
// reference driver's time loop (mg_NS_solver.c:68-83) with one whole step,
// V-cycles included, captured into a hipGraph and replayed.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <string>

#include "ns.hpp"
#include "problem.hpp"

namespace mgm {
int op_vcycle(mgm_problem *pb, int variant);
int mp_prepare_ns(mgm_problem *pb);
int op_correction_ns_mp(mgm_problem *pb, int variant);
}

using namespace mgm;

namespace {

NsDev ns_view(mgm_problem *pb)
{
    Level &L = pb->lv[0];
    NsDev d{};
    d.m = L.m;
    d.dim = pb->cfg.dimension;
    d.S = L.S;
    d.col = L.col.p;
    d.Dx = L.slab[OP_DX].p; d.Dy = L.slab[OP_DY].p; d.Dz = L.slab[OP_DZ].p; d.lap = L.slab[OP_LAP].p;
    d.btag = L.dbtag.p;
    d.nx = L.dnx.p; d.ny = L.dny.p; d.nz = L.dnz.p;
    auto F = [&](int k) { return L.f[k].p; };
    d.u = F(MGM_F_U); d.v = F(MGM_F_V); d.w = F(MGM_F_W);
    d.uo = F(MGM_F_UOLD); d.vo = F(MGM_F_VOLD); d.wo = F(MGM_F_WOLD);
    d.dpdx = F(MGM_F_DPDX); d.dpdy = F(MGM_F_DPDY); d.dpdz = F(MGM_F_DPDZ);
    d.un = F(MGM_F_UNEW); d.vn = F(MGM_F_VNEW); d.wn = F(MGM_F_WNEW);
    d.pold = F(MGM_F_POLD); d.dpdn = F(MGM_F_DPDN);
    d.rho = pb->ns.rho; d.mu = pb->ns.mu; d.dt = pb->ns.dt; d.facRe = pb->ns.facRe;
    d.facdt = pb->ns.facdt; d.field_rho = pb->ns.field_rho;
    return d;
}

double wall()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int spmv(mgm_problem *pb, int op, int in, int out)
{
    Level &L = pb->lv[0];
    return launch_spmv(L.op(op), L.Nd, L.dbtag.p, L.f[in].p, L.f[out].p, stream());
}

// steady-state error sum of squares into scal[0] (hf/timple_solver.h:63-69)
int enqueue_error(mgm_problem *pb, int a0, int a1, int b0, int b1, int c0, int c1)
{
    Level &L = pb->lv[0];
    return launch_vel_change(L.Nd, pb->cfg.dimension, L.f[a0].p, L.f[a1].p, L.f[b0].p, L.f[b1].p,
                             L.f[c0].p, L.f[c1].p, pb->partial.p, pb->scal.p, stream());
}

int read_error(mgm_problem *pb, double *err)
{
    HIP_TRY(hipMemcpyAsync(pb->host_scal, pb->scal.p, sizeof(double), hipMemcpyDeviceToHost, stream()));
    HIP_TRY(hipStreamSynchronize(stream()));
    if (err) *err = sqrt(pb->host_scal[0] / pb->lv[0].N);
    return 0;
}

#define RUN(expr) do { int _rc = (expr); if (_rc) return _rc; } while (0)

// The pressure Poisson solve of a step: num_vcycles V-cycles of the step's
// variant in FP64 (precision 0), or as many FP64 defect corrections, each by
// one FP32 V-cycle of the same variant (precision 1).
int pressure_cycles(mgm_problem *pb, int variant, int precision)
{
    for (int k = 0; k < pb->ns.num_vcycles; k++)
        RUN(precision ? op_correction_ns_mp(pb, variant) : op_vcycle(pb, variant));
    return 0;
}

// One explicit fractional step (hf/fractional_step_solver.h:47-71), error into scal[0].
int enqueue_fs(mgm_problem *pb, int precision)
{
    Level &L = pb->lv[0];
    hipStream_t s = stream();
    const int Nd = L.Nd, dim = pb->cfg.dimension;  // device rows, padding included
    auto F = [&](int k) { return L.f[k].p; };
    const double dt = pb->ns.dt, rho = pb->ns.rho, nu = pb->ns.nu;
    // hf/fractional_step_solver.h:50-55
    RUN(launch_copy(Nd, F(MGM_F_U), F(MGM_F_UOLD), s));
    RUN(launch_copy(Nd, F(MGM_F_V), F(MGM_F_VOLD), s));
    if (dim == 3) RUN(launch_copy(Nd, F(MGM_F_W), F(MGM_F_WOLD), s));
    // calculate_intermediate_velocity (:73-114)
    const int comps[3] = {MGM_F_U, MGM_F_V, MGM_F_W}, outs[3] = {MGM_F_UNEW, MGM_F_VNEW, MGM_F_WNEW};
    for (int c = 0; c < dim; c++) {
        RUN(spmv(pb, OP_DX, comps[c], MGM_F_DPDX));
        RUN(spmv(pb, OP_DY, comps[c], MGM_F_DPDY));
        if (dim == 3) RUN(spmv(pb, OP_DZ, comps[c], MGM_F_DPDZ));
        RUN(spmv(pb, OP_LAP, comps[c], MGM_F_DPDN));
        RUN(launch_fs_predict(Nd, F(comps[c]), F(MGM_F_U), F(MGM_F_V), F(MGM_F_W), F(MGM_F_DPDX),
                              F(MGM_F_DPDY), F(MGM_F_DPDZ), F(MGM_F_DPDN), dt, nu, F(outs[c]), s));
    }
    // calculate_mass_residual (:116-131)
    RUN(spmv(pb, OP_DX, MGM_F_UNEW, MGM_F_DPDX));
    RUN(spmv(pb, OP_DY, MGM_F_VNEW, MGM_F_DPDY));
    if (dim == 3) RUN(spmv(pb, OP_DZ, MGM_F_WNEW, MGM_F_DPDZ));
    RUN(launch_mass_source(Nd, L.dbtag.p, F(MGM_F_DPDX), F(MGM_F_DPDY), F(MGM_F_DPDZ), rho, dt, 0,
                           F(MGM_F_SOURCE), s));
    // calculate_boundary_dpdn (:133-148)
    RUN(launch_fs_boundary_dpdn(Nd, dim, L.dbtag.p, F(MGM_F_U), F(MGM_F_V), F(MGM_F_W), F(MGM_F_UNEW),
                                F(MGM_F_VNEW), F(MGM_F_WNEW), L.dnx.p, L.dny.p, L.dnz.p, rho,
                                dt, F(MGM_F_DPDN), s));
    // multigrid_Poisson_solver (:150-166)
    RUN(pressure_cycles(pb, MGM_VAR_F, precision));
    // update_velocity (:248-275)
    RUN(spmv(pb, OP_DX, MGM_F_P, MGM_F_DPDX));
    RUN(spmv(pb, OP_DY, MGM_F_P, MGM_F_DPDY));
    if (dim == 3) RUN(spmv(pb, OP_DZ, MGM_F_P, MGM_F_DPDZ));
    RUN(launch_fs_correct(Nd, dim, L.dbtag.p, F(MGM_F_UNEW), F(MGM_F_VNEW), F(MGM_F_WNEW), F(MGM_F_DPDX),
                          F(MGM_F_DPDY), F(MGM_F_DPDZ), dt, rho, F(MGM_F_U), F(MGM_F_V), F(MGM_F_W), s));
    RUN(spmv(pb, OP_DX, MGM_F_U, MGM_F_DPDX));
    RUN(spmv(pb, OP_DY, MGM_F_V, MGM_F_DPDY));
    // steady-state error (:63-69)
    return enqueue_error(pb, MGM_F_UOLD, MGM_F_U, MGM_F_VOLD, MGM_F_V, MGM_F_WOLD, MGM_F_W);
}

// One TIMPLE step (hf/timple_solver.h:40-70), error into scal[0].
int enqueue_timple(mgm_problem *pb, int precision)
{
    Level &L = pb->lv[0];
    hipStream_t s = stream();
    const int Nd = L.Nd, dim = pb->cfg.dimension;  // device rows, padding included
    auto F = [&](int k) { return L.f[k].p; };
    const double dt = pb->ns.dt, rho = pb->ns.rho;
    NsDev d = ns_view(pb);
    // hf/timple_solver.h:43-48
    RUN(launch_copy(Nd, F(MGM_F_UNEW), F(MGM_F_UOLD), s));
    RUN(launch_copy(Nd, F(MGM_F_UNEW), F(MGM_F_U), s));
    RUN(launch_copy(Nd, F(MGM_F_VNEW), F(MGM_F_VOLD), s));
    RUN(launch_copy(Nd, F(MGM_F_VNEW), F(MGM_F_V), s));
    RUN(launch_copy(Nd, F(MGM_F_WNEW), F(MGM_F_WOLD), s));
    RUN(launch_copy(Nd, F(MGM_F_WNEW), F(MGM_F_W), s));
    RUN(launch_zero(Nd, L.dbtag.p, 0, F(MGM_F_PPRIME), s));
    for (int it = 0; it < pb->ns.iter_timple; it++) {
        RUN(launch_copy(Nd, F(MGM_F_UNEW), F(MGM_F_U), s));
        RUN(launch_copy(Nd, F(MGM_F_VNEW), F(MGM_F_V), s));
        RUN(launch_copy(Nd, F(MGM_F_WNEW), F(MGM_F_W), s));
        RUN(launch_zero(Nd, L.dbtag.p, 0, F(MGM_F_PPRIME), s));
        // calculate_intermediate_velocity_implicit (:72-164)
        RUN(spmv(pb, OP_DX, MGM_F_POLD, MGM_F_DPDX));
        RUN(spmv(pb, OP_DY, MGM_F_POLD, MGM_F_DPDY));
        if (dim == 3) RUN(spmv(pb, OP_DZ, MGM_F_POLD, MGM_F_DPDZ));
        for (int k = 0; k < pb->ns.iter_momentum; k++) {
            if (pb->cfg.sweep_order == MGM_ORDER_NATURAL) {
                RUN(launch_momentum_natural(d, L.N, L.d_iperm.p, s));
            } else {
                for (int c = 0; c < L.ncol; c++)
                    RUN(launch_momentum_slices(d, L.d_col_slices.p + L.col_soff[c], L.col_soff[c + 1] - L.col_soff[c], s));
            }
        }
        // calculate_mass_residual_implicit (:166-189)
        RUN(spmv(pb, OP_DX, MGM_F_UNEW, MGM_F_DPDX));
        RUN(spmv(pb, OP_DY, MGM_F_VNEW, MGM_F_DPDY));
        if (dim == 3) RUN(spmv(pb, OP_DZ, MGM_F_WNEW, MGM_F_DPDZ));
        RUN(launch_mass_source(Nd, L.dbtag.p, F(MGM_F_DPDX), F(MGM_F_DPDY), F(MGM_F_DPDZ), rho, dt, 1,
                               F(MGM_F_SOURCE), s));
        // multigrid_Poisson_solver (:191-209)
        RUN(pressure_cycles(pb, MGM_VAR_T, precision));
        // update_velocity_implicit (:330-356)
        RUN(spmv(pb, OP_DX, MGM_F_PPRIME, MGM_F_DPDX));
        RUN(spmv(pb, OP_DY, MGM_F_PPRIME, MGM_F_DPDY));
        if (dim == 3) RUN(spmv(pb, OP_DZ, MGM_F_PPRIME, MGM_F_DPDZ));
        RUN(launch_timple_correct(Nd, dim, L.dbtag.p, F(MGM_F_DPDX), F(MGM_F_DPDY), F(MGM_F_DPDZ),
                                  F(MGM_F_PPRIME), dt, rho, F(MGM_F_UNEW), F(MGM_F_VNEW), F(MGM_F_WNEW),
                                  F(MGM_F_POLD), s));
        RUN(launch_pin(Nd, L.dbtag.p, L.ref0.p, F(MGM_F_POLD), pb->scal.p + 8, s));
        // update_boundary_pressure (:358-410)
        RUN(launch_timple_boundary_pressure(d, Nd, s));
    }
    // steady-state error (:63-69)
    return enqueue_error(pb, MGM_F_UOLD, MGM_F_UNEW, MGM_F_VOLD, MGM_F_VNEW, MGM_F_WOLD, MGM_F_WNEW);
}

int check_step(mgm_problem *pb)
{
    if (!pb || !pb->ready) MGM_FAIL(-1, "problem is not set up");
    if (!pb->ns_set) MGM_FAIL(-1, "mgm_ns_set_params has not been called");
    return 0;
}

}  // namespace

extern "C" {

int mgm_ns_set_params(mgm_problem *pb, const mgm_ns_params *ns)
{
    if (!pb || !ns) MGM_FAIL(-1, "null argument");
    if (!(ns->dt > 0) || !(ns->rho > 0)) MGM_FAIL(-1, "ns params: dt and rho must be positive");
    pb->ns = *ns;
    pb->ns_set = true;
    return 0;
}

int mgm_fs_step(mgm_problem *pb, double *steady_error)
{
    RUN(check_step(pb));
    RUN(enqueue_fs(pb, 0));
    return read_error(pb, steady_error);
}

int mgm_timple_step(mgm_problem *pb, double *steady_error)
{
    RUN(check_step(pb));
    RUN(enqueue_timple(pb, 0));
    return read_error(pb, steady_error);
}

int mgm_ns_march(mgm_problem *pb, int scheme, int precision, int max_steps, double steady_tol,
                 double *errors, int *steps_done, double *seconds)
{
    RUN(check_step(pb));
    if (scheme != MGM_VAR_F && scheme != MGM_VAR_T) MGM_FAIL(-1, "ns_march: scheme must be MGM_VAR_F or MGM_VAR_T");
    if (precision != 0 && precision != 1) MGM_FAIL(-1, "ns_march: precision must be 0 or 1");
    if (precision && pb->cfg.sweep_order != MGM_ORDER_COLOUR)
        MGM_FAIL(-1, "ns_march: the FP32 pressure cycle needs the multicolour sweep order");
    if (max_steps < 0) MGM_FAIL(-1, "ns_march: max_steps < 0");
    if (precision) RUN(mp_prepare_ns(pb));
    auto step = [&]() { return scheme == MGM_VAR_F ? enqueue_fs(pb, precision) : enqueue_timple(pb, precision); };
    // MGM_GRAPH=0 asks for plain stream launches (profilers that do not follow graphs)
    const char *ge = getenv("MGM_GRAPH");
    const bool use_graph = !(ge && ge[0] == '0');
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto cleanup = [&]() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    };
    if (use_graph && max_steps > 0) {
        HIP_TRY(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
        int rc = step();
        hipError_t ce = hipStreamEndCapture(stream(), &graph);
        if (rc) { cleanup(); return rc; }
        if (ce != hipSuccess) { cleanup(); MGM_FAIL(-10, std::string("graph capture failed: ") + hipGetErrorString(ce)); }
        hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ie != hipSuccess) { cleanup(); MGM_FAIL(-10, std::string("graph instantiate failed: ") + hipGetErrorString(ie)); }
    }
    HIP_TRY(hipStreamSynchronize(stream()));
    const double t0 = wall();
    int done = 0, rc = 0;
    for (int it = 0; it < max_steps; it++) {
        if (use_graph) {
            hipError_t le = hipGraphLaunch(exec, stream());
            if (le != hipSuccess) { rc = -10; set_error(hipGetErrorString(le)); break; }
        } else if ((rc = step())) {
            break;
        }
        double err = 0;
        if ((rc = read_error(pb, &err))) break;
        if (errors) errors[it] = err;
        done = it + 1;
        if (err < steady_tol) break;  // mg_NS_solver.c:80-83
        if (std::isnan(err)) { rc = -11; set_error("time march diverged (NaN steady-state error)"); break; }
    }
    const double t1 = wall();
    cleanup();
    if (steps_done) *steps_done = done;
    if (seconds) *seconds = t1 - t0;
    return rc;
}

}  // extern "C"
