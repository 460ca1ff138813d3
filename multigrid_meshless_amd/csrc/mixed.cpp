// mixed.cpp -- the reduced-precision inner V-cycle of mgm_solve's defect
// correction.  The FP64 operators of every level are narrowed once to FP32
// (precision 1) or, on levels >= 1, to BF16 (precision 2; level 0 stays
// FP32); vectors, products and sums are FP32.  The cycle is the H variant of
// the FP64 one (hf/heat_conduction.h:25-112 order and boundary handling) in
// multicolour order, started from a zero correction.  It has no reference
// counterpart: the reference solves in FP64 only.
//
// The time steps' pressure solve has an FP32 form too (op_correction_ns_mp):
// each of its V-cycles becomes one FP64 defect correction by an FP32 cycle
// of the step's own variant (T: hf/timple_solver.h:235-328, F:
// hf/fractional_step_solver.h:168-246), applied to the correction e.
#include "problem.hpp"

namespace mgm {

bool learned_active(const mgm_problem *pb, int l, int variant);
int small_level_rows();
int op_learned_sweep_mp(mgm_problem *pb, int l, int (*residual_mp)(Level &));

namespace {

// A level's narrowed operator: gather-sorted where the level has the FP64
// slab sorted (EllT::sorted; plain is the diagonal source), else plain; with
// the level's LDS-window tables where it has them.
template <typename T>
EllT<T> narrowed(const Level &L, const DevBuf<T> &plain, const DevBuf<T> &sorted)
{
    const bool st = stream_slab(L.Nd, L.m, sizeof(T));
    if (L.gs && sorted.p)
        return L.with_window(EllT<T>{L.gcol.p, sorted.p, L.m, L.S, st, plain.p, true, 32 - slot_bits(L.S, L.m)});
    return L.with_window(EllT<T>{L.col.p, plain.p, L.m, L.S, st});
}
template <typename T>
EllT<T> narrowed_prol(const Level &F, const DevBuf<T> &plain, const DevBuf<T> &sorted)
{
    const bool st = stream_slab(F.Nd, F.mp, sizeof(T));
    // (no LDS window: at 4- and 2-byte weights the window pass is bound by its
    // instructions, not bytes, and measured no faster than the gathering one)
    if (F.gsp && sorted.p) return EllT<T>{F.gpcol.p, sorted.p, F.mp, F.Sp, st, plain.p, true,
                                          32 - slot_bits(F.Sp, F.mp)};
    return EllT<T>{F.pcol.p, plain.p, F.mp, F.Sp, st};
}

// A level's narrowed operator in the plain layout (the single-workgroup relax
// reads it so).
template <typename T>
EllT<T> plain_op(const Level &L, const DevBuf<T> &w)
{
    return EllT<T>{L.col.p, w.p, L.m, L.S, stream_slab(L.Nd, L.m, sizeof(T))};
}

// Small levels relax in one single-workgroup launch, as in the FP64 cycle.
bool small_relax(const Level &L)
{
    return L.Nd <= small_level_rows() && relax_small_fits(L.Nd, L.m, L.m, L.S, sizeof(float));
}

template <typename F>
int with_op(const Level &L, F &&f)
{
    if (L.w16) return f(narrowed(L, L.lapp16, L.glapp16));
    return f(narrowed(L, L.lapp32, L.glapp32));
}

int residual_mp(Level &L);

// With the learned smoother active, its FP32 correction takes the place of
// the first sweep.
// from_zero: p32 is +0 on every row that is not padding (the descent of a
// correction), so a first Jacobi sweep needs no stencil pass
// (launch_jacobi_zero, bitwise the full pass; MGM_JACOBI_ZERO=0 runs the pass).
static bool jacobi_zero_on()
{
    const char *e = getenv("MGM_JACOBI_ZERO");
    return !(e && e[0] == '0');
}
int relax_mp(mgm_problem *pb, int l, bool from_zero)
{
    Level &L = pb->lv[l];
    const float w = pb->omega(l);
    const float c = 1.0f - w;
    int it0 = 0;
    if (learned_active(pb, l, MGM_VAR_H) && pb->cfg.num_relax > 0) {
        int rc = op_learned_sweep_mp(pb, l, residual_mp);
        if (rc) return rc;
        it0 = 1;
    }
    if (pb->jacobi(l, MGM_VAR_H))
        return with_op(L, [&](auto A) {
            const float *in = L.p32.p;
            int it = it0;
            if (from_zero && it0 == 0 && pb->cfg.num_relax > 0 && jacobi_zero_on()) {
                int rc = launch_jacobi_zero(A, L.Nd, L.dbtag.p, (const float *)L.src32.p, L.res32.p, w, c, stream());
                if (rc) return rc;
                in = L.res32.p;
                it = 1;
            }
            const int n = pb->cfg.num_relax - it;
            const bool from_p = in == L.p32.p;
            float *const *seq = jacobi_targets(n, from_p, L.p32.p, L.res32.p, L.jtmp32.p);
            for (int k = 0; k < n; k++) {
                int rc = launch_jacobi(A, L.Nd, L.dbtag.p, (const float *)L.src32.p, in, seq[k], w, c, stream());
                if (rc) return rc;
                in = seq[k];
            }
            return in == L.p32.p ? 0 : launch_stream_copy((long)L.Nd * 4 / 16, in, L.p32.p, stream());
        });
    if (small_relax(L) && pb->cfg.num_relax > it0) {
        const int n = pb->cfg.num_relax - it0;
        if (L.w16) {
            const EllT<bf16> A = plain_op(L, L.lapp16);
            return launch_relax_small(A, A, L.d_col_slices.p, L.d_col_soff.p, L.ncol, L.Nd, L.dbtag.p, L.ref0.p,
                                      (const float *)L.src32.p, L.p32.p, L.res32.p, w, c, n, MGM_VAR_H, stream());
        }
        const EllT<float> A = plain_op(L, L.lapp32);
        return launch_relax_small(A, A, L.d_col_slices.p, L.d_col_soff.p, L.ncol, L.Nd, L.dbtag.p, L.ref0.p,
                                  (const float *)L.src32.p, L.p32.p, L.res32.p, w, c, n, MGM_VAR_H, stream());
    }
    return with_op(L, [&](auto A) {
        for (int it = it0; it < pb->cfg.num_relax; it++)
            for (int k = 0; k < L.ncol; k++) {
                const int *list;
                int first, n;
                L.colour_slices(k, &list, &first, &n);
                int rc = launch_sor_slices(A, list, first, n, L.dbtag.p, (const float *)L.src32.p, L.p32.p, w, c,
                                           stream());
                if (rc) return rc;
            }
        return 0;
    });
}

int residual_mp(Level &L)
{
    return with_op(L, [&](auto A) {
        return launch_residual(A, L.Nd, 0, L.dbtag.p, (const float *)L.src32.p, (const float *)L.p32.p,
                               L.res32.p, stream());
    });
}

int restrict_mp(Level &F, Level &C)
{
    if (C.w16)
        return launch_restrict(EllT<bf16>{C.rcol.p, C.rw16.p, C.mr, C.Sr, stream_slab(C.Nd, C.mr, 2)}, C.Nd, C.dbtag.p, 0,
                               (const float *)F.res32.p, C.src32.p, stream());
    return launch_restrict(EllT<float>{C.rcol.p, C.rw32.p, C.mr, C.Sr, stream_slab(C.Nd, C.mr, 4)}, C.Nd, C.dbtag.p, 0,
                           (const float *)F.res32.p, C.src32.p, stream());
}

int prolong_mp(Level &F, Level &C)
{
    int rc;
    if (F.w16)
        rc = launch_prolong(narrowed_prol(F, F.pw16, F.gpw16), F.Nd, F.dbtag.p, (const float *)C.p32.p, F.p32.p,
                            stream());
    else
        rc = launch_prolong(narrowed_prol(F, F.pw32, F.gpw32), F.Nd, F.dbtag.p, (const float *)C.p32.p, F.p32.p,
                            stream());
    if (rc) return rc;
    return launch_zero(C.Nd, C.dbtag.p, 1, C.p32.p, stream());
}

EllT<float> op32(const Level &L, const DevBuf<float> &w, const DevBuf<float> &gw)
{
    return narrowed(L, w, gw);
}

// Source of the lap residual forms of variant F: level 0 has its own
// (r_lap = s - lap p), coarse levels share the SOR source.
const float *srcl(const Level &L, int l) { return l == 0 ? L.srcl32.p : L.src32.p; }

// relaxation of variant T or F on the FP32 correction: SOR sweeps with
// lap_Poison, each followed by the pin and the variant's residual
int relax_ns_mp(mgm_problem *pb, int l, int variant)
{
    Level &L = pb->lv[l];
    const float w = pb->omega(l);
    const float c = 1.0f - w;
    // (level 0 of variant F has its own residual source: per-colour launches)
    if (small_relax(L) && !(variant == MGM_VAR_F && l == 0) && pb->cfg.num_relax > 0)
        return launch_relax_small(plain_op(L, L.lapp32), plain_op(L, L.lap32), L.d_col_slices.p, L.d_col_soff.p,
                                  L.ncol, L.Nd, L.dbtag.p, L.ref0.p, (const float *)L.src32.p, L.p32.p, L.res32.p,
                                  w, c, pb->cfg.num_relax, variant, stream());
    const EllT<float> A = op32(L, L.lapp32, L.glapp32);
    for (int it = 0; it < pb->cfg.num_relax; it++) {
        int rc;
        for (int k = 0; k < L.ncol; k++) {
            const int *list;
            int first, n;
            L.colour_slices(k, &list, &first, &n);
            if ((rc = launch_sor_slices(A, list, first, n, L.dbtag.p, (const float *)L.src32.p, L.p32.p, w, c,
                                        stream())))
                return rc;
        }
        if ((rc = launch_pin(L.Nd, L.dbtag.p, L.ref0.p, L.p32.p, pb->scal.p + 8, stream()))) return rc;
        if (variant == MGM_VAR_T)
            rc = launch_residual(A, L.Nd, 1, L.dbtag.p, (const float *)L.src32.p, (const float *)L.p32.p,
                                 L.res32.p, stream());
        else
            rc = launch_residual(op32(L, L.lap32, L.glap32), L.Nd, 2, L.dbtag.p, srcl(L, l), (const float *)L.p32.p,
                                 L.res32.p, stream());
        if (rc) return rc;
    }
    return 0;
}

int residual_ns_mp(mgm_problem *pb, int l, int variant)
{
    Level &L = pb->lv[l];
    if (variant == MGM_VAR_F)
        return launch_residual(op32(L, L.lap32, L.glap32), L.Nd, 3, L.dbtag.p, srcl(L, l), (const float *)L.p32.p,
                               L.res32.p, stream());
    return launch_residual(op32(L, L.lapp32, L.glapp32), L.Nd, 0, L.dbtag.p, (const float *)L.src32.p,
                           (const float *)L.p32.p, L.res32.p, stream());
}

template <typename T>
int narrow(const DevBuf<double> &src, DevBuf<T> &dst)
{
    HIP_TRY(dst.alloc(src.n));
    if constexpr (sizeof(T) == 4) return launch_narrow_f32((long)src.n, src.p, dst.p, stream());
    else return launch_narrow_bf16((long)src.n, src.p, dst.p, stream());
}

}  // namespace

// Build the narrowed operators and the FP32 fields for `precision` (1 or 2).
int mp_prepare(mgm_problem *pb, int precision)
{
    if (pb->mp_built == precision) return 0;
    for (int l = 0; l < pb->L; l++) {
        Level &L = pb->lv[l];
        L.w16 = (precision == 2 && l > 0);
        int rc;
        if (L.w16) {
            L.lapp32.release(); L.rw32.release(); L.pw32.release(); L.glapp32.release(); L.gpw32.release();
            if ((rc = narrow(L.slab[OP_LAPP], L.lapp16))) return rc;
            if (l > 0 && (rc = narrow(L.rw, L.rw16))) return rc;
            if (l + 1 < pb->L && (rc = narrow(L.pw, L.pw16))) return rc;
            if (L.gs && (rc = narrow(L.gslab[OP_LAPP], L.glapp16))) return rc;
            if (L.gsp && (rc = narrow(L.gpw, L.gpw16))) return rc;
        } else {
            L.lapp16.release(); L.rw16.release(); L.pw16.release(); L.glapp16.release(); L.gpw16.release();
            if ((rc = narrow(L.slab[OP_LAPP], L.lapp32))) return rc;
            if (l > 0 && (rc = narrow(L.rw, L.rw32))) return rc;
            if (l + 1 < pb->L && (rc = narrow(L.pw, L.pw32))) return rc;
            if (L.gs && (rc = narrow(L.gslab[OP_LAPP], L.glapp32))) return rc;
            if (L.gsp && (rc = narrow(L.gpw, L.gpw32))) return rc;
        }
        for (DevBuf<float> *f : {&L.p32, &L.src32, &L.res32, &L.jtmp32}) {
            if (f->n != (size_t)L.Nd) {
                HIP_TRY(f->alloc(L.Nd));
                HIP_TRY(hipMemsetAsync(f->p, 0, sizeof(float) * L.Nd, stream()));
            }
        }
    }
    HIP_TRY(hipStreamSynchronize(stream()));
    pb->mp_built = precision;
    return 0;
}

// One defect-correction step: e = V-cycle(A, r) in reduced precision from
// e = 0, then p += e.  RES of level 0 must hold r (FP64) on entry.
int op_correction_mp(mgm_problem *pb)
{
    Level &L0 = pb->lv[0];
    int rc;
    if ((rc = launch_narrow_f32(L0.Nd, L0.f[MGM_F_RES].p, L0.src32.p, stream()))) return rc;
    for (auto &L : pb->lv)
        if ((rc = launch_zero(L.Nd, L.dbtag.p, 0, L.p32.p, stream()))) return rc;
    const int nl = pb->L;
    for (int l = 0; l < nl; l++) {
        if ((rc = relax_mp(pb, l, true))) return rc;   // every level's p32 was just zeroed
        if ((rc = residual_mp(pb->lv[l]))) return rc;
        if (l != nl - 1 && (rc = restrict_mp(pb->lv[l], pb->lv[l + 1]))) return rc;
    }
    for (int l = nl - 1; l > 0; l--) {
        if ((rc = prolong_mp(pb->lv[l - 1], pb->lv[l]))) return rc;
        if (l != 1 && (rc = relax_mp(pb, l - 1, false))) return rc;
    }
    return launch_add_correction(L0.Nd, L0.dbtag.p, L0.p32.p, L0.f[MGM_F_P].p, stream());
}

// FP32 copies for the time steps' pressure cycle: those of precision 1 plus
// lap on every level and level 0's second source.
int mp_prepare_ns(mgm_problem *pb)
{
    int rc;
    if ((rc = mp_prepare(pb, 1))) return rc;
    if (pb->mp_ns_built) return 0;
    for (auto &L : pb->lv) {
        if ((rc = narrow(L.slab[OP_LAP], L.lap32))) return rc;
        if (L.gs && L.gslab[OP_LAP].p && (rc = narrow(L.gslab[OP_LAP], L.glap32))) return rc;
    }
    Level &L0 = pb->lv[0];
    if (L0.srcl32.n != (size_t)L0.Nd) HIP_TRY(L0.srcl32.alloc(L0.Nd));
    HIP_TRY(hipStreamSynchronize(stream()));
    pb->mp_ns_built = true;
    return 0;
}

// One pressure V-cycle of a time step (variant T on p', F on p) as an FP64
// defect correction: p is pinned in FP64, the residuals of p that the cycle's
// level-0 forms need are formed in FP64 on every row (lap_Poison for the SOR
// source; for F also lap), narrowed to FP32, and the FP32 cycle of the same
// variant runs on the correction e from zero; then p += e on every row.
// Relaxation, residual, restriction and prolongation are linear in (e, source)
// and the pin commutes with a constant shift (rows of lap_Poison and lap sum
// to zero), so in exact arithmetic p + e is what the FP64 cycle produces.
int op_correction_ns_mp(mgm_problem *pb, int variant)
{
    Level &L0 = pb->lv[0];
    double *p = L0.f[variant == MGM_VAR_T ? MGM_F_PPRIME : MGM_F_P].p;
    const double *s = L0.f[MGM_F_SOURCE].p;
    double *res = L0.f[MGM_F_RES].p;
    int rc;
    if ((rc = launch_pin(L0.Nd, L0.dbtag.p, L0.ref0.p, p, pb->scal.p + 8, stream()))) return rc;
    if ((rc = launch_residual(L0.op(OP_LAPP), L0.Nd, 2, L0.dbtag.p, s, (const double *)p, res, stream()))) return rc;
    if ((rc = launch_narrow_f32(L0.Nd, res, L0.src32.p, stream()))) return rc;
    if (variant == MGM_VAR_F) {
        if ((rc = launch_residual(L0.op(OP_LAP), L0.Nd, 2, L0.dbtag.p, s, (const double *)p, res, stream())))
            return rc;
        if ((rc = launch_narrow_f32(L0.Nd, res, L0.srcl32.p, stream()))) return rc;
    }
    // e = 0 on level 0.  Coarse levels start, as in the FP64 cycle, from what
    // the last prolongation left: zero for T; for F zero interior rows and the
    // boundary rows' values of the previous cycle, since F's prolongation zeroes
    // only interior coarse rows (hf/fractional_step_solver.h:229-233).
    const int keep = variant == MGM_VAR_F;
    for (int l = 0; l < pb->L; l++) {
        Level &L = pb->lv[l];
        if ((rc = launch_zero(L.Nd, L.dbtag.p, l > 0 && keep, L.p32.p, stream()))) return rc;
    }
    const int nl = pb->L;
    for (int l = 0; l < nl; l++) {
        if ((rc = relax_ns_mp(pb, l, variant))) return rc;
        if ((rc = residual_ns_mp(pb, l, variant))) return rc;
        if (l != nl - 1) {
            Level &F = pb->lv[l], &C = pb->lv[l + 1];
            if ((rc = launch_restrict(EllT<float>{C.rcol.p, C.rw32.p, C.mr, C.Sr, stream_slab(C.Nd, C.mr, 4)}, C.Nd,
                                      C.dbtag.p, keep, (const float *)F.res32.p, C.src32.p, stream())))
                return rc;
        }
    }
    for (int l = nl - 1; l > 0; l--) {
        Level &F = pb->lv[l - 1], &C = pb->lv[l];
        if ((rc = launch_prolong(narrowed_prol(F, F.pw32, F.gpw32), F.Nd, F.dbtag.p, (const float *)C.p32.p,
                                 F.p32.p, stream())))
            return rc;
        if ((rc = launch_zero(C.Nd, C.dbtag.p, variant != MGM_VAR_T, C.p32.p, stream()))) return rc;
        if (l != 1 && (rc = relax_ns_mp(pb, l - 1, variant))) return rc;
    }
    return launch_add_correction(L0.Nd, L0.dbtag.p, L0.p32.p, p, stream());
}

}  // namespace mgm
