/*
 * ref_harness.c -- exports the compiled reference's state for golden fixtures.
 *
 * TEST INFRASTRUCTURE ONLY.  Compiled by `make -C oracle ref` against the
 * reference headers where they lie (-I <reference tree>); it copies nothing.
 * One build per solver variant, because the reference's solver headers define
 * conflicting symbols (relaxation, calculate_residuals, ...):
 *   -DREF_HEAT   header_files/heat_conduction.h + init/init_heat_conduction.c
 *   -DREF_TIMPLE header_files/timple_solver.h   + init/init_LDC.c
 *   -DREF_FS     header_files/fractional_step_solver.h + init/init_LDC.c
 * The reference's teardown (free_PointStructure) faults, so nothing is freed.
 */
#if defined(REF_HEAT)
#include "header_files/heat_conduction.h"
#include "init/init_heat_conduction.c"
#elif defined(REF_TIMPLE)
#include "header_files/timple_solver.h"
#include "init/init_LDC.c"
#elif defined(REF_FS)
#include "header_files/fractional_step_solver.h"
#include "init/init_LDC.c"
#else
#error "define REF_HEAT, REF_TIMPLE or REF_FS"
#endif
#include <string.h>
#include <unistd.h>

static PointStructure *P = 0;
static FieldVariables *F = 0;
static short L = 0;

/* Mirrors the setup sequence of mg_heat_conduction.c:26-45 / mg_NS_solver.c:27-51. */
int ref_setup(const char *workdir, int build_weights)
{
    if (chdir(workdir) != 0) return -1;
    setvbuf(stdout, NULL, _IOLBF, 0);
    read_flow_parameters("flow_parameters.csv");
    read_grid_filenames(&P, "grid_filenames.csv", &parameters.num_levels);
    L = parameters.num_levels;
    read_complete_mesh_data(P, L);
    if (build_weights)
        for (int ii = 0; ii < L; ii++) create_derivative_matrices(&P[ii]);
    AllocateMemoryFieldVariables(&F, P, L);
    return L;
}

/* info: N, nb, m, n, poly_degree, dimension */
void ref_info(int lev, int *info, double *d_avg)
{
    info[0] = P[lev].num_nodes; info[1] = P[lev].num_boundary_nodes;
    info[2] = P[lev].num_cloud_points; info[3] = P[lev].num_poly_terms;
    info[4] = P[lev].poly_degree; info[5] = parameters.dimension;
    *d_avg = P[lev].d_avg;
}

void ref_geometry(int lev, double *x, double *y, double *z, double *nx, double *ny, double *nz,
                  unsigned char *btag)
{
    for (int i = 0; i < P[lev].num_nodes; i++) {
        x[i] = P[lev].x[i]; y[i] = P[lev].y[i]; z[i] = P[lev].z[i];
        nx[i] = P[lev].x_normal[i]; ny[i] = P[lev].y_normal[i]; nz[i] = P[lev].z_normal[i];
        btag[i] = P[lev].boundary_tag[i];
    }
}

void ref_cloud(int lev, int *out)
{
    int m = P[lev].num_cloud_points;
    for (int i = 0; i < P[lev].num_nodes; i++)
        for (int j = 0; j < m; j++) out[(size_t)i * m + j] = P[lev].cloud_index[i][j];
}

/* which: 0 restriction_points, 1 prolongation_points */
void ref_points(int lev, int which, int *out)
{
    int *src = which ? P[lev].prolongation_points : P[lev].restriction_points;
    for (int i = 0; i < P[lev].num_nodes; i++) out[i] = src[i];
}

/* which: 0 Dx 1 Dy 2 Dz 3 lap 4 lap_Poison 5 restr_mat 6 prol_mat; width = row length */
int ref_matrix(int lev, int which, int width, double *out)
{
    double **M = 0;
    switch (which) {
    case 0: M = P[lev].Dx; break;
    case 1: M = P[lev].Dy; break;
    case 2: M = P[lev].Dz; break;
    case 3: M = P[lev].lap; break;
    case 4: M = P[lev].lap_Poison; break;
    case 5: M = P[lev].restr_mat; break;
    case 6: M = P[lev].prol_mat; break;
    }
    if (!M) return -1;
    for (int i = 0; i < P[lev].num_nodes; i++)
        for (int j = 0; j < width; j++) out[(size_t)i * width + j] = M[i][j];
    return 0;
}

static double *field_ptr(int lev, int which)
{
    switch (which) {
    case 0: return F[lev].p;      case 1: return F[lev].source; case 2: return F[lev].res;
    case 3: return F[lev].pprime; case 4: return F[lev].u;      case 5: return F[lev].v;
    case 6: return F[lev].w;      case 7: return F[lev].u_new;  case 8: return F[lev].v_new;
    case 9: return F[lev].w_new;  case 10: return F[lev].p_old; case 11: return F[lev].T;
    case 12: return F[lev].dpdn;  case 13: return F[lev].dpdx;  case 14: return F[lev].dpdy;
    case 15: return F[lev].dpdz;  case 16: return F[lev].u_old; case 17: return F[lev].v_old;
    case 18: return F[lev].w_old;
    }
    return 0;
}

void ref_get_field(int lev, int which, double *out)
{
    double *f = field_ptr(lev, which);
    for (int i = 0; i < P[lev].num_nodes; i++) out[i] = f[i];
}

void ref_set_field(int lev, int which, const double *in)
{
    double *f = field_ptr(lev, which);
    for (int i = 0; i < P[lev].num_nodes; i++) f[i] = in[i];
}

/* Overwrite lap_Poison (e.g. identity rows for Dirichlet problems). */
void ref_set_lap_poisson(int lev, const double *in)
{
    int m = P[lev].num_cloud_points;
    for (int i = 0; i < P[lev].num_nodes; i++)
        for (int j = 0; j < m; j++) P[lev].lap_Poison[i][j] = in[(size_t)i * m + j];
}

void ref_set_omega_relax(float omega, int num_relax, int num_vcycles)
{
    parameters.omega = omega;
    parameters.num_relax = num_relax;
    parameters.num_vcycles = num_vcycles;
}

void ref_relaxation(int lev) { relaxation(&P[lev], &F[lev]); }
void ref_calculate_residuals(int lev) { calculate_residuals(&P[lev], &F[lev]); }
void ref_restrict(int lev) { restrict_residuals(&P[lev], &P[lev + 1], &F[lev], &F[lev + 1]); }
void ref_prolong(int lev) { prolongate_corrections(&P[lev - 1], &P[lev], &F[lev - 1], &F[lev]); }

/* One V-cycle exactly as the drivers order it (mg_heat_conduction.c:64-75;
 * multigrid_Poisson_solver runs num_vcycles of the same loop). */
void ref_vcycle(void)
{
    for (int ilev = 0; ilev < L; ilev++) {
        relaxation(&P[ilev], &F[ilev]);
        calculate_residuals(&P[ilev], &F[ilev]);
        if (ilev != L - 1) restrict_residuals(&P[ilev], &P[ilev + 1], &F[ilev], &F[ilev + 1]);
    }
    for (int ilev = L - 1; ilev > 0; ilev--) {
        prolongate_corrections(&P[ilev - 1], &P[ilev], &F[ilev - 1], &F[ilev]);
        if (ilev != 1) relaxation(&P[ilev - 1], &F[ilev - 1]);
    }
}

#if defined(REF_HEAT)
void ref_heat_init(int k)
{
    initial_conditions(P, F, L, k);
    boundary_conditions(P, F, L, k);
}
void ref_heat_set_boundary_rhs(int k) { set_boundary_rhs(&P[0], F, L, k); }
#else
void ref_ns_init(void)
{
    initial_conditions(P, F, 1);
    boundary_conditions(P, F, 1);
    parameters.dt = calculate_dt(&P[0]);
}
double ref_dt(void) { return parameters.dt; }
/* d: dt, rho, mu, nu, facRe, facdt, F[0].rho; i: iter_momentum, iter_timple, num_vcycles, num_relax */
void ref_ns_params(double *d, int *i, float *omega)
{
    d[0] = parameters.dt; d[1] = parameters.rho; d[2] = parameters.mu; d[3] = parameters.nu;
    d[4] = parameters.facRe; d[5] = parameters.facdt; d[6] = F[0].rho;
    i[0] = parameters.iter_momentum; i[1] = parameters.iter_timple; i[2] = parameters.num_vcycles;
    i[3] = parameters.num_relax;
    *omega = parameters.omega;
}
void ref_mg_poisson(void) { multigrid_Poisson_solver(P, F); }
#endif

#if defined(REF_TIMPLE)
double ref_step(void) { return time_implicit_solver(P, F); }
#elif defined(REF_FS)
double ref_step(void) { return fractional_step_explicit(P, F); }
#endif

/* The reference's config readers on a file of the caller's (read_flow_parameters,
 * hf/mesh_functions.h:37-100; read_grid_filenames, :102-129).  ref_params copies
 * the global `parameters` the parse left (the struct layout of
 * hf/structures.h:15-41, which the caller mirrors); ref_grid_filenames returns
 * the level count and writes the names, 50 bytes each, into names. */
void ref_params(const char *path, struct parameters *out)
{
    memset(&parameters, 0, sizeof(parameters));
    read_flow_parameters((char *)path);
    *out = parameters;
}

int ref_grid_filenames(const char *path, char *names, int max_levels)
{
    PointStructure *G = 0;
    short n = 0;
    read_grid_filenames(&G, (char *)path, &n);
    for (int l = 0; l < n && l < max_levels; l++) memcpy(names + 50 * l, G[l].mesh_filename, 50);
    free(G);
    return n;
}
