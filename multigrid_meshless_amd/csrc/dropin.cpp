// dropin.cpp -- the reference's C call surface (include/header_files/*.h)
// implemented on libmgm, so mg_heat_conduction.c and mg_NS_solver.c compile
// and link unchanged.
//
// The reference keeps everything in host arrays hung off PointStructure /
// FieldVariables; here the hierarchy is resident on the GPU and the host
// arrays are mirrors.  Each solver call uploads what it reads (fields; the
// lap_Poison rows only when their bytes changed, since init plug-ins rewrite
// boundary rows between V-cycles -- init/init_heat_conduction.c:50-65) and
// downloads what it writes.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/header_files/structures.h"
#include "../../include/header_files/mgm_dropin.h"
#include "problem.hpp"

extern "C" {
struct parameters parameters;   // the one definition (reference: a tentative definition per TU)
}

using namespace mgm;

namespace {

struct DropIn {
    PointStructure *base = nullptr;
    int L = 0;
    mgm_problem *pb = nullptr;
    double multiplier = 0;            // fractional multiplier from flow_parameters.csv
    std::vector<uint64_t> lapP_hash;  // fingerprint of the host lap_Poison rows at the last upload
};
DropIn g;

void die(const char *what)
{
    fprintf(stderr, "libmgm: %s failed: %s\n", what, mgm_last_error());
    exit(1);
}
#define MUST(expr) do { if ((expr) != 0) die(#expr); } while (0)

int level_of(const PointStructure *P)
{
    if (!g.base || P < g.base || P >= g.base + g.L) {
        fprintf(stderr, "libmgm: PointStructure %p was not set up by read_complete_mesh_data\n", (void *)P);
        exit(1);
    }
    return (int)(P - g.base);
}

// Fingerprint of the host lap_Poison rows the init plug-ins may rewrite
// between solver calls: the boundary rows (set_boundary_rhs,
// init/init_heat_conduction.c:50-65, writes those).  Hashed a 64-bit word at
// a time, so at 1M points and m = 70 a call reads the 57k boundary rows (32
// MB) instead of every row.  MGM_DROPIN_CHECK_ALL_ROWS=1 fingerprints every
// row, for plug-ins that also rewrite interior rows.
bool check_all_rows()
{
    static const bool all = [] {
        const char *e = getenv("MGM_DROPIN_CHECK_ALL_ROWS");
        return e && e[0] == '1';
    }();
    return all;
}

uint64_t hash_rows(double **rows, const PointStructure &P)
{
    const int m = P.num_cloud_points;
    const bool all = check_all_rows();
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < P.num_nodes; i++) {
        if (!all && !P.boundary_tag[i]) continue;
        const double *r = rows[i];
        for (int j = 0; j < m; j++) {
            uint64_t w;
            memcpy(&w, &r[j], sizeof(w));
            h = (h ^ w) * 1099511628211ull;
            h ^= h >> 29;
        }
        h = (h ^ (uint64_t)i) * 1099511628211ull;
    }
    return h;
}

std::vector<double> flatten(double **rows, int N, int m)
{
    std::vector<double> v((size_t)N * m);
    for (int i = 0; i < N; i++) memcpy(&v[(size_t)i * m], rows[i], sizeof(double) * m);
    return v;
}

void sync_params()
{
    g.pb->cfg.omega = parameters.omega;
    g.pb->cfg.num_relax = parameters.num_relax;
}

void upload_lapP_if_changed(int l)
{
    PointStructure &P = g.base[l];
    if (!P.lap_Poison) return;
    uint64_t h = hash_rows(P.lap_Poison, P);
    if (h == g.lapP_hash[l]) return;
    std::vector<double> v = flatten(P.lap_Poison, P.num_nodes, P.num_cloud_points);
    MUST(mgm_set_matrix(g.pb, l, MGM_MAT_LAPP, v.data()));
    g.lapP_hash[l] = h;
}

void put(int l, int which, const double *h) { MUST(mgm_set_field(g.pb, l, which, h)); }
void get(int l, int which, double *h) { MUST(mgm_get_field(g.pb, l, which, h)); }

double **rows_from(const double *flat, int N, int m)
{
    double **r = (double **)malloc(sizeof(double *) * N);
    for (int i = 0; i < N; i++) {
        r[i] = (double *)malloc(sizeof(double) * m);
        memcpy(r[i], flat + (size_t)i * m, sizeof(double) * m);
    }
    return r;
}

}  // namespace

extern "C" {

// ---- structures.h -----------------------------------------------------------

void AllocateMemoryPointStructure(PointStructure *s, int nodes)
{
    s->x_normal = (double *)malloc(nodes * sizeof(double));
    s->y_normal = (double *)malloc(nodes * sizeof(double));
    s->z_normal = (double *)malloc(nodes * sizeof(double));
    s->x = (double *)malloc(nodes * sizeof(double));
    s->y = (double *)malloc(nodes * sizeof(double));
    s->z = (double *)malloc(nodes * sizeof(double));
    s->point_index = (int *)malloc(nodes * sizeof(int));
    s->boundary_tag = (bool *)malloc(nodes * sizeof(bool));
    s->corner_tag = (bool *)malloc(nodes * sizeof(bool));
    s->num_nodes = nodes;
    s->num_elem = 0;
}

void AllocateMemoryFieldVariables(FieldVariables **field, PointStructure *P, int num_levels)
{
    *field = (FieldVariables *)calloc(num_levels, sizeof(FieldVariables));
    for (int l = 0; l < num_levels; l++) {
        int N = P[l].num_nodes;
        FieldVariables &F = (*field)[l];
        double **slots[] = {&F.u, &F.v, &F.w, &F.u_new, &F.v_new, &F.w_new, &F.pprime, &F.p,
                            &F.u_old, &F.v_old, &F.w_old, &F.p_old, &F.res, &F.source, &F.dpdn,
                            &F.dpdx, &F.dpdy, &F.dpdz, &F.T};
        for (double **s : slots) *s = (double *)calloc(N, sizeof(double));
    }
}

static void free_rows(void **rows, int n)
{
    if (!rows) return;
    for (int i = 0; i < n; i++) free(rows[i]);
    free(rows);
}

void free_PointStructure(PointStructure *P, int num_levels)
{
    if (!P) return;
    for (int l = 0; l < num_levels; l++) {
        PointStructure &s = P[l];
        free(s.x); free(s.y); free(s.z); free(s.point_index);
        free(s.x_normal); free(s.y_normal); free(s.z_normal);
        free(s.boundary_tag); free(s.corner_tag); free(s.pow_x); free(s.pow_y); free(s.pow_z);
        int N = s.num_nodes;
        free_rows((void **)s.cloud_index, N);
        free_rows((void **)s.Dx, N); free_rows((void **)s.Dy, N); free_rows((void **)s.Dz, N);
        free_rows((void **)s.lap, N); free_rows((void **)s.lap_Poison, N);
        free_rows((void **)s.restr_mat, N); free_rows((void **)s.prol_mat, N);
        free(s.prolongation_points); free(s.restriction_points);
    }
    if (P == g.base) {
        mgm_problem_destroy(g.pb);
        g = DropIn();
    }
    free(P);
}

void free_field(FieldVariables *F, int num_levels)
{
    if (!F) return;
    for (int l = 0; l < num_levels; l++) {
        FieldVariables &f = F[l];
        double *all[] = {f.u, f.v, f.w, f.u_new, f.v_new, f.w_new, f.pprime, f.p, f.u_old, f.v_old,
                         f.w_old, f.p_old, f.res, f.source, f.dpdn, f.dpdx, f.dpdy, f.dpdz, f.T};
        for (double *a : all) free(a);
    }
    free(F);
}

// ---- mat_lib.h (allocation helpers, same semantics as the reference) ------

void create_matrix(double ***A, int n_rows, int n_cols)
{
    *A = (double **)malloc(n_rows * sizeof(double *));
    for (int i = 0; i < n_rows; i++) (*A)[i] = (double *)calloc(n_cols, sizeof(double));
}
void create_matrix_int(int ***A, int n_rows, int n_cols)
{
    *A = (int **)malloc(n_rows * sizeof(int *));
    for (int i = 0; i < n_rows; i++) (*A)[i] = (int *)malloc(n_cols * sizeof(int));
}
double **create_matrix1(int n_rows, int n_cols)
{
    double **A;
    create_matrix(&A, n_rows, n_cols);
    return A;
}
double *create_vector(int n) { return (double *)calloc(n, sizeof(double)); }
void free_matrix(double **A, int n_rows) { free_rows((void **)A, n_rows); }
void free_matrix_int(int **A, int n_rows) { free_rows((void **)A, n_rows); }
void free_vector(double *A) { free(A); }

// ---- mesh_functions.h ------------------------------------------------------

// read_flow_parameters (hf/mesh_functions.h:37-100): positional "name,value"
// lines read with the reference's own conversions (%hd, %lf, %f, %d), so every
// field of `parameters` holds what the reference's parse leaves there.  One
// extension: the cloud-size multiplier is read with %lf and kept whole for
// the stencil size (3.5 -> m = 70 at 20 monomials); parameters.
// cloud_size_multiplier gets its integer part, which is what the reference's
// %hd leaves in it (it stops at the '.', and the ".5" then becomes part of the
// next line's ignored name).  A line that does not parse ends the program
// (the reference checks only the last line).
void read_flow_parameters(char *filename)
{
    FILE *f = fopen(filename, "r");
    if (!f) { printf("Error: Unable to open the file\n"); exit(1); }
    char name[128];
    int line = 0;
    auto bad = [&]() {
        printf("Error: Unable to read the file (line %d)\n", line);
        exit(1);
    };
    auto rd_short = [&](short *v) {
        line++;
        if (fscanf(f, "%127[^,],%hd\n", name, v) != 2) bad();
        printf("PARAMETERS: %s = %hd\n", name, *v);
    };
    auto rd_double = [&](double *v, const char *fmt) {
        line++;
        if (fscanf(f, "%127[^,],%lf\n", name, v) != 2) bad();
        printf(fmt, name, *v);
    };
    auto rd_int = [&](int *v) {
        line++;
        if (fscanf(f, "%127[^,],%d\n", name, v) != 2) bad();
        printf("PARAMETERS: %s = %d\n", name, *v);
    };
    rd_short(&parameters.dimension);
    rd_short(&parameters.poly_degree);
    rd_short(&parameters.phs_degree);
    double mult = 0;
    line++;
    if (fscanf(f, "%127[^,],%lf\n", name, &mult) != 2 || !(mult > 0) || mult > 32767) bad();
    parameters.cloud_size_multiplier = (short)mult;
    g.multiplier = mult;
    printf("PARAMETERS: %s = %g\n", name, mult);
    rd_short(&parameters.test);
    rd_double(&parameters.courant_number, "PARAMETERS: %s = %lf\n");
    rd_double(&parameters.steady_state_tolerance, "PARAMETERS: %s = %e\n");
    rd_double(&parameters.poisson_solver_tolerance, "PARAMETERS: %s = %e\n");
    line++;
    if (fscanf(f, "%127[^,],%f\n", name, &parameters.omega) != 2) bad();
    printf("PARAMETERS: %s = %f\n", name, parameters.omega);
    rd_short(&parameters.num_vcycles);
    rd_short(&parameters.num_relax);
    rd_int(&parameters.num_time_steps);
    rd_int(&parameters.write_interval);
    int flag = 0;
    rd_int(&flag);
    parameters.neumann_flag_boundary = flag == 1;
    rd_double(&parameters.dt, "PARAMETERS: %s = %lf\n");
    rd_short(&parameters.iter_momentum);
    rd_short(&parameters.iter_timple);
    rd_double(&parameters.Re, "PARAMETERS: %s = %lf\n");
    rd_double(&parameters.facRe, "PARAMETERS: %s = %lf\n");
    rd_double(&parameters.facdt, "PARAMETERS: %s = %lf\n");
    fclose(f);
    parameters.rho = 1;
    parameters.mu = parameters.rho / parameters.Re;
    parameters.nu = parameters.mu / parameters.rho;
}

// read_grid_filenames (hf/mesh_functions.h:102-129): "name,levels" then one
// mesh file name per level (at most 49 characters, the size of
// PointStructure.mesh_filename; the reference's %s has no bound).
void read_grid_filenames(PointStructure **P, char *filename, short *num_levels)
{
    FILE *f = fopen(filename, "r");
    if (!f) { printf("Error: Unable to open the file\n"); exit(1); }
    char name[128];
    if (fscanf(f, "%127[^,],%hd\n", name, num_levels) != 2 || *num_levels < 1) {
        printf("Error: Unable to read the file\n");
        exit(1);
    }
    *P = (PointStructure *)calloc(*num_levels, sizeof(PointStructure));
    for (int l = 0; l < *num_levels; l++)
        if (fscanf(f, "%49s\n", (*P)[l].mesh_filename) != 1) {
            printf("Error: Unable to read the meshfile names\n");
            exit(1);
        }
    fclose(f);
}

void calculate_parameters(PointStructure *P)
{
    int dim = parameters.dimension;
    P->num_poly_terms = (short)poly_terms(dim, P->poly_degree, nullptr, nullptr, nullptr);
    double mult = g.multiplier > 0 ? g.multiplier : parameters.cloud_size_multiplier;
    P->num_cloud_points = (short)floor(mult * P->num_poly_terms + 1e-9);
    P->pow_x = (short *)malloc(sizeof(short) * P->num_poly_terms);
    P->pow_y = (short *)malloc(sizeof(short) * P->num_poly_terms);
    P->pow_z = (short *)malloc(sizeof(short) * P->num_poly_terms);
    poly_terms(dim, P->poly_degree, P->pow_x, P->pow_y, P->pow_z);
}

void read_complete_mesh_data(PointStructure *P, short num_levels)
{
    if (mgm_init(0)) die("mgm_init");
    mgm_config cfg;
    mgm_config_default(&cfg);
    cfg.dimension = parameters.dimension;
    cfg.poly_degree = parameters.poly_degree;
    cfg.coarse_poly_degree = 3;
    cfg.phs_degree = parameters.phs_degree;
    cfg.cloud_size_multiplier = g.multiplier > 0 ? g.multiplier : parameters.cloud_size_multiplier;
    cfg.omega = parameters.omega;
    cfg.num_relax = parameters.num_relax;
    cfg.transfer_rows = MGM_ROWS_REFERENCE;
    const char *ord = getenv("MGM_SWEEP_ORDER");
    cfg.sweep_order = (ord && !strcmp(ord, "colour")) ? MGM_ORDER_COLOUR : MGM_ORDER_NATURAL;
    if (g.pb) mgm_problem_destroy(g.pb);
    MUST(mgm_problem_create(&cfg, num_levels, &g.pb));
    for (int l = 0; l < num_levels; l++) {
        mgm_points *pts = nullptr;
        printf("\nReading mesh data for level %d\n", l + 1);
        MUST(mgm_points_read_gmsh(P[l].mesh_filename, parameters.dimension, &pts));
        MUST(mgm_problem_set_level(g.pb, l, pts));
        mgm_points_free(pts);
    }
    MUST(mgm_problem_setup(g.pb));
    g.base = P;
    g.L = num_levels;
    g.lapP_hash.assign(num_levels, 0);
    for (int l = 0; l < num_levels; l++) {
        PointStructure &s = P[l];
        mgm_level_info inf;
        MUST(mgm_level_info_get(g.pb, l, &inf));
        int N = inf.N, m = inf.m;
        AllocateMemoryPointStructure(&s, N);
        s.num_corners = 0;
        s.num_boundary_nodes = inf.nb;
        s.d_avg = inf.d_avg;
        s.poly_degree = (short)inf.poly_degree;
        calculate_parameters(&s);
        std::vector<uint8_t> bt(N);
        MUST(mgm_get_geometry(g.pb, l, s.x, s.y, s.z, s.x_normal, s.y_normal, s.z_normal, bt.data()));
        const Level &Lv = g.pb->lv[l];
        s.num_elem = Lv.in.num_elem;
        for (int i = 0; i < N; i++) {
            s.boundary_tag[i] = bt[i] != 0;
            s.corner_tag[i] = Lv.ctag[i] != 0;
            s.point_index[i] = i;
        }
        std::vector<int32_t> cl((size_t)N * m);
        MUST(mgm_get_cloud(g.pb, l, cl.data()));
        s.cloud_index = (int **)malloc(sizeof(int *) * N);
        for (int i = 0; i < N; i++) {
            s.cloud_index[i] = (int *)malloc(sizeof(int) * m);
            memcpy(s.cloud_index[i], &cl[(size_t)i * m], sizeof(int) * m);
        }
        s.prolongation_points = (int *)malloc(sizeof(int) * N);
        s.restriction_points = (int *)malloc(sizeof(int) * N);
        MUST(mgm_get_points(g.pb, l, MGM_PTS_PROLONGATION, s.prolongation_points));
        MUST(mgm_get_points(g.pb, l, MGM_PTS_RESTRICTION, s.restriction_points));
        s.prolongation_points2 = s.restriction_points2 = nullptr;
        s.Dx = s.Dy = s.Dz = s.lap = s.lap_Poison = nullptr;
        s.restr_mat = s.prol_mat = nullptr;
    }
    for (int l = 0; l < num_levels; l++) {
        PointStructure &s = P[l];
        if (l > 0) {
            std::vector<double> R((size_t)s.num_nodes * P[l - 1].num_cloud_points);
            MUST(mgm_get_matrix(g.pb, l, MGM_MAT_RESTR, R.data()));
            s.restr_mat = rows_from(R.data(), s.num_nodes, P[l - 1].num_cloud_points);
        }
        if (l < num_levels - 1) {
            std::vector<double> Pm((size_t)s.num_nodes * P[l + 1].num_cloud_points);
            MUST(mgm_get_matrix(g.pb, l, MGM_MAT_PROL, Pm.data()));
            s.prol_mat = rows_from(Pm.data(), s.num_nodes, P[l + 1].num_cloud_points);
        }
    }
}

// ---- kdtree_functions.h ----------------------------------------------------

void find_cloud_index(PointStructure *P)
{
    // Stencils are built with the rest of the hierarchy; refresh the host copy.
    int l = level_of(P);
    std::vector<int32_t> cl((size_t)P->num_nodes * P->num_cloud_points);
    MUST(mgm_get_cloud(g.pb, l, cl.data()));
    for (int i = 0; i < P->num_nodes; i++)
        memcpy(P->cloud_index[i], &cl[(size_t)i * P->num_cloud_points], sizeof(int) * P->num_cloud_points);
}

int *find_nearest_point(PointStructure *P1, PointStructure *P2, int)
{
    int l1 = level_of(P1), l2 = level_of(P2);
    int *out = (int *)malloc(sizeof(int) * P1->num_nodes);
    int which = (l2 == l1 + 1 || (l2 == l1 && l1 == g.L - 1)) ? MGM_PTS_PROLONGATION : MGM_PTS_RESTRICTION;
    MUST(mgm_get_points(g.pb, l1, which, out));
    return out;
}

// ---- rbf.h -------------------------------------------------------------------

void create_derivative_matrices(PointStructure *P)
{
    int l = level_of(P);
    int N = P->num_nodes, m = P->num_cloud_points;
    std::vector<double> v((size_t)N * m);
    auto fill = [&](int which, double ***dst) {
        MUST(mgm_get_matrix(g.pb, l, which, v.data()));
        *dst = rows_from(v.data(), N, m);
    };
    fill(MGM_MAT_DX, &P->Dx);
    fill(MGM_MAT_DY, &P->Dy);
    if (parameters.dimension == 3) fill(MGM_MAT_DZ, &P->Dz);
    fill(MGM_MAT_LAP, &P->lap);
    fill(MGM_MAT_LAPP, &P->lap_Poison);
    g.lapP_hash[l] = hash_rows(P->lap_Poison, *P);
}

// ---- general_functions.h ---------------------------------------------------

double calculate_dt(PointStructure *P)   // hf/general_functions.h:34-41
{
    double d = P->d_avg;
    double termd = 2 * parameters.nu / (d * d);
    double termc = 1 / d;
    double dt = 0.5 / (termd + termc);
    return dt * parameters.courant_number;
}

double l2_norm_gen(PointStructure *P, double *A, double *B, int n)   // hf/general_functions.h:181-192
{
    double r = 0.0;
    for (int i = 0; i < n; i++)
        if (P->boundary_tag[i] == 0) { double d = A[i] - B[i]; r += d * d; }
    return sqrt(r / n);
}

void test_derivatives(PointStructure *P, short num_levels, short dim)
{
    for (int l = 0; l < num_levels; l++) {
        printf("Testing level %d\n", l);
        int N = P[l].num_nodes;
        const double k2 = 2 * 3.14;   // the reference's 2*3.14*k with k = 1
        std::vector<double> f(N), F[4];
        for (auto &a : F) a.resize(N);
        for (int i = 0; i < N; i++) {
            double sx = sin(k2 * P[l].x[i]), sy = sin(k2 * P[l].y[i]), sz = sin(k2 * P[l].z[i]);
            double cx = cos(k2 * P[l].x[i]), cy = cos(k2 * P[l].y[i]), cz = cos(k2 * P[l].z[i]);
            if (dim > 2) {
                f[i] = sx * sy * sz;
                F[0][i] = k2 * cx * sy * sz; F[1][i] = k2 * sx * cy * sz; F[2][i] = k2 * sx * sy * cz;
                F[3][i] = -12 * 3.14 * 3.14 * sx * sy * sz;
            } else {
                f[i] = sx * sy;
                F[0][i] = k2 * cx * sy; F[1][i] = k2 * sx * cy;
                F[3][i] = -8 * 3.14 * 3.14 * sx * sy;
            }
        }
        put(l, MGM_F_T, f.data());
        double err[4] = {0, 0, 0, 0};
        std::vector<double> out(N);
        int mats[4] = {MGM_MAT_DX, MGM_MAT_DY, MGM_MAT_DZ, MGM_MAT_LAP};
        for (int q = 0; q < 4; q++) {
            if (q == 2 && dim < 3) continue;
            MUST(mgm_spmv(g.pb, l, mats[q], MGM_F_T, MGM_F_DPDN));
            get(l, MGM_F_DPDN, out.data());
            err[q] = l2_norm_gen(&P[l], F[q].data(), out.data(), N);
        }
        if (dim > 2)
            printf("L2 error in f, fx, fy, fz, lapf:\n%lf %lf %lf %lf %lf\n", 0.0, err[0], err[1], err[2], err[3]);
        else
            printf("L2 error in f, fx, fy, lapf:\n%lf %lf %lf %lf\n", 0.0, err[0], err[1], err[3]);
    }
}

// ---- openACC_functions.h ---------------------------------------------------
// The hierarchy is resident after read_complete_mesh_data; solver calls
// upload the host arrays they read, so these have nothing left to move.
void copyin_parameters_to_gpu(void) {}
void copyin_pointstructure_to_gpu(PointStructure *) {}
void copyin_field_to_gpu(FieldVariables *, PointStructure *) {}
void copyin_essentials_to_gpu(PointStructure *) {}
void copypout_pointstructure_from_gpu(PointStructure *) {}
void copypout_field_from_gpu(FieldVariables *, PointStructure *) {}

// ---- mat_lib.h SpMV ----------------------------------------------------------

static void spmv_rows(double **D, double *f, double *out, int **cloud, int n_rows, int n_cols)
{
    // Resident operator rows are applied from their device slabs; any other
    // row set is uploaded for the call.
    for (int l = 0; g.base && l < g.L; l++) {
        PointStructure &s = g.base[l];
        if (s.num_nodes != n_rows || s.num_cloud_points != n_cols || cloud != s.cloud_index) continue;
        int which = (D == s.Dx) ? MGM_MAT_DX : (D == s.Dy) ? MGM_MAT_DY : (D == s.Dz) ? MGM_MAT_DZ
                  : (D == s.lap) ? MGM_MAT_LAP : (D == s.lap_Poison) ? MGM_MAT_LAPP : -1;
        if (which < 0) continue;
        if (which == MGM_MAT_LAPP) upload_lapP_if_changed(l);
        put(l, MGM_F_T, f);
        MUST(mgm_spmv(g.pb, l, which, MGM_F_T, MGM_F_DPDN));
        get(l, MGM_F_DPDN, out);
        return;
    }
    std::vector<double> w((size_t)n_rows * n_cols);
    std::vector<int32_t> c((size_t)n_rows * n_cols);
    for (int i = 0; i < n_rows; i++) {
        memcpy(&w[(size_t)i * n_cols], D[i], sizeof(double) * n_cols);
        memcpy(&c[(size_t)i * n_cols], cloud[i], sizeof(int) * n_cols);
    }
    // the source vector spans the largest column index + 1
    int ncol = 0;
    for (int32_t v : c) ncol = std::max(ncol, (int)v + 1);
    MUST(mgm_spmv_rows(n_rows, n_cols, w.data(), c.data(), ncol, f, out));
}

void multiply_sparse_matrix_vector(double **D, double *f, double *out, int **cloud, int n_rows, int n_cols)
{
    spmv_rows(D, f, out, cloud, n_rows, n_cols);
}

void multiply_sparse_matrix_vector_gpu(double **D, double *f, double *out, int **cloud, int n_rows, int n_cols)
{
    spmv_rows(D, f, out, cloud, n_rows, n_cols);
}

// ---- variant solver primitives ---------------------------------------------

static int pf_of(int variant) { return variant == 1 ? MGM_F_PPRIME : MGM_F_P; }
static double *pf_host(FieldVariables *F, int variant) { return variant == 1 ? F->pprime : F->p; }

void mgm_dropin_relaxation(PointStructure *P, FieldVariables *F, int variant)
{
    int l = level_of(P);
    sync_params();
    upload_lapP_if_changed(l);
    put(l, pf_of(variant), pf_host(F, variant));
    put(l, MGM_F_SOURCE, F->source);
    if (variant == 2) put(l, MGM_F_RES, F->res);
    MUST(mgm_relax(g.pb, l, variant));
    get(l, pf_of(variant), pf_host(F, variant));
    if (variant != 0) get(l, MGM_F_RES, F->res);
    if (variant == 0) {   // the reference prints the (stale) residual norm (hf/heat_conduction.h:54)
        std::vector<double> zeros(P->num_nodes, 0.0);
        printf("Pressure Residual: %e\n", l2_norm_gen(P, F->res, zeros.data(), P->num_nodes));
    }
}

void mgm_dropin_calculate_residuals(PointStructure *P, FieldVariables *F, int variant)
{
    int l = level_of(P);
    upload_lapP_if_changed(l);
    put(l, pf_of(variant), pf_host(F, variant));
    put(l, MGM_F_SOURCE, F->source);
    put(l, MGM_F_RES, F->res);
    MUST(mgm_residual(g.pb, l, variant));
    get(l, MGM_F_RES, F->res);
}

void mgm_dropin_restrict_residuals(PointStructure *Pf, PointStructure *Pc, FieldVariables *Ff,
                                   FieldVariables *Fc, int variant)
{
    int lf = level_of(Pf), lc = level_of(Pc);
    if (lc != lf + 1) { fprintf(stderr, "libmgm: restrict_residuals needs adjacent levels\n"); exit(1); }
    put(lf, MGM_F_RES, Ff->res);
    put(lc, MGM_F_SOURCE, Fc->source);
    MUST(mgm_restrict(g.pb, lf, variant));
    get(lc, MGM_F_SOURCE, Fc->source);
}

void mgm_dropin_prolongate_corrections(PointStructure *Pf, PointStructure *Pc, FieldVariables *Ff,
                                       FieldVariables *Fc, int variant)
{
    int lf = level_of(Pf), lc = level_of(Pc);
    if (lc != lf + 1) { fprintf(stderr, "libmgm: prolongate_corrections needs adjacent levels\n"); exit(1); }
    put(lf, pf_of(variant), pf_host(Ff, variant));
    put(lc, pf_of(variant), pf_host(Fc, variant));
    MUST(mgm_prolong(g.pb, lc, variant));
    get(lf, pf_of(variant), pf_host(Ff, variant));
    get(lc, pf_of(variant), pf_host(Fc, variant));
}

static void ns_params_from_globals(FieldVariables *F)
{
    mgm_ns_params ns{};
    ns.dt = parameters.dt; ns.rho = parameters.rho; ns.mu = parameters.mu; ns.nu = parameters.nu;
    ns.facRe = parameters.facRe; ns.facdt = parameters.facdt;
    ns.field_rho = F[0].rho;
    ns.iter_momentum = parameters.iter_momentum; ns.iter_timple = parameters.iter_timple;
    ns.num_vcycles = parameters.num_vcycles;
    MUST(mgm_ns_set_params(g.pb, &ns));
}

static double *field_slot(FieldVariables &f, int which)
{
    switch (which) {
    case MGM_F_P: return f.p;           case MGM_F_SOURCE: return f.source; case MGM_F_RES: return f.res;
    case MGM_F_PPRIME: return f.pprime; case MGM_F_U: return f.u;           case MGM_F_V: return f.v;
    case MGM_F_W: return f.w;           case MGM_F_UNEW: return f.u_new;    case MGM_F_VNEW: return f.v_new;
    case MGM_F_WNEW: return f.w_new;    case MGM_F_POLD: return f.p_old;    case MGM_F_T: return f.T;
    case MGM_F_DPDN: return f.dpdn;     case MGM_F_DPDX: return f.dpdx;     case MGM_F_DPDY: return f.dpdy;
    case MGM_F_DPDZ: return f.dpdz;     case MGM_F_UOLD: return f.u_old;    case MGM_F_VOLD: return f.v_old;
    case MGM_F_WOLD: return f.w_old;
    }
    return nullptr;
}

static void sync_all_fields(FieldVariables *F, bool upload)
{
    for (int l = 0; l < g.L; l++) {
        if (upload) upload_lapP_if_changed(l);
        for (int k = 0; k < MGM_NUM_FIELDS; k++) {
            if (l > 0 && k != MGM_F_P && k != MGM_F_PPRIME && k != MGM_F_SOURCE && k != MGM_F_RES) continue;
            double *h = field_slot(F[l], k);
            if (upload) put(l, k, h); else get(l, k, h);
        }
    }
}

double mgm_dropin_time_implicit_solver(PointStructure *P, FieldVariables *F)
{
    if (P != g.base) level_of(P);
    sync_params();
    ns_params_from_globals(F);
    sync_all_fields(F, true);
    double err = 0;
    MUST(mgm_timple_step(g.pb, &err));
    sync_all_fields(F, false);
    return err;
}

double mgm_dropin_fractional_step_explicit(PointStructure *P, FieldVariables *F)
{
    if (P != g.base) level_of(P);
    sync_params();
    ns_params_from_globals(F);
    sync_all_fields(F, true);
    double err = 0;
    MUST(mgm_fs_step(g.pb, &err));
    sync_all_fields(F, false);
    return err;
}

void mgm_dropin_multigrid_Poisson_solver(PointStructure *P, FieldVariables *F, int variant)
{
    if (P != g.base) level_of(P);
    sync_params();
    for (int l = 0; l < g.L; l++) {
        upload_lapP_if_changed(l);
        put(l, pf_of(variant), pf_host(&F[l], variant));
        put(l, MGM_F_SOURCE, F[l].source);
        put(l, MGM_F_RES, F[l].res);
    }
    MUST(mgm_vcycle(g.pb, variant, parameters.num_vcycles));
    for (int l = 0; l < g.L; l++) {
        get(l, pf_of(variant), pf_host(&F[l], variant));
        get(l, MGM_F_SOURCE, F[l].source);
        get(l, MGM_F_RES, F[l].res);
    }
}

}  // extern "C"
