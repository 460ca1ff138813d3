/*
 * mgm_oracle.h -- CPU FP64 restatement of the reference RBF-FD multigrid hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in multigrid_meshless_amd/.  Only tests/, the smoke() entry
 * point and bench.py's cpu_baseline leg may load it.  The product path never
 * links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference tree, header_files/ abbreviated hf/).  Arrays are flat and
 * row-major: a reference `double** D` with N rows of m entries is `D[i*m+j]`.
 */
#ifndef MGM_ORACLE_H
#define MGM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Gmsh v2 reader (hf/mesh_functions.h:172-476) -------------------- */
typedef struct {
    int N;            /* nodes after corner removal */
    int nb;           /* boundary nodes (count(boundary) - count(corner)) */
    int num_elem;     /* element count, all types */
    int num_corners_removed;
    double d_avg;     /* bbox estimate (:274-281) */
    double *x, *y, *z, *nx, *ny, *nz;    /* normals: accumulated, unnormalised */
    uint8_t *btag, *ctag;                 /* boundary / corner tags (shifted) */
} orc_mesh;

int  orc_read_gmsh(const char *path, int dim, orc_mesh *out);
void orc_free_mesh(orc_mesh *m);

/* ---- monomial table (hf/mesh_functions.h:131-170) --------------------- */
int orc_poly_terms(int dim, int deg, short *px, short *py, short *pz);

/* ---- kNN stencils (hf/kdtree_functions.h:61-175, hf/kdtree.c) --------- */
int orc_find_cloud_index(int N, int nb, const double *x, const double *y, const double *z,
                         const uint8_t *btag, const uint8_t *ctag, double d_avg, int m, int *cloud);
int orc_find_nearest_point(int N1, const double *x1, const double *y1, const double *z1,
                           const uint8_t *ctag1,
                           int N2, const double *x2, const double *y2, const double *z2,
                           const uint8_t *ctag2, double d_avg2, int *out);

/* ---- RCM / normals / d_avg (hf/mesh_functions.h:478-527, 774-910) ------ */
int    orc_rcm(int N, int m, double *x, double *y, double *z, double *nx, double *ny, double *nz,
               uint8_t *btag, int *cloud, int *queue1_out);
void   orc_correct_normals(int N, int m, const double *x, const double *y, const double *z,
                           double *nx, double *ny, double *nz, const uint8_t *btag, const int *cloud);
double orc_avg_dx(int N, int m, int dim, const double *x, const double *y, const double *z,
                  const uint8_t *btag, const int *cloud);

/* ---- RBF-FD weights (hf/rbf.h:32-421, hf/mat_lib.h:214-228, 423-491) -- */
/* method 0: unpivoted Gauss-Jordan inverse exactly as the reference;
 * method 1: LU with partial pivoting (textbook; used as a cross-check).  */
int orc_weights(int N, int m, int n, int dim, int phs, const short *px, const short *py,
                const short *pz, const double *x, const double *y, const double *z,
                const uint8_t *ctag, const int *cloud, int method,
                double *Dx, double *Dy, double *Dz, double *lap);
void orc_lap_poisson(int N, int m, int dim, const double *Dx, const double *Dy, const double *Dz,
                     const double *lap, const double *nx, const double *ny, const double *nz,
                     const uint8_t *btag, const uint8_t *ctag, double *lapP);

/* ---- transfer rows (hf/mesh_functions.h:533-670) ---------------------- */
/* row_mode 0: rows for indices >= nb (reference); 1: rows for every node
 * whose boundary tag is false (documented deviation used by the solver).  */
int orc_restriction(int Nf, int mf, int nf, int dim, int phs, const short *px, const short *py,
                    const short *pz, const double *xf, const double *yf, const double *zf,
                    const int *cloud_f, int Nc, int nbc, const double *xc, const double *yc,
                    const double *zc, const uint8_t *btag_c, const int *restr_pts, int row_mode,
                    int method, double *R);
int orc_prolongation(int Nf, int nbf, const double *xf, const double *yf, const double *zf,
                     const uint8_t *btag_f, const int *prol_pts, int Nc, int mc, int nc, int dim,
                     int phs, const short *px, const short *py, const short *pz,
                     const double *xc, const double *yc, const double *zc, const int *cloud_c,
                     int row_mode, int method, double *P);

/* ---- V-cycle primitives (hf/heat_conduction.h, hf/timple_solver.h,
 *      hf/fractional_step_solver.h) ------------------------------------ */
typedef struct {
    int N, m;
    const int *cloud;        /* N*m */
    const double *Lp;        /* lap_Poison N*m */
    const double *lap;       /* lap N*m (variant F residual) */
    const uint8_t *btag;
    const int *order;        /* sweep order, length N; NULL = natural */
    double *p, *s, *res;     /* p is field.p (H, F) or field.pprime (T) */
    const int *rp;           /* restriction_points (this level as coarse) */
    const double *R;         /* restr_mat N*m_fine */
    const int *pp;           /* prolongation_points (this level as fine) */
    const double *P;         /* prol_mat N*m_coarse */
    const int *cls;          /* optional: colour class c is order[cls[c] .. cls[c+1]-1] */
    int ncls;
} orc_level;

/* Threads the OpenMP loops use (results do not depend on it). */
void orc_set_num_threads(int n);

enum { ORC_VAR_H = 0, ORC_VAR_T = 1, ORC_VAR_F = 2 };

void   orc_relax(int variant, orc_level *L, float omega, int num_relax);
void   orc_residual(int variant, orc_level *L);
void   orc_restrict(int variant, orc_level *f, orc_level *c);
void   orc_prolong(int variant, orc_level *f, orc_level *c);
void   orc_vcycle(int variant, int nlev, orc_level *levs, float omega, int num_relax);
void   orc_vcycle_levels(int variant, int nlev, orc_level *levs, const float *omega, int num_relax);
void   orc_vcycle_smooth(int nlev, orc_level *levs, const float *omega, const int *kind, int num_relax);
double orc_l2_norm_gen(int N, const uint8_t *btag, const double *a, const double *b);
void   orc_spmv(int N, int m, const double *W, const int *cloud, const double *f, double *out);

/* ---- greedy multicolouring of the symmetrised stencil graph ----------- */
int orc_greedy_colour(int N, int m, const int *cloud, int *colour);

/* ---- SplitMix64 (project RNG for synthetic clouds) -------------------- */
uint64_t orc_splitmix64(uint64_t *state);

/* ---- OpenMP thread count actually used by the parallel loops ---------- */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
