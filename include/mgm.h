/*
 * mgm.h -- C ABI of libmgm, the MI355X-native RBF-FD meshless multigrid path.
 *
 * Plain C types only (pointers, sizes, scalars); no torch or HIP types.
 * Every entry point returns 0 on success and a negative code on failure;
 * mgm_last_error() describes the last failure on the calling thread.
 * Host arrays are in the reference's node order (the order after its BFS
 * reordering, hf/mesh_functions.h:774-872) and row-major (a reference
 * `double** D` of N rows x m is D[i*m+j]).  Device data is owned by the
 * problem object and never exposed.
 *
 * Reference interfaces replaced (file:line in the reference tree):
 *   mgm_points_read_gmsh      read_PointStructure      header_files/mesh_functions.h:172
 *   mgm_problem_setup         read_complete_mesh_data  header_files/mesh_functions.h:914
 *                             + create_derivative_matrices header_files/rbf.h:401
 *   mgm_relax                 relaxation               header_files/heat_conduction.h:25,
 *                                                      header_files/timple_solver.h:235,
 *                                                      header_files/fractional_step_solver.h:168
 *   mgm_residual              calculate_residuals      header_files/heat_conduction.h:101 (+T :274, F :236)
 *   mgm_restrict              restrict_residuals       header_files/heat_conduction.h:58 (+T :292, F :198)
 *   mgm_prolong               prolongate_corrections   header_files/heat_conduction.h:79 (+T :308, F :214)
 *   mgm_vcycle                V-cycle loop             mg_heat_conduction.c:64-75,
 *                                                      header_files/timple_solver.h:191
 *   mgm_spmv                  multiply_sparse_matrix_vector_gpu header_files/mat_lib.h:81
 *   mgm_l2_norm_gen           l2_norm_gen              header_files/general_functions.h:181
 */
#ifndef MGM_H
#define MGM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enums ------------------------------------------------------------ */
enum { MGM_VAR_H = 0, MGM_VAR_T = 1, MGM_VAR_F = 2 };  /* heat / TIMPLE / fractional-step semantics */
enum { MGM_ROWS_REFERENCE = 0, MGM_ROWS_BY_TAG = 1 };  /* transfer rows for i>=nb, or for !boundary */
enum { MGM_ORDER_COLOUR = 0, MGM_ORDER_NATURAL = 1 };  /* SOR sweep order */
enum { MGM_SMOOTH_SOR = 0, MGM_SMOOTH_JACOBI = 1 };    /* smoother of a level (mgm_set_smoother_kinds) */
enum {
    MGM_MAT_DX = 0, MGM_MAT_DY = 1, MGM_MAT_DZ = 2, MGM_MAT_LAP = 3, MGM_MAT_LAPP = 4,
    MGM_MAT_RESTR = 5, MGM_MAT_PROL = 6
};
enum {
    MGM_F_P = 0, MGM_F_SOURCE = 1, MGM_F_RES = 2, MGM_F_PPRIME = 3, MGM_F_U = 4, MGM_F_V = 5,
    MGM_F_W = 6, MGM_F_UNEW = 7, MGM_F_VNEW = 8, MGM_F_WNEW = 9, MGM_F_POLD = 10,
    MGM_F_T = 11, MGM_F_DPDN = 12, MGM_F_DPDX = 13, MGM_F_DPDY = 14, MGM_F_DPDZ = 15,
    MGM_F_UOLD = 16, MGM_F_VOLD = 17, MGM_F_WOLD = 18, MGM_NUM_FIELDS = 19
};
enum { MGM_PTS_RESTRICTION = 0, MGM_PTS_PROLONGATION = 1 };

/* ---- configuration (mirror of struct parameters, header_files/structures.h:15-41) */
typedef struct {
    int dimension;                 /* 2 or 3 */
    int poly_degree;               /* finest level */
    int coarse_poly_degree;        /* other levels; the reference hard-codes 3 (mesh_functions.h:917) */
    int phs_degree;
    double cloud_size_multiplier;  /* m = floor(multiplier * n) */
    float omega;                   /* SOR factor; float as in the reference */
    int num_relax;
    int transfer_rows;             /* MGM_ROWS_* */
    int sweep_order;               /* MGM_ORDER_* */
    /* No reference counterpart: wider stencils on boundary nodes.  0 (the
     * default) gives every node m = floor(cloud_size_multiplier * n) points,
     * as the reference does (kdtree_functions.h:139-149).  A larger value
     * gives each boundary node a second stencil of m_b = floor(value * n)
     * points (itself, then its m_b - 1 nearest interior nodes, the same rule
     * as the reference's boundary stencils) and wide derivative operators on
     * it (mgm_get_boundary_rows, mgm_spmv_wide); every other operator keeps
     * the m-point stencils.  m_b + n <= 128. */
    double boundary_cloud_size_multiplier;
} mgm_config;

typedef struct {
    int N, nb, m, n, poly_degree, num_colours;
    double d_avg;                  /* after calculate_avg_dx (mesh_functions.h:895) */
    double d_avg_bbox;             /* reader's bbox estimate (mesh_functions.h:274-281) */
    int knn_ties;                  /* queries whose stencil needed the kd-order tie-break */
    /* device layout (no reference counterpart) */
    int device_rows;               /* rows of the slabs, padding included */
    int slice_height;              /* rows per ELL slice of the operator slabs */
    int blocked;                   /* rows ordered in Morton blocks (colour steps take slice lists) */
    int gather_sorted;             /* lap / lap_Poison also stored gather-sorted */
    int small_relax;               /* relaxes in one workgroup (at the current MGM_SMALL_ROWS) */
    int window_blocks;             /* full passes read x from LDS windows of this many blocks (0: none) */
    int window_max;                /* largest window (values); places are stored as 16-bit */
    int prolong_window_max;        /* the prolongation into this level reads LDS windows of at most this
                                      many coarse values (0: it gathers from memory) */
} mgm_level_info;

typedef struct mgm_points mgm_points;
typedef struct mgm_problem mgm_problem;

/* ---- status / device ------------------------------------------------------ */
const char *mgm_last_error(void);
int mgm_device_count(void);        /* 0 when no GPU is visible */
int mgm_init(int device);          /* select device, create the library stream */

/* ---- host-only helpers (no GPU needed) ------------------------------------ */
void mgm_config_default(mgm_config *cfg);
int mgm_poly_terms(int dim, int deg, short *px, short *py, short *pz);
int mgm_points_read_gmsh(const char *path, int dim, mgm_points **out);
/* Jittered Cartesian cloud on [0,1]^dim, n per axis (seeded SplitMix64);
 * optionally also written to gmsh_path as Gmsh v2. */
int mgm_points_generate(int dim, int n, uint64_t seed, double jitter, int with_volume_elements,
                        const char *gmsh_path, mgm_points **out);
/* Points from arrays laid out as the reader leaves them (after corner removal). */
int mgm_points_from_arrays(int N, int nb, int num_elem, double d_avg, const double *x,
                           const double *y, const double *z, const double *nx, const double *ny,
                           const double *nz, const uint8_t *btag, const uint8_t *ctag,
                           mgm_points **out);
int mgm_points_info(const mgm_points *p, int *N, int *nb, int *num_elem, double *d_avg);
int mgm_points_copy(const mgm_points *p, double *x, double *y, double *z, double *nx,
                    double *ny, double *nz, uint8_t *btag, uint8_t *ctag);
void mgm_points_free(mgm_points *p);
uint64_t mgm_splitmix64(uint64_t *state);

/* ---- problem: setup -------------------------------------------------------- */
int mgm_problem_create(const mgm_config *cfg, int num_levels, mgm_problem **out);
void mgm_problem_destroy(mgm_problem *p);
int mgm_problem_set_level(mgm_problem *p, int lev, const mgm_points *pts);
/* kNN stencils, BFS reordering, normals, spacing, transfer points, colouring,
 * weights (Dx, Dy, Dz, Laplacian, lap_Poison), transfer rows, ELL slabs. */
int mgm_problem_setup(mgm_problem *p);
int mgm_level_info_get(const mgm_problem *p, int lev, mgm_level_info *out);
/* seconds spent in setup phases: [0] kNN [1] reorder+colour [2] transfer points
 * [3] weights+transfer rows [4] slab assembly [5] total */
int mgm_setup_timings(const mgm_problem *p, double *t6);
/* Rebuild the weight rows of one level (batched local solves) and return the
 * kernel time in seconds and the number of systems solved (the wide boundary
 * systems included). */
int mgm_rebuild_weights(mgm_problem *p, int lev, double *seconds, long *systems);
/* No reference counterpart: the variant of the batched weight kernel that
 * ran the last build of systems of size mpn (= m + n; the autotuner's choice
 * unless the environment forced one): the
 * FP64 Gauss-Jordan kernel with *tile x 16 rows and columns of register tiles
 * and *waves_per_simd its register budget (1: unconstrained), or, when
 * *waves_per_simd is 0, the matrix-core kernel (FP32 LU on MFMA with FP64
 * iterative refinement, *tile tiles of 16 x 16).  Returns -1 if no system of
 * that size was built yet. */
int mgm_weight_kernel_choice(int mpn, int *tile, int *waves_per_simd);

/* ---- problem: data exchange (reference node order) ----------------------- */
int mgm_get_geometry(const mgm_problem *p, int lev, double *x, double *y, double *z,
                     double *nx, double *ny, double *nz, uint8_t *btag);
int mgm_get_cloud(const mgm_problem *p, int lev, int32_t *cloud);
int mgm_get_points(const mgm_problem *p, int lev, int which, int32_t *out);
int mgm_get_colours(const mgm_problem *p, int lev, int32_t *colour);
/* width: m of this level (DX..LAPP), m of level lev-1 (RESTR), m of level lev+1 (PROL) */
int mgm_get_matrix(const mgm_problem *p, int lev, int which, double *out);
int mgm_set_matrix(mgm_problem *p, int lev, int which, const double *in);
/* Identity rows (Dirichlet) in lap_Poison for every boundary node of every level. */
int mgm_dirichlet_rows(mgm_problem *p);
int mgm_get_field(const mgm_problem *p, int lev, int which, double *out);
int mgm_set_field(mgm_problem *p, int lev, int which, const double *in);
int mgm_zero_fields(mgm_problem *p);

/* ---- output (reference node order, the drivers' formats) ------------------ */
/* Solution file of level lev: one line per node, "x, y[, z], f1, f2, ...",
 * every value printed with "%lf" and ", " between them, as the drivers write
 * Solution.csv: with_z = 0 and fields {P} is mg_heat_conduction.c:89-93,
 * with_z = 1 and fields {U, V, W, POLD} is mg_NS_solver.c:84-90.  fields are
 * MGM_F_* (at most 16). */
int mgm_write_solution(const mgm_problem *p, int lev, const char *path, int with_z, int nfields,
                       const int *fields);
/* Convergence file: line i is "i, value" with the value printed "%e"
 * (scientific != 0, mg_NS_solver.c:77-78) or "%lf" (mg_heat_conduction.c:77). */
int mgm_write_history(const char *path, int n, const double *values, int scientific);

/* Smoother parameters used by subsequent relaxations (parameters.omega and
 * parameters.num_relax in the reference, header_files/structures.h:26,31);
 * omega applies to every level (per-level factors are cleared). */
int mgm_set_smoother(mgm_problem *p, float omega, int num_relax);

/* ---- problem: operators (reference variant semantics) --------------------- */
int mgm_relax(mgm_problem *p, int lev, int variant);
int mgm_residual(mgm_problem *p, int lev, int variant);
int mgm_restrict(mgm_problem *p, int lev, int variant);      /* lev -> lev+1 */
int mgm_prolong(mgm_problem *p, int lev, int variant);       /* lev -> lev-1 */
int mgm_vcycle(mgm_problem *p, int variant, int ncycles);
/* out = M * in over fixed-width rows (M: MGM_MAT_DX..MGM_MAT_LAPP). */
int mgm_spmv(mgm_problem *p, int lev, int which_matrix, int in_field, int out_field);
/* No reference counterpart (wide boundary stencils, boundary_cloud_size_multiplier):
 * the number of boundary nodes with a wide stencil and its width m_b (0, 0 when
 * the level has none). */
int mgm_boundary_info(const mgm_problem *p, int lev, int *nbw, int *mb);
/* No reference counterpart: the wide boundary rows of Dx, Dy, Dz or lap
 * (MGM_MAT_DX .. MGM_MAT_LAP): nodes[nbw] (reference order), cloud[nbw*m_b]
 * (node 0 of each row is the node itself) and w[nbw*m_b]; null outputs are
 * skipped.  Solved in the same batched weight build as the m-point systems. */
int mgm_get_boundary_rows(const mgm_problem *p, int lev, int which, int *nodes, int *cloud, double *w);
/* No reference counterpart: mgm_spmv of Dx, Dy, Dz or lap with the boundary
 * rows taken from the wide stencils (ordered sums, as
 * multiply_sparse_matrix_vector, header_files/mat_lib.h:70-89, over the wider
 * rows). */
int mgm_spmv_wide(mgm_problem *p, int lev, int which_matrix, int in_field, int out_field);
int mgm_l2_norm_gen(const mgm_problem *p, int lev, int field_a, int field_b, double *out);
/* out = D f over caller-supplied fixed-width rows (uploaded for the call):
 * D and cloud are n_rows x n_cols row-major, f has ncol entries.
 * (multiply_sparse_matrix_vector_gpu for row sets that are not resident.) */
int mgm_spmv_rows(int n_rows, int n_cols, const double *D, const int32_t *cloud, int ncol,
                  const double *f, double *out);
/* No counterpart in the reference: mgm_spmv_rows for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and result, bit for bit;
 * the device staging of D, cloud, f and out lives in a 64 MiB array per GPU
 * that the library holds until the process ends, so calls whose
 * 12 * n_rows * n_cols + 8 * (ncol + n_rows) bytes (plus alignment) exceed it
 * are rejected.  Calls on one GPU share that array: calls that overlap on one
 * GPU (other host threads, other streams) race.  mgm_init must have selected
 * the device. */
int mgm_spmv_rows_static(int n_rows, int n_cols, const double *D, const int32_t *cloud, int ncol,
                         const double *f, double *out);

/* ---- time steps that call the V-cycle (level-0 fields, resident) ----------- */
typedef struct {
    double dt, rho, mu, nu, facRe, facdt;
    double field_rho;             /* FieldVariables.rho read by timple_solver.h:389 */
    int iter_momentum, iter_timple, num_vcycles;
} mgm_ns_params;
int mgm_ns_set_params(mgm_problem *p, const mgm_ns_params *ns);
/* fractional_step_explicit (header_files/fractional_step_solver.h:47): predictor,
 * divergence source, boundary dp/dn, F-variant V-cycles on P, projection. */
int mgm_fs_step(mgm_problem *p, double *steady_error);
/* time_implicit_solver (header_files/timple_solver.h:40): implicit momentum
 * Gauss-Seidel (sweep order as configured), divergence source, T-variant
 * V-cycles on PPRIME, velocity/pressure update, Neumann boundary pressure. */
int mgm_timple_step(mgm_problem *p, double *steady_error);
/* The time loop of mg_NS_solver.c:68-83: up to max_steps steps of scheme
 * MGM_VAR_F (fractional_step_explicit) or MGM_VAR_T (time_implicit_solver),
 * stopping after the first step whose steady-state error is below steady_tol.
 * One whole step, V-cycles included, is captured into a hipGraph once and
 * replayed (the environment variable MGM_GRAPH=0 launches on the stream
 * instead); errors[k] receives step k's error, *seconds the loop's wall time.
 * precision 0: FP64 V-cycles, bitwise those of mgm_fs_step / mgm_timple_step.
 * precision 1 (no reference counterpart; multicolour order): each pressure
 * V-cycle becomes an FP64 defect correction by one FP32 V-cycle of the same
 * variant; divergence, projection and the outer residual stay FP64. */
int mgm_ns_march(mgm_problem *p, int scheme, int precision, int max_steps, double steady_tol,
                 double *errors, int *steps_done, double *seconds);

/* ---- learned smoother (no reference counterpart) -------------------------------
 * A per-node network replaces the first SOR sweep of every level's relaxation
 * (H variant, multicolour order, FP64 solve).  Input of node i: the m stencil
 * offsets (dim values each, divided by the level's mean spacing) and the m
 * stencil residuals divided by their RMS sigma_i; layers in -> 64 -> 64 -> 1
 * with ReLU; the correction is output * sigma_i / a_ii on interior nodes.
 * Weights: FP32 masters, laid out W1[64][in], b1[64], W2[64][64], b2[64],
 * w3[64], b3; inference uses BF16-rounded W1 (residual columns), W2, w3 with
 * FP32 accumulation.  mgm_solve runs its first cycle plain and falls back to
 * plain SOR for the rest of the solve as soon as a learned cycle reduces the
 * residual less than that plain cycle did. */
typedef struct {
    int snapshots;              /* random sources; each gives 3 (error, residual) states */
    int samples_per_snapshot;   /* interior nodes sampled per state */
    int epochs, batch;
    float lr;                   /* Adam step size */
    uint64_t seed;
} mgm_train_config;
int mgm_mlp_inputs(const mgm_problem *p);             /* m*dim + m of level 0 */
size_t mgm_mlp_weight_count(int inputs);
int mgm_mlp_set_weights(mgm_problem *p, const float *w);
int mgm_mlp_get_weights(const mgm_problem *p, float *w);
/* Network on caller rows X (n x inputs, row-major): y = f(X); and the loss
 * mean((f(X) - t)^2) with its gradient (layout of the weights), FP32. */
int mgm_mlp_forward(mgm_problem *p, int n, const float *X, float *y);
int mgm_mlp_gradient(mgm_problem *p, int n, const float *X, const float *t, double *loss, float *grad);
/* Training rows of level-0 nodes (reference-order indices) from caller
 * fields in reference order: X[b] (inputs values) as the network reads node
 * nodes[b] with residual field res, and the target t[b] = err[i] a_ii / sigma_i
 * (0 where sigma_i = 0), unclipped.  FP32 outputs. */
int mgm_mlp_features(mgm_problem *p, int n, const int32_t *nodes, const double *res, const double *err, float *X,
                     float *t);
/* Adam (beta 0.9 / 0.999, eps 1e-8, bias-corrected step size) on the current
 * weights over caller rows: each epoch shuffles the rows (Fisher-Yates from
 * seed, splitmix64) and takes floor(n / batch) steps of batch rows; loss over
 * all rows before and after. */
int mgm_mlp_fit(mgm_problem *p, int n, const float *X, const float *t, int epochs, int batch, float lr,
                uint64_t seed, double *loss_first, double *loss_last);
/* Train on level 0 from (error, residual) pairs of the plain V-cycle (the
 * fields of the problem are overwritten): the rows of mgm_mlp_features with
 * targets clipped to |t| <= 8, He-normal initial weights, then mgm_mlp_fit
 * with the config's epochs, batch, lr and the seed's stream continued; loss
 * over all samples before and after training. */
int mgm_train_smoother(mgm_problem *p, const mgm_train_config *cfg, double *loss_first, double *loss_last);
int mgm_use_learned_smoother(mgm_problem *p, int enable);
int mgm_learned_fell_back(const mgm_problem *p);      /* 1 if the last solve fell back to SOR */
/* The levels whose first sweep the network replaces (use[l] != 0), n = the
 * number of levels; n = 0: every level whose stencils fit the network (the
 * default).  The others keep their plain smoother in every sweep. */
int mgm_set_learned_levels(mgm_problem *p, const int *use, int n);

/* ---- resident solve -------------------------------------------------------- */
/* V-cycles on level-0 (field P, SOURCE) until l2(r)/l2(f) <= tol (l2_norm_gen
 * norms, r = source - lap_Poison*p on interior nodes) or max_cycles.  history
 * (may be NULL) receives max_cycles+1 ratios (index 0 = before the first cycle).
 * precision: 0 FP64 V-cycles (any variant, either sweep order);
 *            1 defect correction: FP64 residual and update, each correction one
 *              FP32 V-cycle (FP32 weights, vectors and sums) from zero;
 *            2 as 1 with BF16-stored weights on levels >= 1.
 *            1 and 2 support the H variant in the multicolour order. */
int mgm_solve(mgm_problem *p, int variant, int precision, double tol, int max_cycles,
              int *cycles_done, double *history, double *seconds);
/* SOR factor autotuner (no reference counterpart: the reference reads omega
 * from flow_parameters.csv, hf/mesh_functions.h:37-100).  Runs mgm_solve from
 * the current level-0 iterate and source for omega = lo, lo + step, ..., hi,
 * then at the midpoints on either side of the best; keeps the omega that
 * reaches tol in the fewest cycles (ties: the smaller final residual) and sets
 * it as the smoother's factor.  A candidate whose solve diverges or needs more
 * cycles than the best so far is stopped early.  The level-0 iterate and
 * source are restored afterwards and every other field is left zero.  Fails
 * (-12, omega unchanged) when no candidate reaches tol within max_cycles. */
int mgm_tune_omega(mgm_problem *p, int variant, int precision, double tol, int max_cycles, float lo, float hi,
                   float step, float *omega, int *cycles);
/* A SOR factor per level (no reference counterpart: the reference uses one
 * factor on every level).  n = the number of levels, or 0 to return to the
 * single factor of mgm_set_smoother (which also clears per-level factors).
 * mgm_get_smoother_levels writes the factor each level uses. */
int mgm_set_smoother_levels(mgm_problem *p, const float *omega, int n);
int mgm_get_smoother_levels(const mgm_problem *p, float *omega);
/* The smoother of each level (no reference counterpart: the reference relaxes
 * by SOR, hf/heat_conduction.h:25-56): MGM_SMOOTH_SOR (the default, in the
 * configured sweep order) or MGM_SMOOTH_JACOBI -- each sweep updates every row
 * from the previous iterate with the SOR formula, so a sweep is one full pass
 * instead of one launch per colour class.  Jacobi applies to the H variant
 * (mgm_vcycle, mgm_solve, the relaxation of the H variant); the T and F
 * variants keep SOR.  n = the number of levels, or 0 for SOR everywhere. */
int mgm_set_smoother_kinds(mgm_problem *p, const int *kind, int n);
int mgm_get_smoother_kinds(const mgm_problem *p, int *kind);
/* Per-level autotuner: coordinate descent from the current factors -- each
 * level's factor over lo, lo + step, ..., hi with the others held, up to
 * `sweeps` passes over the levels (stopping after a pass that changes
 * nothing) -- keeping the candidate that reaches tol in the fewest cycles
 * (ties: the smaller final residual), with the same early stop and state
 * handling as mgm_tune_omega.  omega receives the per-level factors (number
 * of levels), which stay set. */
int mgm_tune_omega_levels(mgm_problem *p, int variant, int precision, double tol, int max_cycles, float lo,
                          float hi, float step, int sweeps, float *omega, int *cycles);
/* As mgm_tune_omega_levels, with each level's smoother kind searched too:
 * for each level, every allowed kind (bit 1 << MGM_SMOOTH_*) over the factor
 * grid.  With both kinds allowed the candidates differ in their cost per
 * cycle, so they are ranked by measured solve time (a candidate must be 0.5%
 * faster to win), not by cycles.  omega and kind receive the per-level
 * choice, which stays set. */
int mgm_tune_smoother(mgm_problem *p, int variant, int precision, double tol, int max_cycles, float lo, float hi,
                      float step, int sweeps, int kinds_allowed, float *omega, int *kind, int *cycles);
/* Seconds per copy of `bytes` (rounded down to 16) between two device buffers
 * (a streaming-copy kernel): the measured HBM reference of the roofline. */
int mgm_stream_copy(size_t bytes, int reps, double *seconds);
/* Time `reps` applications of lap_Poison on level lev; returns seconds per apply. */
int mgm_time_spmv(mgm_problem *p, int lev, int reps, double *seconds);

#ifdef __cplusplus
}
#endif
#endif
