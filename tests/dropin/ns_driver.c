/*
 * A lid-driven-cavity driver written against the drop-in headers
 * (include/header_files), following the call sequence of the reference's
 * mg_NS_solver.c:15-60: read_flow_parameters, read_grid_filenames,
 * read_complete_mesh_data, create_derivative_matrices per level,
 * AllocateMemoryFieldVariables, the case's initial and boundary conditions,
 * dt from calculate_dt, the copyin_* calls, then one solver call per time
 * step.  Built with -DFRACTIONAL_STEP it steps with fractional_step_explicit
 * (fractional_step_solver.h), otherwise with time_implicit_solver
 * (timple_solver.h).
 *
 * The conditions are this file's own: all fields zero, then u = 1 (and, for
 * the TIMPLE step, which reads its velocity from u_new, u_new = 1) on the
 * level-0 nodes with |y - 1| < 1e-9; field rho = parameters.rho.
 *
 * Run from a directory holding flow_parameters.csv, grid_filenames.csv and
 * the meshes it names.  Writes steps.bin (dt, then each step's steady-state
 * error) and out.bin (level-0 x, y, u, v, p, u_new, v_new, p_old, pprime; N
 * doubles each).
 */
#ifdef FRACTIONAL_STEP
#include "header_files/fractional_step_solver.h"
#define TIME_STEP fractional_step_explicit
#else
#include "header_files/timple_solver.h"
#define TIME_STEP time_implicit_solver
#endif
#include "header_files/openACC_functions.h"

static void cavity(PointStructure *P, FieldVariables *F)
{
    F[0].rho = parameters.rho;
    for (int i = 0; i < P[0].num_nodes; i++)
        if (fabs(P[0].y[i] - 1.0) < 1e-9) {
            F[0].u[i] = 1.0;
#ifndef FRACTIONAL_STEP
            F[0].u_new[i] = 1.0;
#endif
        }
}

int main(void)
{
    PointStructure *P;
    FieldVariables *F;

    read_flow_parameters("flow_parameters.csv");
    read_grid_filenames(&P, "grid_filenames.csv", &parameters.num_levels);
    read_complete_mesh_data(P, parameters.num_levels);
    for (int l = 0; l < parameters.num_levels; l++) create_derivative_matrices(&P[l]);
    AllocateMemoryFieldVariables(&F, P, parameters.num_levels);
    cavity(P, F);
    parameters.dt = calculate_dt(&P[0]);
    copyin_essentials_to_gpu(P);
    copyin_field_to_gpu(F, P);
    copyin_parameters_to_gpu();

    FILE *s = fopen("steps.bin", "wb");
    if (!s) return 1;
    int bad = fwrite(&parameters.dt, sizeof(double), 1, s) != 1;
    for (int it = 0; it < parameters.num_time_steps; it++) {
        double err = TIME_STEP(P, F);
        printf("Time step: %d, Steady state error: %e\n", it, err);
        bad |= fwrite(&err, sizeof(double), 1, s) != 1;
    }
    bad |= fclose(s) != 0;

    FILE *f = fopen("out.bin", "wb");
    if (!f) return 1;
    const double *a[] = {P[0].x, P[0].y, F[0].u, F[0].v, F[0].p, F[0].u_new, F[0].v_new, F[0].p_old, F[0].pprime};
    for (int k = 0; k < 9; k++) bad |= fwrite(a[k], sizeof(double), P[0].num_nodes, f) != (size_t)P[0].num_nodes;
    bad |= fclose(f) != 0;

    int L = parameters.num_levels;
    free_PointStructure(P, L);
    free_field(F, L);
    return bad;
}
