/*
 * A heat-conduction driver written against the drop-in headers
 * (include/header_files), following the call sequence of the reference's
 * mg_heat_conduction.c:18-86 -- read_flow_parameters, read_grid_filenames,
 * read_complete_mesh_data, create_derivative_matrices per level,
 * AllocateMemoryFieldVariables, then V-cycles built from relaxation /
 * calculate_residuals / restrict_residuals / prolongate_corrections, with a
 * plug-in that rewrites boundary rows of the host lap_Poison arrays and the
 * boundary source before every cycle (the pattern of the reference's
 * init/init_heat_conduction.c:50-65).
 *
 * The plug-in here is this file's own: Dirichlet identity rows on the wall
 * x == 0 with value 1, a smooth Neumann-type value on the other boundary
 * nodes, and a smooth interior source.
 *
 * Run from a directory holding flow_parameters.csv, grid_filenames.csv and
 * the meshes it names.  Writes out.bin: level-0 x, y, source and p, N
 * doubles each, for comparison with the resident API.
 */
#include "header_files/heat_conduction.h"
#include "header_files/openACC_functions.h"

static const double PI = 3.14159265358979323846;

static void initial_fields(PointStructure *P, FieldVariables *F, int num_levels)
{
    for (int l = 0; l < num_levels; l++)
        for (int i = 0; i < P[l].num_nodes; i++) {
            F[l].res[i] = 0.0;
            F[l].p[i] = 0.0;
            F[l].T[i] = 0.0;
            F[l].source[i] = l == 0 ? -2 * PI * PI * sin(PI * P[0].x[i]) * sin(PI * P[0].y[i]) : 0.0;
        }
}

static void boundary_rhs(PointStructure *P, FieldVariables *F)
{
    for (int i = 0; i < P->num_nodes; i++) {
        if (!P->boundary_tag[i]) continue;
        if (P->x[i] == 0) {
            F[0].source[i] = 1.0;
            for (int j = 0; j < P->num_cloud_points; j++) P->lap_Poison[i][j] = 0.0;
            P->lap_Poison[i][0] = 1.0;
        } else {
            F[0].source[i] = PI * (cos(PI * P->x[i]) * sin(PI * P->y[i]) * P->x_normal[i] +
                                   sin(PI * P->x[i]) * cos(PI * P->y[i]) * P->y_normal[i]);
        }
    }
}

static int write_level0(const char *name, PointStructure *P, FieldVariables *F)
{
    FILE *f = fopen(name, "wb");
    if (!f) return 1;
    const double *a[] = {P->x, P->y, F->source, F->p};
    int bad = 0;
    for (int k = 0; k < 4; k++) bad |= fwrite(a[k], sizeof(double), P->num_nodes, f) != (size_t)P->num_nodes;
    return (fclose(f) != 0) | bad;
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
    initial_fields(P, F, parameters.num_levels);
    copyin_essentials_to_gpu(P);
    copyin_field_to_gpu(F, P);
    copyin_parameters_to_gpu();

    int L = parameters.num_levels;
    for (int it = 0; it < parameters.num_time_steps; it++) {
        for (int c = 0; c < parameters.num_vcycles; c++) {
            boundary_rhs(&P[0], F);
            for (int l = 0; l < L; l++) {
                relaxation(&P[l], &F[l]);
                calculate_residuals(&P[l], &F[l]);
                if (l != L - 1) restrict_residuals(&P[l], &P[l + 1], &F[l], &F[l + 1]);
            }
            for (int l = L - 1; l > 0; l--) {
                prolongate_corrections(&P[l - 1], &P[l], &F[l - 1], &F[l]);
                if (l != 1) relaxation(&P[l - 1], &F[l - 1]);
            }
        }
    }

    int bad = write_level0("out.bin", &P[0], &F[0]);
    free_PointStructure(P, L);
    free_field(F, L);
    return bad;
}
