/*
 * structures.h -- drop-in replacement for the reference's header_files/structures.h.
 *
 * The three structs keep the reference's field order and types exactly
 * (header_files/structures.h:15-112 in the reference), because drivers and
 * init plug-ins index their members directly.  Unlike the reference this
 * header only declares: the single `parameters` object and the functions live
 * in libmgm.so.  Link drivers with -lmgm.
 */
#ifndef STRUCTURES_H
#define STRUCTURES_H

#include <stdbool.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

struct parameters {                  /* header_files/structures.h:15-41 */
    short dimension;
    short poly_degree;
    short phs_degree;
    short cloud_size_multiplier;
    short test;
    double courant_number;
    double steady_state_tolerance;
    double poisson_solver_tolerance;
    short num_vcycles;
    short num_relax;
    int num_time_steps;
    short num_levels;
    int write_interval;
    bool neumann_flag_boundary;
    float omega;
    double dt;
    short iter_momentum;
    short iter_timple;
    double rho;
    double mu;
    double Re;
    double facRe;
    double facdt;
    double nu;
};
extern struct parameters parameters;   /* defined once, in libmgm.so */

typedef struct PointStructure {      /* header_files/structures.h:44-78 */
    char mesh_filename[50];
    int num_nodes;
    int num_corners;
    int num_boundary_nodes;
    int num_elem;
    double d_avg;
    short num_poly_terms;
    short num_cloud_points;
    short poly_degree;
    double *x;
    double *y;
    double *z;
    int *point_index;
    double *x_normal;
    double *y_normal;
    double *z_normal;
    bool *boundary_tag;
    bool *corner_tag;
    short *pow_x;
    short *pow_y;
    short *pow_z;
    int **cloud_index;
    int *prolongation_points;
    int *restriction_points;
    double **Dx;
    double **Dy;
    double **Dz;
    double **lap;
    double **lap_Poison;
    double **restr_mat;
    double **prol_mat;
    int **prolongation_points2;
    int **restriction_points2;
} PointStructure;

typedef struct FieldVariables {      /* header_files/structures.h:80-106 */
    double *u;
    double *v;
    double *w;
    double *p;
    double *u_new;
    double *v_new;
    double *w_new;
    double *u_old;
    double *v_old;
    double *w_old;
    double *p_old;
    double *pprime;
    double *res;
    double *source;
    double *T;
    double *dpdn;
    double *dpdx;
    double *dpdy;
    double *dpdz;
    double rho;
    double mu;
    double nu;
    double Re;
} FieldVariables;

typedef struct Point {               /* header_files/structures.h:109-112 */
    double coords[3];
    int index;
} Point;

void AllocateMemoryPointStructure(PointStructure *myPointStruct, int nodes);
void AllocateMemoryFieldVariables(FieldVariables **field, PointStructure *myPointStruct, int num_levels);
/* Frees only what was allocated (the reference's version faults on Dz in 2D
 * and on the end levels' transfer rows). */
void free_PointStructure(PointStructure *myPointStruct, int num_levels);
void free_field(FieldVariables *field, int num_levels);

#ifdef __cplusplus
}
#endif
#endif
