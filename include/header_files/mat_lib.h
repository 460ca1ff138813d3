/*
 * mat_lib.h -- drop-in replacement for the reference's header_files/mat_lib.h
 * (the functions its drivers, init plug-ins and solver headers call).
 * multiply_sparse_matrix_vector_gpu runs on the GPU through libmgm: operator
 * rows produced by create_derivative_matrices are recognised and applied from
 * their resident device slabs; other row sets are uploaded for the call.
 */
#ifndef MATH_LIBRARY_H
#define MATH_LIBRARY_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

void create_matrix(double ***A, int n_rows, int n_cols);              /* mat_lib.h:131 */
void create_matrix_int(int ***A, int n_rows, int n_cols);             /* mat_lib.h:148 */
double **create_matrix1(int n_rows, int n_cols);                      /* mat_lib.h:114 */
double *create_vector(int n_rows);                                    /* mat_lib.h:158 */
void free_matrix(double **A, int n_rows);                             /* mat_lib.h:169 */
void free_matrix_int(int **A, int n_rows);
void free_vector(double *A);
/* Both run on the GPU (the host variant is kept for its callers' sake). */
void multiply_sparse_matrix_vector(double **D_coeff, double *f, double *dfdx, int **cloud,
                                   int n_rows_D, int n_cols_D);       /* mat_lib.h:70 */
void multiply_sparse_matrix_vector_gpu(double **D_coeff, double *f, double *dfdx, int **cloud,
                                       int n_rows_D, int n_cols_D);   /* mat_lib.h:81 */

#ifdef __cplusplus
}
#endif
#endif
