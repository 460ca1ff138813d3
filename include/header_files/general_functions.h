/*
 * general_functions.h -- drop-in replacement for the reference's
 * header_files/general_functions.h (functions on the hot path's call surface).
 */
#ifndef GENERAL_FUNCTIONS_H
#define GENERAL_FUNCTIONS_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "structures.h"
#include "mat_lib.h"

#ifdef __cplusplus
extern "C" {
#endif

double calculate_dt(PointStructure *myPointStruct);                             /* :34 */
/* Applies Dx, Dy, (Dz), lap on the GPU to the reference's manufactured
 * function and prints the l2_norm_gen errors per level, as :44-101 does. */
void test_derivatives(PointStructure *myPointStruct, short num_levels, short domain_dimension);
double l2_norm_gen(PointStructure *myPointStruct, double *A, double *B, int n_rows_A); /* :181 */

#ifdef __cplusplus
}
#endif
#endif
