/*
 * mesh_functions.h -- drop-in replacement for the reference's
 * header_files/mesh_functions.h.  read_complete_mesh_data builds the whole
 * hierarchy on the GPU (stencils, reordering, normals, spacing, transfer
 * points, weights and transfer rows) and fills every host field the reference
 * fills.  cloud_size_multiplier may be fractional in flow_parameters.csv
 * (3.5 -> m = 70 with cubic polynomials in 3D); parameters keeps the truncated
 * short as the reference's struct does.
 */
#ifndef MESH_FUNCTIONS_H
#define MESH_FUNCTIONS_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "structures.h"
#include "kdtree_functions.h"
#include "general_functions.h"
#include "rbf.h"

#ifdef __cplusplus
extern "C" {
#endif

void read_flow_parameters(char *filename);                                     /* :37 */
void read_grid_filenames(PointStructure **myPointStruct, char *filename, short *num_levels); /* :102 */
void calculate_parameters(PointStructure *myPointStruct);                      /* :131 */
void read_complete_mesh_data(PointStructure *myPointStruct, short num_levels); /* :914 */

#ifdef __cplusplus
}
#endif
#endif
