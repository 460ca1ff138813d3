/*
 * kdtree_functions.h -- drop-in replacement for the reference's
 * header_files/kdtree_functions.h.  Stencils are found on the GPU with the
 * reference's ordering (sqrt distance, ties in the order its kd-tree range
 * query reports them).
 */
#ifndef KDTREE_FUNCTIONS_H
#define KDTREE_FUNCTIONS_H

#include "structures.h"

#ifdef __cplusplus
extern "C" {
#endif

void find_cloud_index(PointStructure *myPointStruct1);                          /* :113 */
int *find_nearest_point(PointStructure *myPointStruct1, PointStructure *myPointStruct2,
                        int num_cloud_points);                                  /* :153 */

#ifdef __cplusplus
}
#endif
#endif
