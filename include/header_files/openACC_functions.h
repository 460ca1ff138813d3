/*
 * openACC_functions.h -- drop-in replacement for the reference's
 * header_files/openACC_functions.h.  The hierarchy is already resident on the
 * GPU after read_complete_mesh_data; these calls mark the host copies as the
 * current state so the next solver call uploads them.
 */
#ifndef OPENACC_FUNCTIONS_H
#define OPENACC_FUNCTIONS_H

#include "structures.h"
#include "general_functions.h"
#include "mesh_functions.h"
#include "kdtree_functions.h"
#include "mat_lib.h"

#ifdef __cplusplus
extern "C" {
#endif

void copyin_parameters_to_gpu(void);                                           /* :26 */
void copyin_pointstructure_to_gpu(PointStructure *myPointStruct);              /* :30 */
void copyin_field_to_gpu(FieldVariables *field, PointStructure *myPointStruct); /* :67 */
void copyin_essentials_to_gpu(PointStructure *myPointStruct);                  /* :96 */
void copypout_pointstructure_from_gpu(PointStructure *myPointStruct);          /* :63 */
void copypout_field_from_gpu(FieldVariables *field, PointStructure *myPointStruct); /* :92 */

#ifdef __cplusplus
}
#endif
#endif
