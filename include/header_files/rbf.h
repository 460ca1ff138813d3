/*
 * rbf.h -- drop-in replacement for the reference's header_files/rbf.h.
 * The local PHS + polynomial systems are solved as one batched GPU launch per
 * level during read_complete_mesh_data; create_derivative_matrices allocates
 * the reference's double** rows and fills them from the device.
 */
#ifndef RBF_H
#define RBF_H

#include "structures.h"

#ifdef __cplusplus
extern "C" {
#endif

void create_derivative_matrices(PointStructure *myPointStruct);                /* :401 */

#ifdef __cplusplus
}
#endif
#endif
