/*
 * heat_conduction.h -- drop-in replacement for the reference's
 * header_files/heat_conduction.h: the heat/Poisson V-cycle primitives
 * (:25 relaxation, :58 restrict_residuals, :79 prolongate_corrections,
 * :101 calculate_residuals) on the GPU.
 */
#ifndef HEAT_CONDUCTION_H
#define HEAT_CONDUCTION_H

#include <assert.h>
#include <math.h>
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "structures.h"
#include "general_functions.h"
#include "mesh_functions.h"
#include "mat_lib.h"
#include "kdtree_functions.h"
#include "rbf.h"
#include "mgm_dropin.h"

static inline void relaxation(PointStructure *p, FieldVariables *f) { mgm_dropin_relaxation(p, f, 0); }
static inline void calculate_residuals(PointStructure *p, FieldVariables *f) { mgm_dropin_calculate_residuals(p, f, 0); }
static inline void restrict_residuals(PointStructure *pf, PointStructure *pc, FieldVariables *ff, FieldVariables *fc)
{ mgm_dropin_restrict_residuals(pf, pc, ff, fc, 0); }
static inline void prolongate_corrections(PointStructure *pf, PointStructure *pc, FieldVariables *ff, FieldVariables *fc)
{ mgm_dropin_prolongate_corrections(pf, pc, ff, fc, 0); }

#endif
