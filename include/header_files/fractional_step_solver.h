/*
 * fractional_step_solver.h -- drop-in replacement for the reference's
 * header_files/fractional_step_solver.h: the explicit fractional step (:47)
 * and its multigrid pressure solve (:150) with the F-variant primitives
 * (:168 relaxation sweeping lap_Poison with residual from lap, :198, :214, :236).
 */
#ifndef FRACTIONAL_STEP_SOLVERS_H
#define FRACTIONAL_STEP_SOLVERS_H

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

static inline double fractional_step_explicit(PointStructure *p, FieldVariables *f) { return mgm_dropin_fractional_step_explicit(p, f); }
static inline void multigrid_Poisson_solver(PointStructure *p, FieldVariables *f) { mgm_dropin_multigrid_Poisson_solver(p, f, 2); }
static inline void relaxation(PointStructure *p, FieldVariables *f) { mgm_dropin_relaxation(p, f, 2); }
static inline void calculate_residuals(PointStructure *p, FieldVariables *f) { mgm_dropin_calculate_residuals(p, f, 2); }
static inline void restrict_residuals(PointStructure *pf, PointStructure *pc, FieldVariables *ff, FieldVariables *fc)
{ mgm_dropin_restrict_residuals(pf, pc, ff, fc, 2); }
static inline void prolongate_corrections(PointStructure *pf, PointStructure *pc, FieldVariables *ff, FieldVariables *fc)
{ mgm_dropin_prolongate_corrections(pf, pc, ff, fc, 2); }

#endif
