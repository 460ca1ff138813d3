/*
 * timple_solver.h -- drop-in replacement for the reference's
 * header_files/timple_solver.h: the TIMPLE time step (:40) and its
 * multigrid pressure correction (:191) with the T-variant primitives
 * (:235 relaxation with pressure pin and fused residual, :274, :292, :308).
 */
#ifndef TIMPLE_SOLVERS_H
#define TIMPLE_SOLVERS_H

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

static inline double time_implicit_solver(PointStructure *p, FieldVariables *f) { return mgm_dropin_time_implicit_solver(p, f); }
static inline void multigrid_Poisson_solver(PointStructure *p, FieldVariables *f) { mgm_dropin_multigrid_Poisson_solver(p, f, 1); }
static inline void relaxation(PointStructure *p, FieldVariables *f) { mgm_dropin_relaxation(p, f, 1); }
static inline void calculate_residuals(PointStructure *p, FieldVariables *f) { mgm_dropin_calculate_residuals(p, f, 1); }
static inline void restrict_residuals(PointStructure *pf, PointStructure *pc, FieldVariables *ff, FieldVariables *fc)
{ mgm_dropin_restrict_residuals(pf, pc, ff, fc, 1); }
static inline void prolongate_corrections(PointStructure *pf, PointStructure *pc, FieldVariables *ff, FieldVariables *fc)
{ mgm_dropin_prolongate_corrections(pf, pc, ff, fc, 1); }

#endif
