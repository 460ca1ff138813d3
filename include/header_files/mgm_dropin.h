/*
 * mgm_dropin.h -- variant-specific entry points behind the drop-in solver
 * headers.  The reference defines relaxation, calculate_residuals,
 * restrict_residuals, prolongate_corrections and multigrid_Poisson_solver
 * once per solver header with different semantics; libmgm exports one symbol
 * per variant and heat_conduction.h / timple_solver.h /
 * fractional_step_solver.h map the reference names onto them.
 * variant: 0 heat (H), 1 TIMPLE (T), 2 fractional step (F).
 *
 * Host/device coherence: each call uploads the host arrays it reads (fields,
 * and lap_Poison rows when their contents changed since the last upload) and
 * downloads what it writes, so drivers and init plug-ins may keep reading and
 * writing the host arrays between calls as they do with the reference.
 */
#ifndef MGM_DROPIN_H
#define MGM_DROPIN_H

#include "structures.h"

#ifdef __cplusplus
extern "C" {
#endif

void mgm_dropin_relaxation(PointStructure *P, FieldVariables *F, int variant);
void mgm_dropin_calculate_residuals(PointStructure *P, FieldVariables *F, int variant);
void mgm_dropin_restrict_residuals(PointStructure *Pf, PointStructure *Pc, FieldVariables *Ff,
                                   FieldVariables *Fc, int variant);
void mgm_dropin_prolongate_corrections(PointStructure *Pf, PointStructure *Pc, FieldVariables *Ff,
                                       FieldVariables *Fc, int variant);
void mgm_dropin_multigrid_Poisson_solver(PointStructure *P, FieldVariables *F, int variant);
double mgm_dropin_time_implicit_solver(PointStructure *P, FieldVariables *F);
double mgm_dropin_fractional_step_explicit(PointStructure *P, FieldVariables *F);

#ifdef __cplusplus
}
#endif
#endif
