/*
 * CP_AL LP solver: public C entry points of libcp_al.so (MI355X build).
 *
 * Each function replaces the reference entry point of the same name and
 * signature in LP/CP_AL/include/cp_al.h:28-47; `optimize` replaces
 * LP/CP_AL/internal/solver.h:27, which the reference CLI links directly.
 * Inputs are deep-copied by create_lp_problem; results are heap objects the
 * caller frees with cp_al_result_free (which, unlike the reference, also frees
 * reduced_cost).  Errors return NULL instead of terminating the process.
 */
#pragma once

#include "cp_al_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference: cp_al.h:30-36, cp_al.c:26-116 */
lp_problem_t *create_lp_problem(const double *objective_c,
                                const matrix_desc_t *A_desc,
                                const double *con_lb,
                                const double *con_ub,
                                const double *var_lb,
                                const double *var_ub,
                                const double *objective_constant);

/* reference: cp_al.h:39, cp_al.c:146-178 */
void set_start_values(lp_problem_t *prob, const double *primal, const double *dual);

/* reference: cp_al.h:42, cp_al.c:180-209.  NULL params -> defaults. */
cp_al_result_t *solve_lp_problem(lp_problem_t *prob, const cp_al_parameters_t *params);

/* reference: cp_al.h:45, utils.cu:361-394 */
void set_default_parameters(cp_al_parameters_t *params);

/* reference: cp_al.h:47, cp_al.c:118-128 */
void cp_al_result_free(cp_al_result_t *results);

/* reference: cp_al.h:49, cp_al.c:130-144 */
void lp_problem_free(lp_problem_t *prob);

/* reference: internal/solver.h:27, solver.cu:1175-1736 */
cp_al_result_t *optimize(const cp_al_parameters_t *params, lp_problem_t *original_problem);

#ifdef __cplusplus
}
#endif
