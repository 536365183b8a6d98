/*
 * RHR-FA-CP, the planted-KKT FA_CP QP solver: public C entry points of
 * librhr_fa_cp.so (MI355X build).
 *
 * Each function replaces the reference entry point of the same name and
 * signature in QP/PLANTED_FA_CP/include/fa_cp.h (built there as
 * librhr_fa_cp.so, QP/PLANTED_FA_CP/CMakeLists.txt:119-128); `optimize`
 * replaces QP/PLANTED_FA_CP/internal/solver.h.
 * create_qp_problem deep-copies its inputs; solve_qp_problem leaves the
 * problem untouched; fa_cp_result_free also frees reduced_cost.  Errors
 * return NULL instead of terminating the process.
 */
#pragma once

#include "planted_fa_cp_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference: fa_cp.h:28-34, fa_cp.c:28-288 */
qp_problem_t *create_qp_problem(const double *objective_c,
                                const matrix_desc_t *Q_desc,
                                const matrix_desc_t *R_desc,
                                const matrix_desc_t *A_desc,
                                const double *con_lb, const double *con_ub,
                                const double *var_lb, const double *var_ub,
                                const double *objective_constant);

/* reference: fa_cp.h:37-38, fa_cp.c:324-350 */
void set_start_values(qp_problem_t *prob, const double *primal, const double *dual);

/* reference: fa_cp.h:41-42, fa_cp.c:352-376 */
fa_cp_result_t *solve_qp_problem(const qp_problem_t *prob, const fa_cp_parameters_t *params);

/* reference: QP/PLANTED_FA_CP/src/utils.cu:471-512 */
void set_default_parameters(fa_cp_parameters_t *params);

/* reference: fa_cp.h:47, fa_cp.c:290-298 */
void fa_cp_result_free(fa_cp_result_t *results);

/* reference: fa_cp.h:49, fa_cp.c:300-322 */
void qp_problem_free(qp_problem_t *prob);

/* reference: QP/PLANTED_FA_CP/src/solver.cu:106-190 */
fa_cp_result_t *optimize(const fa_cp_parameters_t *params, const qp_problem_t *original_problem);

/* Extras of this build, no counterpart in the reference: reseed the
 * power-iteration generators, and the name of a termination reason. */
void pdal_planted_fa_cp_reset_rng(void);
const char *pdal_planted_fa_cp_termination_reason_to_string(termination_reason_t reason);

#ifdef __cplusplus
}
#endif
