/*
 * RHR-CP-AL, the planted-KKT CP-AL QP solver: public C entry points of
 * librhr_cp_al.so (MI355X build).
 *
 * Each function replaces the reference entry point of the same name and
 * signature in QP/PLANTED_CP_AL/include/cp_al.h (built there as
 * librhr_cp_al.so, QP/PLANTED_CP_AL/CMakeLists.txt:119-128); `optimize`
 * replaces QP/PLANTED_CP_AL/internal/solver.h, which the reference's CLI calls.
 * create_qp_problem deep-copies its inputs; solve_qp_problem leaves the
 * problem untouched; cp_al_result_free also frees reduced_cost.  Errors
 * return NULL instead of terminating the process.
 */
#pragma once

#include "planted_cp_al_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference: QP/PLANTED_CP_AL/include/cp_al.h, src/cp_al.c */
qp_problem_t *create_qp_problem(const double *objective_c,
                                const matrix_desc_t *Q_desc,
                                const matrix_desc_t *R_desc,
                                const matrix_desc_t *A_desc,
                                const double *con_lb, const double *con_ub,
                                const double *var_lb, const double *var_ub,
                                const double *objective_constant);

void set_start_values(qp_problem_t *prob, const double *primal, const double *dual);

/* reference: QP/PLANTED_CP_AL/src/cp_al.c (solve_qp_problem) */
cp_al_result_t *solve_qp_problem(const qp_problem_t *prob, const cp_al_parameters_t *params);

/* reference: QP/PLANTED_CP_AL/src/utils.cu:481-513 */
void set_default_parameters(cp_al_parameters_t *params);

void cp_al_result_free(cp_al_result_t *results);

void qp_problem_free(qp_problem_t *prob);

/* reference: QP/PLANTED_CP_AL/src/solver.cu:111-204 */
cp_al_result_t *optimize(const cp_al_parameters_t *params, const qp_problem_t *original_problem);

/* Extras of this build, no counterpart in the reference (as pdal_qp.h for
 * FA_CP): reseed the power-iteration generators (std::mt19937(1) and srand(1),
 * process-global there), and the name of a termination reason. */
void pdal_planted_cp_al_reset_rng(void);
const char *pdal_planted_cp_al_termination_reason_to_string(termination_reason_t reason);

#ifdef __cplusplus
}
#endif
