// Public C entry points of libfa_cp.so, and with -DPDAL_LIN_CP_AL of
// librhr_lin_cp_al.so, -DPDAL_PLANTED_FA_CP of librhr_fa_cp.so and
// -DPDAL_PLANTED_CP_AL of librhr_cp_al.so (the same problem construction,
// e.g. QP/LIN_CP_AL/src/cp_al.c; each its own defaults and result type).
//
// Reference: QP/FA_CP/src/fa_cp.c:28-376 (problem construction, start values,
// solve, frees), QP/FA_CP/src/utils.cu:494-537 (defaults), :661-678 (final
// summary), :1135-1358 (dense / CSC / COO -> CSR with zero_tolerance);
// QP/LIN_CP_AL/src/utils.cu:466-497 (its defaults).
#include "fa_cp_internal.hpp"
#if !defined(PDAL_CP_AL_TYPES) && !defined(PDAL_PLANTED_FA_CP)
#include "pdal_qp.h"
#endif

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <vector>

#define PDAL_API extern "C" __attribute__((visibility("default")))

namespace {

// Triplets (row-major order of arrival) -> CSR by a stable counting sort, so
// duplicates are kept in input order as the reference's converters do.
bool triplets_to_csr(int m, const std::vector<int> &ri, const std::vector<int> &cj, const std::vector<double> &vv,
                     CsrComponent *out, int *nnz_out)
{
    const int nnz = (int)vv.size();
    out->row_ptr = (int *)calloc((size_t)m + 1, sizeof(int));
    out->col_ind = (int *)malloc((size_t)std::max(nnz, 1) * sizeof(int));
    out->val = (double *)malloc((size_t)std::max(nnz, 1) * sizeof(double));
    if (!out->row_ptr || !out->col_ind || !out->val) return false;
    for (int k = 0; k < nnz; ++k) out->row_ptr[ri[k] + 1]++;
    for (int i = 0; i < m; ++i) out->row_ptr[i + 1] += out->row_ptr[i];
    std::vector<int> next(out->row_ptr, out->row_ptr + m);
    for (int k = 0; k < nnz; ++k) {
        const int p = next[ri[k]]++;
        out->col_ind[p] = cj[k];
        out->val[p] = vv[k];
    }
    *nnz_out = nnz;
    return true;
}

bool desc_to_csr(const matrix_desc_t *d, CsrComponent *out, int *nnz_out, const char *name)
{   // dense_to_csr / csc_to_csr / coo_to_csr, utils.cu:1144-1358
    const int m = d->m, n = d->n;
    if (m < 0 || n < 0) return false;
    std::vector<int> ri, cj;
    std::vector<double> vv;
    switch (d->fmt) {
    case matrix_dense: {
        const double tol = d->zero_tolerance > 0 ? d->zero_tolerance : 1e-12;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < n; ++j) {
                const double a = d->data.dense.A[(size_t)i * n + j];
                if (std::fabs(a) > tol) {
                    ri.push_back(i);
                    cj.push_back(j);
                    vv.push_back(a);
                }
            }
        break;
    }
    case matrix_csr: {
        const int nnz = d->data.csr.nnz;
        out->row_ptr = (int *)malloc((size_t)(m + 1) * sizeof(int));
        out->col_ind = (int *)malloc((size_t)std::max(nnz, 1) * sizeof(int));
        out->val = (double *)malloc((size_t)std::max(nnz, 1) * sizeof(double));
        if (!out->row_ptr || !out->col_ind || !out->val) return false;
        memcpy(out->row_ptr, d->data.csr.row_ptr, (size_t)(m + 1) * sizeof(int));
        if (nnz > 0) {
            memcpy(out->col_ind, d->data.csr.col_ind, (size_t)nnz * sizeof(int));
            memcpy(out->val, d->data.csr.vals, (size_t)nnz * sizeof(double));
        }
        *nnz_out = nnz;
        return true;
    }
    case matrix_csc: {
        const double tol = d->zero_tolerance > 0 ? d->zero_tolerance : 0.0;
        const int *cp = d->data.csc.col_ptr;
        for (int j = 0; j < n; ++j)
            for (int k = cp[j]; k < cp[j + 1]; ++k) {
                const int r = d->data.csc.row_ind[k];
                if (r < 0 || r >= m) {
                    fprintf(stderr, "[interface] CSC: row index out of range\n");
                    return false;
                }
                const double a = d->data.csc.vals[k];
                if (tol > 0 && std::fabs(a) <= tol) continue;
                ri.push_back(r);
                cj.push_back(j);
                vv.push_back(a);
            }
        break;
    }
    case matrix_coo: {
        const double tol = d->zero_tolerance > 0 ? d->zero_tolerance : 0.0;
        for (int k = 0; k < d->data.coo.nnz; ++k) {
            const double a = d->data.coo.vals[k];
            if (tol > 0 && std::fabs(a) <= tol) continue;
            const int r = d->data.coo.row_ind[k], c = d->data.coo.col_ind[k];
            if (r < 0 || r >= m) {
                fprintf(stderr, "[interface] COO: row index out of range\n");
                return false;
            }
            if (c < 0 || c >= n) {
                fprintf(stderr, "[interface] COO: col index out of range\n");
                return false;
            }
            ri.push_back(r);
            cj.push_back(c);
            vv.push_back(a);
        }
        break;
    }
    default: fprintf(stderr, "[interface] %s matrix: unsupported format %d.\n", name, (int)d->fmt); return false;
    }
    return triplets_to_csr(m, ri, cj, vv, out, nnz_out);
}

bool fill_or_copy(double **dst, int n, const double *src, double fill)
{   // utils.cu:1135-1142
    *dst = (double *)malloc((size_t)std::max(n, 1) * sizeof(double));
    if (!*dst) return false;
    if (src) {
        if (n > 0) memcpy(*dst, src, (size_t)n * sizeof(double));
    } else {
        for (int i = 0; i < n; ++i) (*dst)[i] = fill;
    }
    return true;
}

void free_component(CsrComponent *c)
{
    if (!c) return;
    free(c->row_ptr);
    free(c->col_ind);
    free(c->val);
    free(c);
}

using pdal::qp::params_t;
using pdal::qp::result_t;

void final_log(const result_t *r, const params_t *p)
{   // fa_cp_final_log, utils.cu:661-678
    const char *quiet = getenv("PDAL_QUIET");
    if (quiet && quiet[0] == '1') return;
    if (p->verbose >= 2)
        printf("---------------------------------------------------------------------------------------\n");
    if (p->verbose < 1) return;
    printf("Solution Summary\n");
    printf("  Status             : %s\n", pdal::qp::termination_reason_name(r->termination_reason));
    printf("  Solve time         : %.3g sec\n", r->cumulative_time_sec);
    printf("  Iterations         : %d\n", r->total_count);
    printf("  Primal objective   : %.10g\n", r->primal_objective_value);
    printf("  Dual objective     : %.10g\n", r->dual_objective_value);
    printf("  Objective gap      : %.3e\n", r->relative_objective_gap);
    printf("  Primal infeas      : %.3e\n", r->relative_primal_residual);
    printf("  Dual infeas        : %.3e\n", r->relative_dual_residual);
}

}  // namespace

#ifdef PDAL_LIN_CP_AL
PDAL_API void set_default_parameters(cp_al_parameters_t *p)
{   // QP/LIN_CP_AL/src/utils.cu:466-497
    memset(p, 0, sizeof(*p));
    p->l_inf_ruiz_iterations = 10;
    p->has_pock_chambolle_alpha = true;
    p->pock_chambolle_alpha = 1.0;
    p->bound_objective_rescaling = true;
    p->use_al_qp = true;
    p->sigma_init = 0.0;
    p->verbose = 1;
    p->termination_evaluation_frequency = 200;
    p->feasibility_polishing = false;
    p->reflection_coefficient = 0.0;
    p->sv_max_iter = 5000;
    p->sv_tol = 1e-4;
    p->termination_criteria.eps_optimal_relative = 1e-4;
    p->termination_criteria.eps_feasible_relative = 1e-4;
    p->termination_criteria.eps_infeasible = 1e-10;
    p->termination_criteria.time_sec_limit = 3600.0;
    p->termination_criteria.iteration_limit = INT32_MAX;
    p->termination_criteria.eps_feas_polish_relative = 1e-6;
    p->restart_params.artificial_restart_threshold = 0.36;
    p->restart_params.sufficient_reduction_for_restart = 0.2;
    p->restart_params.necessary_reduction_for_restart = 0.8;
    p->restart_params.k_p = 0.99;
    p->restart_params.k_i = 0.01;
    p->restart_params.k_d = 0.0;
    p->restart_params.i_smooth = 0.3;
    p->optimality_norm = NORM_TYPE_L_INF;
}
#elif defined(PDAL_PLANTED_CP_AL)
PDAL_API void set_default_parameters(cp_al_parameters_t *p)
{   // QP/PLANTED_CP_AL/src/utils.cu:481-513
    memset(p, 0, sizeof(*p));
    p->l_inf_ruiz_iterations = 10;
    p->has_pock_chambolle_alpha = true;
    p->pock_chambolle_alpha = 1.0;
    p->bound_objective_rescaling = true;
    p->use_al_qp = true;
    p->sigma_init = 0.0;
    p->trajectory_csv = nullptr;
    p->verbose = 1;
    p->termination_evaluation_frequency = 200;
    p->feasibility_polishing = false;
    p->reflection_coefficient = 1.0;
    p->sv_max_iter = 5000;
    p->sv_tol = 1e-4;
    p->termination_criteria.eps_optimal_relative = 1e-4;
    p->termination_criteria.eps_feasible_relative = 1e-4;
    p->termination_criteria.eps_infeasible = 1e-10;
    p->termination_criteria.time_sec_limit = 3600.0;
    p->termination_criteria.iteration_limit = INT32_MAX;
    p->termination_criteria.eps_feas_polish_relative = 1e-6;
    p->restart_params.artificial_restart_threshold = 0.36;
    p->restart_params.sufficient_reduction_for_restart = 0.2;
    p->restart_params.necessary_reduction_for_restart = 0.8;
    p->restart_params.k_p = 0.99;
    p->restart_params.k_i = 0.01;
    p->restart_params.k_d = 0.0;
    p->restart_params.i_smooth = 0.3;
    p->optimality_norm = NORM_TYPE_L_INF;
}
#elif defined(PDAL_PLANTED_FA_CP)
PDAL_API void set_default_parameters(fa_cp_parameters_t *p)
{   // QP/PLANTED_FA_CP/src/utils.cu:471-512
    memset(p, 0, sizeof(*p));
    p->l_inf_ruiz_iterations = 10;
    p->has_pock_chambolle_alpha = true;
    p->pock_chambolle_alpha = 1.0;
    p->bound_objective_rescaling = true;
    p->use_al_qp = true;
    p->al_sigma = 0.1;
    p->adaptive_al_sigma = false;
    p->adaptive_al_sigma_mode = ADAPTIVE_AL_SIGMA_MODE_BIDIRECTIONAL;
    p->adaptive_al_sigma_rule = ADAPTIVE_AL_SIGMA_RULE_TRIGGER;
    p->al_sigma_min = 5e-3;
    p->al_sigma_max = 3e-1;
    p->al_sigma_increase_factor = 1.5;
    p->al_sigma_decrease_factor = 0.7;
    p->al_sigma_primal_trigger = 3.0;
    p->al_sigma_gap_trigger = 5.0;
    p->trajectory_csv = nullptr;
    p->verbose = 1;
    p->termination_evaluation_frequency = 200;
    p->feasibility_polishing = false;
    p->reflection_coefficient = 1.0;
    p->sv_max_iter = 5000;
    p->sv_tol = 1e-4;
    p->termination_criteria.eps_optimal_relative = 1e-4;
    p->termination_criteria.eps_feasible_relative = 1e-4;
    p->termination_criteria.eps_infeasible = 1e-10;
    p->termination_criteria.time_sec_limit = 3600.0;
    p->termination_criteria.iteration_limit = INT32_MAX;
    p->termination_criteria.eps_feas_polish_relative = 1e-6;
    p->restart_params.artificial_restart_threshold = 0.36;
    p->restart_params.sufficient_reduction_for_restart = 0.2;
    p->restart_params.necessary_reduction_for_restart = 0.8;
    p->restart_params.k_p = 0.99;
    p->restart_params.k_i = 0.01;
    p->restart_params.k_d = 0.0;
    p->restart_params.i_smooth = 0.3;
    p->optimality_norm = NORM_TYPE_L_INF;
}
#else
PDAL_API void set_default_parameters(fa_cp_parameters_t *p)
{   // utils.cu:494-537
    memset(p, 0, sizeof(*p));
    p->l_inf_ruiz_iterations = 10;
    p->has_pock_chambolle_alpha = true;
    p->pock_chambolle_alpha = 1.0;
    p->bound_objective_rescaling = true;
    p->trajectory_csv = nullptr;
    p->use_al_qp = true;
    p->al_sigma = 0.0;
    p->sigma_update_rule = FA_CP_SIGMA_UPDATE_GUARDED;
    p->sigma_update_gain = 0.5;
    p->sigma_clip_low = 0.6;
    p->sigma_clip_high = 1.7;
    p->sigma_min = 1e-8;
    p->sigma_max = 10.0;
    p->sigma_rollback_tol = 1.25;
    p->sigma_zero_threshold = 1e-8;
    p->sigma_activation_threshold = 2.0;
    p->sigma_normal_motion_threshold = 1e-10;
    p->sigma_activation_value = 0.1;
    p->verbose = 1;
    p->termination_evaluation_frequency = 200;
    p->feasibility_polishing = false;
    p->reflection_coefficient = 1.0;
    p->sv_max_iter = 5000;
    p->sv_tol = 1e-4;
    p->termination_criteria.eps_optimal_relative = 1e-4;
    p->termination_criteria.eps_feasible_relative = 1e-4;
    p->termination_criteria.eps_infeasible = 1e-10;
    p->termination_criteria.time_sec_limit = 3600.0;
    p->termination_criteria.iteration_limit = INT32_MAX;
    p->termination_criteria.eps_feas_polish_relative = 1e-6;
    p->restart_params.artificial_restart_threshold = 0.36;
    p->restart_params.sufficient_reduction_for_restart = 0.2;
    p->restart_params.necessary_reduction_for_restart = 0.8;
    p->restart_params.k_p = 0.99;
    p->restart_params.k_i = 0.01;
    p->restart_params.k_d = 0.0;
    p->restart_params.i_smooth = 0.3;
    p->optimality_norm = NORM_TYPE_L_INF;
}
#endif

PDAL_API void qp_problem_free(qp_problem_t *p)
{   // fa_cp.c:300-318
    if (!p) return;
    free_component(p->constraint_matrix);
    free_component(p->objective_sparse_matrix);
    free_component(p->objective_lowrank_matrix);
    free(p->variable_lower_bound);
    free(p->variable_upper_bound);
    free(p->objective_vector);
    free(p->constraint_lower_bound);
    free(p->constraint_upper_bound);
    free(p->primal_start);
    free(p->dual_start);
    free(p);
}

PDAL_API qp_problem_t *create_qp_problem(const double *c, const matrix_desc_t *Q, const matrix_desc_t *R,
                                         const matrix_desc_t *A, const double *con_lb, const double *con_ub,
                                         const double *var_lb, const double *var_ub, const double *obj_const)
{   // fa_cp.c:28-268: n from A, else Q, else R; m from A only
    int n, m = 0;
    if (A) {
        n = A->n;
        m = A->m;
    } else if (Q) {
        n = Q->n;
    } else if (R) {
        n = R->n;
    } else {
        fprintf(stderr,
                "[interface] Error: At least one matrix (A, Q, or R) must be provided for non-trivil optimization "
                "problem.\n");
        return nullptr;
    }
    if (n == 0) fprintf(stderr, "[interface] Warning: Matrix dimensions seem to be zero or inconsistent.\n");
    qp_problem_t *p = (qp_problem_t *)calloc(1, sizeof(qp_problem_t));
    if (!p) return nullptr;
    p->num_variables = n;
    p->num_constraints = m;
    p->constraint_matrix = (CsrComponent *)calloc(1, sizeof(CsrComponent));
    p->objective_sparse_matrix = (CsrComponent *)calloc(1, sizeof(CsrComponent));
    p->objective_lowrank_matrix = (CsrComponent *)calloc(1, sizeof(CsrComponent));
    bool ok = p->constraint_matrix && p->objective_sparse_matrix && p->objective_lowrank_matrix;
    auto empty = [](CsrComponent *cc) { cc->row_ptr = (int *)calloc(1, sizeof(int)); };
    if (ok) {
        if (A) ok = desc_to_csr(A, p->constraint_matrix, &p->constraint_matrix_num_nonzeros, "A");
        else empty(p->constraint_matrix);
    }
    if (ok) {
        if (Q) {
            if (Q->m != n || Q->n != n) {
                fprintf(stderr, "[interface] Q matrix must be %d x %d.\n", n, n);
                ok = false;
            } else {
                ok = desc_to_csr(Q, p->objective_sparse_matrix, &p->objective_sparse_matrix_num_nonzeros, "Q");
            }
        } else {
            empty(p->objective_sparse_matrix);
        }
    }
    if (ok) {
        if (R) {
            if (R->n != n) {
                fprintf(stderr, "[interface] R matrix must have %d columns.\n", n);
                ok = false;
            } else {
                p->num_rank_lowrank_obj = R->m;
                ok = desc_to_csr(R, p->objective_lowrank_matrix, &p->objective_lowrank_matrix_num_nonzeros, "R");
            }
        } else {
            empty(p->objective_lowrank_matrix);
        }
    }
    if (ok) {
        p->objective_constant = obj_const ? *obj_const : 0.0;
        ok = fill_or_copy(&p->objective_vector, n, c, 0.0) &&
             fill_or_copy(&p->variable_lower_bound, n, var_lb, -INFINITY) &&
             fill_or_copy(&p->variable_upper_bound, n, var_ub, INFINITY) &&
             fill_or_copy(&p->constraint_lower_bound, m, con_lb, -INFINITY) &&
             fill_or_copy(&p->constraint_upper_bound, m, con_ub, INFINITY);
    }
    if (!ok) {
        qp_problem_free(p);
        return nullptr;
    }
    return p;
}

PDAL_API void set_start_values(qp_problem_t *p, const double *primal, const double *dual)
{   // fa_cp.c:320-350
    if (!p) return;
    free(p->primal_start);
    free(p->dual_start);
    p->primal_start = p->dual_start = nullptr;
    if (primal) {
        p->primal_start = (double *)malloc((size_t)std::max(p->num_variables, 1) * sizeof(double));
        if (p->primal_start && p->num_variables > 0)
            memcpy(p->primal_start, primal, (size_t)p->num_variables * sizeof(double));
    }
    if (dual) {
        p->dual_start = (double *)malloc((size_t)std::max(p->num_constraints, 1) * sizeof(double));
        if (p->dual_start && p->num_constraints > 0)
            memcpy(p->dual_start, dual, (size_t)p->num_constraints * sizeof(double));
    }
}

#ifdef PDAL_CP_AL_TYPES
PDAL_API void cp_al_result_free(cp_al_result_t *r)
#else
PDAL_API void fa_cp_result_free(fa_cp_result_t *r)
#endif
{   // fa_cp.c:270-277; reduced_cost is freed too (the reference leaks it)
    if (!r) return;
    free(r->primal_solution);
    free(r->dual_solution);
    free(r->reduced_cost);
    free(r);
}

PDAL_API result_t *optimize(const params_t *params, const qp_problem_t *prob)
{   // solver.cu:74-168.  Errors return NULL instead of exiting the process.
    if (!params || !prob) return nullptr;
    if (params->termination_evaluation_frequency <= 0) {
        fprintf(stderr, "[fa_cp] termination_evaluation_frequency must be positive.\n");
        return nullptr;
    }
    try {
        result_t *r = pdal::qp::optimize_gpu(params, prob);
        if (r) final_log(r, params);
        return r;
    } catch (const std::exception &e) {
        fprintf(stderr, "[fa_cp] %s\n", e.what());
        return nullptr;
    }
}

PDAL_API result_t *solve_qp_problem(const qp_problem_t *prob, const params_t *params)
{   // fa_cp.c:352-376
    if (!prob) {
        fprintf(stderr, "[interface] solve_qp_problem: invalid arguments.\n");
        return nullptr;
    }
    params_t local;
    if (params) local = *params;
    else set_default_parameters(&local);
    result_t *res = optimize(&local, prob);
    if (!res) fprintf(stderr, "[interface] optimize returned NULL.\n");
    return res;
}

#if defined(PDAL_LIN_CP_AL)
PDAL_API void pdal_lin_cp_al_reset_rng(void) { pdal::qp::reset_rng(); }

PDAL_API const char *pdal_lin_cp_al_termination_reason_to_string(termination_reason_t r)
{
    return pdal::qp::termination_reason_name(r);
}
#elif defined(PDAL_PLANTED_CP_AL)
PDAL_API void pdal_planted_cp_al_reset_rng(void) { pdal::qp::reset_rng(); }

PDAL_API const char *pdal_planted_cp_al_termination_reason_to_string(termination_reason_t r)
{
    return pdal::qp::termination_reason_name(r);
}
#elif defined(PDAL_PLANTED_FA_CP)
PDAL_API void pdal_planted_fa_cp_reset_rng(void) { pdal::qp::reset_rng(); }

PDAL_API const char *pdal_planted_fa_cp_termination_reason_to_string(termination_reason_t r)
{
    return pdal::qp::termination_reason_name(r);
}
#else
PDAL_API void pdal_fa_cp_reset_rng(void) { pdal::qp::reset_rng(); }

PDAL_API const char *pdal_fa_cp_termination_reason_to_string(termination_reason_t r)
{
    return pdal::qp::termination_reason_name(r);
}
#endif
