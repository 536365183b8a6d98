// Public C entry points of libcp_al.so.
//
// Reference: LP/CP_AL/src/cp_al.c:26-209 (problem construction, start values,
// solve, frees), LP/CP_AL/src/utils.cu:361-394 (defaults), :396-478 (verbose
// zero filter), :498-611 (initial info and final summary), :1032-1192
// (dense / CSC / COO -> CSR conversions).
#include "cp_al.h"
#include "cp_al_internal.hpp"
#include "pdal_batch.h"
#include "presolve.hpp"
#include "pdal_lp.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <vector>

#define PDAL_API extern "C" __attribute__((visibility("default")))

#ifndef PDAL_VERSION
#define PDAL_VERSION "0.1.0-mi355x"
#endif

namespace pdal {
namespace lp {

double env_double(const char *name, double def)
{
    const char *raw = getenv(name);
    if (!raw || !raw[0]) return def;
    char *end = nullptr;
    const double v = strtod(raw, &end);
    if (end == raw || !std::isfinite(v)) return def;
    return v;
}

int env_int(const char *name, int def)
{
    const char *raw = getenv(name);
    if (!raw || !raw[0]) return def;
    char *end = nullptr;
    const long v = strtol(raw, &end, 10);
    if (end == raw) return def;
    return (int)v;
}

bool env_set(const char *name)
{
    const char *raw = getenv(name);
    return raw && raw[0] != '\0';
}

void warn_unsupported_controllers(bool batch)
{
    // Controllers of the reference that are switched on only through these
    // variables.  The single-instance solver runs all of them; the batched
    // solver runs none of them (it keeps the restart knobs, utils.cu:324-332,
    // Curtis-Reid without its early stop, and sigma mode).
    static const char *const batched[] = {"CP_AL_LOG_FREQ",
                                          "CP_AL_SIGMA_TRIGGER",
                                          "CP_AL_ADAPTIVE_LAMBDA",
                                          "CP_AL_LAMBDA_PROFILE",
                                          "CP_AL_DYNAMIC_EVAL",
                                          "CP_AL_DYNAMIC_EVAL_AFTER_SEC",
                                          "CP_AL_DYNAMIC_EVAL_AFTER_ITERS",
                                          "CP_AL_CR_EARLY_STOP",
                                          "CP_AL_ALPHA_ROLLBACK_TRIAL",
                                          "CP_AL_ALPHA_DIRECT_PULSE",
                                          "CP_AL_RESTART_SELECT"};
    if (!batch) return;
    for (const char *knob : batched) {
        const char *raw = getenv(knob);
        if (raw && raw[0] != '\0' && strcmp(raw, "0") != 0)
            fprintf(stderr, "[cp_al-mi355x] %s=%s requests a controller the batched solver does not implement; "
                            "ignored\n",
                    knob, raw);
    }
}

namespace {

void fill_or_copy(double **dst, int n, const double *src, double fill)
{
    *dst = (double *)malloc((size_t)std::max(n, 1) * sizeof(double));
    if (src) memcpy(*dst, src, (size_t)n * sizeof(double));
    else
        for (int i = 0; i < n; ++i) (*dst)[i] = fill;
}

// Stable row bucketing of (row, col, val) triplets given in storage order.
bool triplets_to_csr(int m, int n, int nnz, const int *rows, const int *cols, const double *vals, int **rp_out,
                     int **ci_out, double **v_out, const char *what)
{
    int *rp = (int *)calloc((size_t)m + 1, sizeof(int));
    for (int k = 0; k < nnz; ++k) {
        if (rows[k] < 0 || rows[k] >= m) {
            fprintf(stderr, "[interface] %s: row index out of range\n", what);
            free(rp);
            return false;
        }
        if (cols[k] < 0 || cols[k] >= n) {
            fprintf(stderr, "[interface] %s: col index out of range\n", what);
            free(rp);
            return false;
        }
        rp[rows[k] + 1]++;
    }
    for (int i = 0; i < m; ++i) rp[i + 1] += rp[i];
    int *ci = (int *)malloc((size_t)std::max(nnz, 1) * sizeof(int));
    double *v = (double *)malloc((size_t)std::max(nnz, 1) * sizeof(double));
    std::vector<int> next(rp, rp + m);
    for (int k = 0; k < nnz; ++k) {
        const int p = next[rows[k]]++;
        ci[p] = cols[k];
        v[p] = vals[k];
    }
    *rp_out = rp;
    *ci_out = ci;
    *v_out = v;
    return true;
}

void filter_matrix(lp_problem_t *pr, double tol, bool verbose)
{   // filter_constraint_matrix_entries, utils.cu:396-478
    const int rows = pr->num_constraints, nnz = pr->constraint_matrix_num_nonzeros;
    if (rows == 0 || nnz == 0) return;
    int kept = 0;
    for (int k = 0; k < nnz; ++k)
        if (std::fabs(pr->constraint_matrix_values[k]) > tol) ++kept;
    if (kept == nnz) return;
    int *rp = (int *)malloc((size_t)(rows + 1) * sizeof(int));
    int *ci = (int *)malloc((size_t)std::max(kept, 1) * sizeof(int));
    double *v = (double *)malloc((size_t)std::max(kept, 1) * sizeof(double));
    int pos = 0;
    rp[0] = 0;
    for (int i = 0; i < rows; ++i) {
        for (int k = pr->constraint_matrix_row_pointers[i]; k < pr->constraint_matrix_row_pointers[i + 1]; ++k) {
            const double a = pr->constraint_matrix_values[k];
            if (std::fabs(a) <= tol) continue;
            ci[pos] = pr->constraint_matrix_col_indices[k];
            v[pos] = a;
            ++pos;
        }
        rp[i + 1] = pos;
    }
    free(pr->constraint_matrix_row_pointers);
    free(pr->constraint_matrix_col_indices);
    free(pr->constraint_matrix_values);
    pr->constraint_matrix_row_pointers = rp;
    pr->constraint_matrix_col_indices = ci;
    pr->constraint_matrix_values = v;
    pr->constraint_matrix_num_nonzeros = kept;
    if (verbose)
        printf("Dropped %d near-zero %s (|value| <= %.1e) from the constraint matrix\n", nnz - kept,
               nnz - kept == 1 ? "entry" : "entries", tol);
}

const char *reason_name(termination_reason_t r)
{   // termination_reason_to_string, utils.cu:230-253
    switch (r) {
    case TERMINATION_REASON_OPTIMAL: return "OPTIMAL";
    case TERMINATION_REASON_PRIMAL_INFEASIBLE: return "PRIMAL_INFEASIBLE";
    case TERMINATION_REASON_DUAL_INFEASIBLE: return "DUAL_INFEASIBLE";
    case TERMINATION_REASON_INFEASIBLE_OR_UNBOUNDED: return "INFEASIBLE_OR_UNBOUNDED";
    case TERMINATION_REASON_TIME_LIMIT: return "TIME_LIMIT";
    case TERMINATION_REASON_ITERATION_LIMIT: return "ITERATION_LIMIT";
    case TERMINATION_REASON_UNSPECIFIED: return "UNSPECIFIED";
    case TERMINATION_REASON_FEAS_POLISH_SUCCESS: return "FEAS_POLISH_SUCCESS";
    default: return "UNKNOWN";
    }
}

void print_initial_info(const cp_al_parameters_t *p, lp_problem_t *prob)
{   // print_initial_info, utils.cu:498-560
    if (!p->verbose) return;
    printf("---------------------------------------------------------------------------------------\n");
    printf("                        CP_AL v%s (MI355X / gfx950 HIP build)\n", PDAL_VERSION);
    printf("---------------------------------------------------------------------------------------\n");
    filter_matrix(prob, p->matrix_zero_tol, true);
    printf("Problem: %d rows, %d columns, %d nonzeros\n", prob->num_constraints, prob->num_variables,
           prob->constraint_matrix_num_nonzeros);
    printf("Settings:\n");
    printf("  iter_limit         : %d\n", p->termination_criteria.iteration_limit);
    printf("  time_limit         : %.2f sec\n", p->termination_criteria.time_sec_limit);
    printf("  eps_opt            : %.1e\n", p->termination_criteria.eps_optimal_relative);
    printf("  eps_feas           : %.1e\n", p->termination_criteria.eps_feasible_relative);
    if (p->optimality_norm != NORM_TYPE_L2) printf("  optimality_norm    : L_inf\n");
}

void final_log(const cp_al_result_t *r, const cp_al_parameters_t *p)
{   // cp_al_final_log, utils.cu:564-585
    if (p->verbose) printf("---------------------------------------------------------------------------------------\n");
    printf("Solution Summary\n");
    printf("  Status                 : %s\n", reason_name(r->termination_reason));
    printf("  Precondition time      : %.5g sec\n", r->rescaling_time_sec);
    printf("  Solve time             : %.3g sec\n", r->cumulative_time_sec);
    printf("  Iterations             : %d\n", r->total_count);
    printf("  Primal objective       : %.10g\n", r->primal_objective_value);
    printf("  Dual objective         : %.10g\n", r->dual_objective_value);
    printf("  Objective gap          : %.3e\n", r->relative_objective_gap);
    printf("  Primal infeas          : %.3e\n", r->relative_primal_residual);
    printf("  Dual infeas            : %.3e\n", r->relative_dual_residual);
    fflush(stdout);
}

bool quiet() { return env_int("PDAL_QUIET", 0) != 0; }

}  // namespace
}  // namespace lp
}  // namespace pdal

using namespace pdal::lp;

PDAL_API void set_default_parameters(cp_al_parameters_t *p)
{   // utils.cu:361-394
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->l_inf_ruiz_iterations = 10;
    p->has_pock_chambolle_alpha = true;
    p->pock_chambolle_alpha = 1.0;
    p->bound_objective_rescaling = true;
    p->verbose = false;
    p->termination_evaluation_frequency = 200;
    p->feasibility_polishing = false;
    p->reflection_coefficient = env_double("CP_AL_REFLECTION_COEFFICIENT", 1.0);
    p->reflection_coefficient = std::fmax(0.0, std::fmin(p->reflection_coefficient, 2.0));
    p->sv_max_iter = 5000;
    p->sv_tol = 1e-4;
    p->termination_criteria.eps_optimal_relative = 1e-4;
    p->termination_criteria.eps_feasible_relative = 1e-4;
    p->termination_criteria.time_sec_limit = 3600.0;
    p->termination_criteria.iteration_limit = INT32_MAX;
    p->termination_criteria.eps_feas_polish_relative = 1e-6;
    p->restart_params.artificial_restart_threshold = 0.36;
    p->restart_params.sufficient_reduction_for_restart = 0.2;
    p->restart_params.necessary_reduction_for_restart = 0.5;
    p->restart_params.k_p = 0.99;
    p->restart_params.k_i = 0.01;
    p->restart_params.k_d = 0.0;
    p->restart_params.i_smooth = 0.3;
    p->optimality_norm = NORM_TYPE_L2;
    p->presolve = true;
    p->matrix_zero_tol = 1e-9;
}

PDAL_API lp_problem_t *create_lp_problem(const double *c, const matrix_desc_t *A, const double *con_lb,
                                         const double *con_ub, const double *var_lb, const double *var_ub,
                                         const double *obj_const)
{   // cp_al.c:26-116
    if (!A || A->m < 0 || A->n < 0) return nullptr;
    const int m = A->m, n = A->n;
    int *rp = nullptr, *ci = nullptr;
    double *v = nullptr;
    int nnz = 0;
    switch (A->fmt) {
    case matrix_dense: {  // dense_to_csr, utils.cu:1042-1078: exact zeros dropped
        for (long i = 0; i < (long)m * n; ++i)
            if (std::fabs(A->data.dense.A[i]) > 0.0) ++nnz;
        rp = (int *)malloc((size_t)(m + 1) * sizeof(int));
        ci = (int *)malloc((size_t)std::max(nnz, 1) * sizeof(int));
        v = (double *)malloc((size_t)std::max(nnz, 1) * sizeof(double));
        int nz = 0;
        for (int i = 0; i < m; ++i) {
            rp[i] = nz;
            for (int j = 0; j < n; ++j) {
                const double a = A->data.dense.A[(long)i * n + j];
                if (std::fabs(a) > 0.0) {
                    ci[nz] = j;
                    v[nz] = a;
                    ++nz;
                }
            }
        }
        rp[m] = nz;
        break;
    }
    case matrix_csr:
        nnz = A->data.csr.nnz;
        if (nnz < 0 || (nnz > 0 && (!A->data.csr.row_ptr || !A->data.csr.col_ind || !A->data.csr.vals))) return nullptr;
        rp = (int *)malloc((size_t)(m + 1) * sizeof(int));
        ci = (int *)malloc((size_t)std::max(nnz, 1) * sizeof(int));
        v = (double *)malloc((size_t)std::max(nnz, 1) * sizeof(double));
        if (A->data.csr.row_ptr) memcpy(rp, A->data.csr.row_ptr, (size_t)(m + 1) * sizeof(int));
        else
            for (int i = 0; i <= m; ++i) rp[i] = 0;
        if (nnz) {
            memcpy(ci, A->data.csr.col_ind, (size_t)nnz * sizeof(int));
            memcpy(v, A->data.csr.vals, (size_t)nnz * sizeof(double));
        }
        break;
    case matrix_csc: {  // csc_to_csr, utils.cu:1080-1136
        nnz = A->data.csc.nnz;
        std::vector<int> rows(nnz), cols(nnz);
        for (int j = 0; j < n; ++j)
            for (int k = A->data.csc.col_ptr[j]; k < A->data.csc.col_ptr[j + 1]; ++k) {
                rows[k] = A->data.csc.row_ind[k];
                cols[k] = j;
            }
        if (!triplets_to_csr(m, n, nnz, rows.data(), cols.data(), A->data.csc.vals, &rp, &ci, &v, "CSC")) {
            fprintf(stderr, "[interface] CSC->CSR failed.\n");
            return nullptr;
        }
        break;
    }
    case matrix_coo:  // coo_to_csr, utils.cu:1138-1192
        nnz = A->data.coo.nnz;
        if (!triplets_to_csr(m, n, nnz, A->data.coo.row_ind, A->data.coo.col_ind, A->data.coo.vals, &rp, &ci, &v,
                             "COO")) {
            fprintf(stderr, "[interface] COO->CSR failed.\n");
            return nullptr;
        }
        break;
    default:
        fprintf(stderr, "[interface] make_problem_from_matrix: unsupported matrix format %d.\n", (int)A->fmt);
        return nullptr;
    }
    lp_problem_t *p = (lp_problem_t *)calloc(1, sizeof(lp_problem_t));
    p->num_variables = n;
    p->num_constraints = m;
    p->constraint_matrix_row_pointers = rp;
    p->constraint_matrix_col_indices = ci;
    p->constraint_matrix_values = v;
    p->constraint_matrix_num_nonzeros = nnz;
    p->objective_constant = obj_const ? *obj_const : 0.0;
    fill_or_copy(&p->objective_vector, n, c, 0.0);
    fill_or_copy(&p->variable_lower_bound, n, var_lb, -INFINITY);
    fill_or_copy(&p->variable_upper_bound, n, var_ub, INFINITY);
    fill_or_copy(&p->constraint_lower_bound, m, con_lb, -INFINITY);
    fill_or_copy(&p->constraint_upper_bound, m, con_ub, INFINITY);
    return p;
}

PDAL_API void set_start_values(lp_problem_t *p, const double *primal, const double *dual)
{   // cp_al.c:146-178
    if (!p) return;
    free(p->primal_start);
    free(p->dual_start);
    p->primal_start = nullptr;
    p->dual_start = nullptr;
    if (primal) {
        p->primal_start = (double *)malloc((size_t)std::max(p->num_variables, 1) * sizeof(double));
        memcpy(p->primal_start, primal, (size_t)p->num_variables * sizeof(double));
    }
    if (dual) {
        p->dual_start = (double *)malloc((size_t)std::max(p->num_constraints, 1) * sizeof(double));
        memcpy(p->dual_start, dual, (size_t)p->num_constraints * sizeof(double));
    }
}

PDAL_API void cp_al_result_free(cp_al_result_t *r)
{   // cp_al.c:118-128 (+ reduced_cost, which the reference leaks)
    if (!r) return;
    free(r->primal_solution);
    free(r->dual_solution);
    free(r->reduced_cost);
    free(r);
}

PDAL_API void lp_problem_free(lp_problem_t *p)
{   // cp_al.c:130-144
    if (!p) return;
    free(p->constraint_matrix_row_pointers);
    free(p->constraint_matrix_col_indices);
    free(p->constraint_matrix_values);
    free(p->variable_lower_bound);
    free(p->variable_upper_bound);
    free(p->objective_vector);
    free(p->constraint_lower_bound);
    free(p->constraint_upper_bound);
    free(p->primal_start);
    free(p->dual_start);
    memset(p, 0, sizeof(*p));
    free(p);
}

PDAL_API cp_al_result_t *optimize(const cp_al_parameters_t *params, lp_problem_t *prob)
{   // solver.cu:1175-1736
    if (!params || !prob) return nullptr;
    if (params->termination_evaluation_frequency < 3) {  // check_params_validity, :1757-1766
        fprintf(stderr, "Error: termination_evaluation_frequency must be >= 3 (got %d).\n",
                params->termination_evaluation_frequency);
        return nullptr;
    }
    print_initial_info(params, prob);
    cp_al_result_t *r = nullptr;
    if (params->presolve) {  // solver.cu:1183-1194, 1726-1730; presolve.c:61-226
        if (prob->primal_start || prob->dual_start)
            printf("Warning: Warm-starting is currently not supported when presolve is enabled.\n"
                   "The provided initial solutions will be ignored.\n");
        if (params->verbose) printf("\nRunning presolver (built-in reductions)...\n");
        Presolved ps;
        presolve(prob, ps);
        if (params->verbose) {
            printf("  %-15s : %s\n", "status", presolve_status_name(ps.status));
            printf("  %-15s : %.3g sec\n", "presolve time", ps.seconds);
            printf("  %-15s : %d rows, %d columns, %d nonzeros\n", "reduced problem", (int)ps.rows.size(),
                   (int)ps.cols.size(), ps.reduced ? ps.reduced->constraint_matrix_num_nonzeros : 0);
        }
        const int n = prob->num_variables, m = prob->num_constraints;
        if (!ps.reduced) {  // solved during presolve: infeasible, unbounded, or no column left
            r = (cp_al_result_t *)calloc(1, sizeof(cp_al_result_t));
            r->num_variables = n;
            r->num_constraints = m;
            r->num_nonzeros = prob->constraint_matrix_num_nonzeros;
            r->num_reduced_variables = (int)ps.cols.size();
            r->num_reduced_constraints = (int)ps.rows.size();
            r->num_reduced_nonzeros = 0;
            r->presolve_status = ps.status;
            r->presolve_time = ps.seconds;
            r->primal_solution = (double *)calloc((size_t)std::max(n, 1), sizeof(double));
            r->dual_solution = (double *)calloc((size_t)std::max(m, 1), sizeof(double));
            r->reduced_cost = (double *)calloc((size_t)std::max(n, 1), sizeof(double));
            if (ps.status & (kPresolveInfeasible | kPresolveUnbndOrInfeas)) {
                r->termination_reason = (ps.status & kPresolveInfeasible) ? TERMINATION_REASON_PRIMAL_INFEASIBLE
                                                                          : TERMINATION_REASON_INFEASIBLE_OR_UNBOUNDED;
                r->absolute_primal_residual = r->relative_primal_residual = INFINITY;
                r->absolute_dual_residual = r->relative_dual_residual = INFINITY;
                r->primal_objective_value = INFINITY;
                r->dual_objective_value = -INFINITY;
                r->objective_gap = r->relative_objective_gap = INFINITY;
            } else {
                r->termination_reason = TERMINATION_REASON_OPTIMAL;
                postsolve(ps, prob, nullptr, nullptr, r->primal_solution, r->dual_solution, r->reduced_cost);
                double obj = prob->objective_constant;
                for (int j = 0; j < n; ++j) obj += prob->objective_vector[j] * r->primal_solution[j];
                r->primal_objective_value = r->dual_objective_value = obj;
            }
            if (!quiet()) final_log(r, params);
            return r;
        }
        try {
            r = optimize_gpu(params, ps.reduced);
        } catch (const std::exception &e) {
            fprintf(stderr, "[cp_al-mi355x] solve failed: %s\n", e.what());
            return nullptr;
        }
        if (!r) return nullptr;
        double *x = (double *)malloc((size_t)std::max(n, 1) * sizeof(double));
        double *y = (double *)malloc((size_t)std::max(m, 1) * sizeof(double));
        double *z = (double *)malloc((size_t)std::max(n, 1) * sizeof(double));
        postsolve(ps, prob, r->primal_solution, r->dual_solution, x, y, z);
        free(r->primal_solution);
        free(r->dual_solution);
        free(r->reduced_cost);
        r->primal_solution = x;
        r->dual_solution = y;
        r->reduced_cost = z;
        r->num_reduced_variables = ps.reduced->num_variables;
        r->num_reduced_constraints = ps.reduced->num_constraints;
        r->num_reduced_nonzeros = ps.reduced->constraint_matrix_num_nonzeros;
        r->num_variables = n;
        r->num_constraints = m;
        r->num_nonzeros = prob->constraint_matrix_num_nonzeros;
        r->presolve_status = ps.status;
        r->presolve_time = ps.seconds;
        if (!quiet()) final_log(r, params);
        return r;
    }
    try {
        r = optimize_gpu(params, prob);
    } catch (const std::exception &e) {
        fprintf(stderr, "[cp_al-mi355x] solve failed: %s\n", e.what());
        return nullptr;
    }
    if (r && !quiet()) final_log(r, params);
    return r;
}

PDAL_API cp_al_result_t *solve_lp_problem(lp_problem_t *prob, const cp_al_parameters_t *params)
{   // cp_al.c:180-209
    if (!prob) {
        fprintf(stderr, "[interface] solve_lp_problem: invalid arguments.\n");
        return nullptr;
    }
    cp_al_parameters_t local;
    if (params) local = *params;
    else set_default_parameters(&local);
    cp_al_result_t *r = optimize(&local, prob);
    if (!r) fprintf(stderr, "[interface] optimize returned NULL.\n");
    return r;
}

PDAL_API const char *termination_reason_to_string(termination_reason_t r) { return reason_name(r); }

PDAL_API void pdal_cp_al_reset_rng(void) { reset_rng(); }

PDAL_API int pdal_cp_al_time_kernels(const lp_problem_t *prob, const cp_al_parameters_t *params, int iters,
                                     double *out)
{
    if (!prob || !out || iters <= 0) return -1;
    cp_al_parameters_t local;
    if (params) local = *params;
    else set_default_parameters(&local);
    try {
        return time_kernels_gpu(&local, prob, iters, out);
    } catch (const std::exception &e) {
        fprintf(stderr, "[cp_al-mi355x] kernel timing failed: %s\n", e.what());
        return -2;
    }
}

PDAL_API void *pdal_cp_al_session_create(lp_problem_t *prob, const cp_al_parameters_t *params)
{
    if (!prob) return nullptr;
    cp_al_parameters_t local;
    if (params) local = *params;
    else set_default_parameters(&local);
    if (local.termination_evaluation_frequency < 3) {
        fprintf(stderr, "Error: termination_evaluation_frequency must be >= 3 (got %d).\n",
                local.termination_evaluation_frequency);
        return nullptr;
    }
    try {
        return session_create(&local, prob);
    } catch (const std::exception &e) {
        fprintf(stderr, "[cp_al-mi355x] session set-up failed: %s\n", e.what());
        return nullptr;
    }
}

PDAL_API cp_al_result_t *pdal_cp_al_session_solve(void *session)
{
    if (!session) return nullptr;
    try {
        return session_solve(session);
    } catch (const std::exception &e) {
        fprintf(stderr, "[cp_al-mi355x] session solve failed: %s\n", e.what());
        return nullptr;
    }
}

PDAL_API int pdal_cp_al_session_time_kernels(void *session, int iters, double *out)
{
    if (!session || !out || iters <= 0) return -1;
    try {
        return session_time_kernels(session, iters, out);
    } catch (const std::exception &e) {
        fprintf(stderr, "[cp_al-mi355x] kernel timing failed: %s\n", e.what());
        return -2;
    }
}

PDAL_API double pdal_cp_al_session_setup_seconds(void *session)
{
    return session ? session_setup_seconds(session) : -1.0;
}

PDAL_API void pdal_cp_al_session_free(void *session)
{
    if (session) session_free(session);
}

PDAL_API double pdal_gather_calibration(long long count, long long len, int reps)
{
    try {
        return gather_calibration(count, len, reps);
    } catch (const std::exception &e) {
        fprintf(stderr, "[cp_al-mi355x] gather calibration failed: %s\n", e.what());
        return -2.0;
    }
}

PDAL_API double pdal_copy_calibration(long long bytes, int reps)
{
    try {
        return copy_calibration(bytes, reps);
    } catch (const std::exception &e) {
        fprintf(stderr, "[cp_al-mi355x] copy calibration failed: %s\n", e.what());
        return -2.0;
    }
}

// ---------------------------------------------------------------------------
// Batched solver (pdal_batch.h)

PDAL_API pdal_cp_al_batch_t *pdal_cp_al_batch_create(lp_problem_t *const *problems, int count,
                                                     const cp_al_parameters_t *params)
{
    if (!problems || count <= 0) return nullptr;
    for (int i = 0; i < count; ++i)
        if (!problems[i]) return nullptr;
    cp_al_parameters_t local;
    if (params) local = *params;
    else set_default_parameters(&local);
    if (local.termination_evaluation_frequency < 3) {
        fprintf(stderr, "Error: termination_evaluation_frequency must be >= 3 (got %d).\n",
                local.termination_evaluation_frequency);
        return nullptr;
    }
    try {
        return static_cast<pdal_cp_al_batch_t *>(batch_create(problems, count, &local));
    } catch (const std::exception &e) {
        fprintf(stderr, "[cp_al-mi355x] batch set-up failed: %s\n", e.what());
        return nullptr;
    }
}

PDAL_API int pdal_cp_al_batch_set_starts(pdal_cp_al_batch_t *batch, const double *x, const double *y)
{
    if (!batch) return -1;
    try {
        batch_set_starts(batch, x, y);
        return 0;
    } catch (const std::exception &e) {
        fprintf(stderr, "[cp_al-mi355x] batch start update failed: %s\n", e.what());
        return -2;
    }
}

PDAL_API int pdal_cp_al_batch_start_buffers(pdal_cp_al_batch_t *batch, double **x, double **y)
{
    if (!batch || !x || !y) return -1;
    try {
        batch_start_buffers(batch, x, y);
        return 0;
    } catch (const std::exception &e) {
        fprintf(stderr, "[cp_al-mi355x] batch start buffers failed: %s\n", e.what());
        return -2;
    }
}

PDAL_API int pdal_cp_al_batch_set_starts_from_buffers(pdal_cp_al_batch_t *batch)
{
    if (!batch) return -1;
    try {
        batch_set_starts_from_buffers(batch);
        return 0;
    } catch (const std::exception &e) {
        fprintf(stderr, "[cp_al-mi355x] batch start update failed: %s\n", e.what());
        return -2;
    }
}

PDAL_API int pdal_cp_al_batch_solve(pdal_cp_al_batch_t *batch)
{
    if (!batch) return -1;
    try {
        return batch_solve(batch);
    } catch (const std::exception &e) {
        fprintf(stderr, "[cp_al-mi355x] batch solve failed: %s\n", e.what());
        return -1;
    }
}

PDAL_API cp_al_result_t *pdal_cp_al_batch_result(pdal_cp_al_batch_t *batch, int i)
{
    if (!batch) return nullptr;
    try {
        return batch_result(batch, i);
    } catch (const std::exception &e) {
        fprintf(stderr, "[cp_al-mi355x] batch result failed: %s\n", e.what());
        return nullptr;
    }
}

PDAL_API void pdal_cp_al_batch_stats(const pdal_cp_al_batch_t *batch, double *out)
{
    if (!batch || !out) return;
    batch_stats(const_cast<pdal_cp_al_batch_t *>(batch), out);
}

PDAL_API int pdal_cp_al_batch_time_kernels(pdal_cp_al_batch_t *batch, int iters, double *out)
{
    if (!batch || !out || iters < 2) return -1;
    try {
        batch_time_kernels(batch, iters, out);
        return 0;
    } catch (const std::exception &e) {
        fprintf(stderr, "[cp_al-mi355x] batch kernel timing failed: %s\n", e.what());
        return -2;
    }
}

PDAL_API int pdal_cp_al_batch_primal_weights(const pdal_cp_al_batch_t *batch, double *out)
{
    if (!batch || !out) return -1;
    batch_primal_weights(const_cast<pdal_cp_al_batch_t *>(batch), out);
    return 0;
}

PDAL_API int pdal_cp_al_batch_set_primal_weights(pdal_cp_al_batch_t *batch, const double *pw)
{
    if (!batch || !pw) return -1;
    batch_set_primal_weights(batch, pw);
    return 0;
}

static int batch_vectors(pdal_cp_al_batch_t *batch, const double *c, const double *con_lb, const double *con_ub,
                         const double *var_lb, const double *var_ub, bool device)
{
    if (!batch) return -1;
    try {
        batch_set_vectors(batch, c, con_lb, con_ub, var_lb, var_ub, device);
        return 0;
    } catch (const std::exception &e) {
        fprintf(stderr, "[cp_al-mi355x] batch vector update failed: %s\n", e.what());
        return -2;
    }
}

PDAL_API int pdal_cp_al_batch_set_vectors(pdal_cp_al_batch_t *batch, const double *c, const double *con_lb,
                                          const double *con_ub, const double *var_lb, const double *var_ub)
{
    return batch_vectors(batch, c, con_lb, con_ub, var_lb, var_ub, false);
}

PDAL_API int pdal_cp_al_batch_set_vectors_device(pdal_cp_al_batch_t *batch, const double *c, const double *con_lb,
                                                 const double *con_ub, const double *var_lb, const double *var_ub)
{
    return batch_vectors(batch, c, con_lb, con_ub, var_lb, var_ub, true);
}

PDAL_API void pdal_cp_al_batch_free(pdal_cp_al_batch_t *batch)
{
    if (batch) batch_free(batch);
}
