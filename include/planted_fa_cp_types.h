/*
 * Public types of RHR-FA-CP, the planted-KKT FA_CP QP solver (MI355X build,
 * librhr_fa_cp.so).
 *
 * Layout-compatible with the reference's QP/PLANTED_FA_CP/include/
 * fa_cp_types.h: field order and types are kept so a caller compiled against
 * that header links against librhr_fa_cp.so unchanged.  The problem is
 *   min 1/2 x^T (Q + R^T R) x + c^T x + c0  s.t.  A x = b, lb <= x <= ub
 * in the AL branch (b is the constraint lower bound); the plain branch keeps
 * FA_CP's l <= A x <= u.  The type names are FA_CP's with another parameter
 * and result layout: include one of the two headers per translation unit.
 */
#pragma once

#include <stdbool.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

/* reference: QP/PLANTED_FA_CP/include/fa_cp_types.h (as QP/FA_CP's):28-37 */
typedef enum {
    TERMINATION_REASON_UNSPECIFIED,
    TERMINATION_REASON_OPTIMAL,
    TERMINATION_REASON_PRIMAL_INFEASIBLE,
    TERMINATION_REASON_DUAL_INFEASIBLE,
    TERMINATION_REASON_INFEASIBLE_OR_UNBOUNDED,
    TERMINATION_REASON_TIME_LIMIT,
    TERMINATION_REASON_ITERATION_LIMIT,
    TERMINATION_REASON_FEAS_POLISH_SUCCESS
} termination_reason_t;

/* reference: QP/PLANTED_FA_CP/include/fa_cp_types.h (as QP/FA_CP's):39 */
typedef enum { NORM_TYPE_L2 = 0, NORM_TYPE_L_INF = 1 } norm_type_t;

/* reference: QP/PLANTED_FA_CP/include/fa_cp_types.h (as QP/FA_CP's):41-45 */
typedef struct {
    int *row_ptr;
    int *col_ind;
    double *val;
} CsrComponent;

/* reference: QP/PLANTED_FA_CP/include/fa_cp_types.h (as QP/FA_CP's):47-53 */
typedef enum {
    FA_CP_SPARSE_Q,
    FA_CP_DIAG_Q,
    FA_CP_LOW_RANK_PLUS_SPARSE_Q,
    FA_CP_LOW_RANK_Q,
    FA_CP_NON_Q
} quad_obj_type_t;

/* reference: QP/PLANTED_FA_CP/include/fa_cp_types.h (as QP/FA_CP's):55-79 */
typedef struct {
    int num_variables;
    int num_constraints;
    int num_rank_lowrank_obj;
    double *variable_lower_bound;
    double *variable_upper_bound;
    double *objective_vector;
    double objective_constant;

    CsrComponent *constraint_matrix;
    int constraint_matrix_num_nonzeros;

    CsrComponent *objective_sparse_matrix;
    int objective_sparse_matrix_num_nonzeros;

    CsrComponent *objective_lowrank_matrix;
    int objective_lowrank_matrix_num_nonzeros;

    double *constraint_lower_bound;
    double *constraint_upper_bound;

    double *primal_start;
    double *dual_start;
} qp_problem_t;

/* reference: QP/PLANTED_FA_CP/include/fa_cp_types.h (as QP/FA_CP's):81-89 */
typedef struct {
    double artificial_restart_threshold;
    double sufficient_reduction_for_restart;
    double necessary_reduction_for_restart;
    double k_p;
    double k_i;
    double k_d;
    double i_smooth;
} restart_parameters_t;

/* reference: QP/PLANTED_FA_CP/include/fa_cp_types.h (as QP/FA_CP's):91-98 */
typedef struct {
    double eps_optimal_relative;
    double eps_feasible_relative;
    double eps_feas_polish_relative;
    double eps_infeasible;
    double time_sec_limit;
    int iteration_limit;
} termination_criteria_t;

/* reference: QP/PLANTED_FA_CP/include/fa_cp_types.h:100-110 */
typedef enum {
    ADAPTIVE_AL_SIGMA_MODE_BIDIRECTIONAL = 0,
    ADAPTIVE_AL_SIGMA_MODE_UP_ONLY = 1,
    ADAPTIVE_AL_SIGMA_MODE_DOWN_ONLY = 2
} adaptive_al_sigma_mode_t;

typedef enum {
    ADAPTIVE_AL_SIGMA_RULE_TRIGGER = 0,
    ADAPTIVE_AL_SIGMA_RULE_SQRT = 1,
    ADAPTIVE_AL_SIGMA_RULE_BANGBANG = 2
} adaptive_al_sigma_rule_t;

/* reference: QP/PLANTED_FA_CP/include/fa_cp_types.h:112-138 */
typedef struct {
    int l_inf_ruiz_iterations;
    bool has_pock_chambolle_alpha;
    double pock_chambolle_alpha;
    bool bound_objective_rescaling;
    bool use_al_qp;
    double al_sigma;
    bool adaptive_al_sigma;
    adaptive_al_sigma_mode_t adaptive_al_sigma_mode;
    adaptive_al_sigma_rule_t adaptive_al_sigma_rule;
    double al_sigma_min;
    double al_sigma_max;
    double al_sigma_increase_factor;
    double al_sigma_decrease_factor;
    double al_sigma_primal_trigger;
    double al_sigma_gap_trigger;
    const char *trajectory_csv;
    int verbose;
    int termination_evaluation_frequency;
    int sv_max_iter;
    double sv_tol;
    termination_criteria_t termination_criteria;
    restart_parameters_t restart_params;
    double reflection_coefficient;
    bool feasibility_polishing;
    norm_type_t optimality_norm;
} fa_cp_parameters_t;

/* reference: QP/PLANTED_FA_CP/include/fa_cp_types.h:140-172 */
typedef struct {
    int num_variables;
    int num_constraints;
    int num_nonzeros;

    int num_reduced_variables;
    int num_reduced_constraints;
    int num_reduced_nonzeros;

    double *primal_solution;
    double *dual_solution;
    double *reduced_cost;

    int total_count;
    int total_inner_count;
    double rescaling_time_sec;
    double cumulative_time_sec;

    double absolute_primal_residual;
    double relative_primal_residual;
    double absolute_dual_residual;
    double relative_dual_residual;
    double primal_objective_value;
    double dual_objective_value;
    double objective_gap;
    double relative_objective_gap;
    double max_primal_ray_infeasibility;
    double max_dual_ray_infeasibility;
    double primal_ray_linear_objective;
    double dual_ray_objective;
    termination_reason_t termination_reason;
    double feasibility_polishing_time;
    int feasibility_iteration;
} fa_cp_result_t;

/* reference: QP/PLANTED_FA_CP/include/fa_cp_types.h (as QP/FA_CP's):179-185 */
typedef enum {
    matrix_dense = 0,
    matrix_csr = 1,
    matrix_csc = 2,
    matrix_coo = 3
} matrix_format_t;

/* reference: QP/PLANTED_FA_CP/include/fa_cp_types.h (as QP/FA_CP's):187-222.  |x| <= zero_tolerance entries are
 * dropped from dense (default 1e-12), CSC and COO input (when > 0). */
typedef struct {
    int m;
    int n;
    matrix_format_t fmt;

    double zero_tolerance;

    union MatrixData {
        struct MatrixDense {
            const double *A;
        } dense;

        struct MatrixCSR {
            int nnz;
            const int *row_ptr;
            const int *col_ind;
            const double *vals;
        } csr;

        struct MatrixCSC {
            int nnz;
            const int *col_ptr;
            const int *row_ind;
            const double *vals;
        } csc;

        struct MatrixCOO {
            int nnz;
            const int *row_ind;
            const int *col_ind;
            const double *vals;
        } coo;
    } data;
} matrix_desc_t;

#ifdef __cplusplus
}
#endif
