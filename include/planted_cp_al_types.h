/*
 * Public types of the planted-KKT CP-AL QP solver, RHR-CP-AL (MI355X build,
 * librhr_cp_al.so).
 *
 * Layout-compatible with the reference's QP/PLANTED_CP_AL/include/cp_al_types.h:
 * field order and types are kept so a caller compiled against that header
 * links against librhr_cp_al.so unchanged.  The types are those of the
 * linearized solver (lin_cp_al_types.h) with one more parameter,
 * trajectory_csv (the per-checkpoint trajectory file).  The
 * problem is
 *   min 1/2 x^T (Q + R^T R) x + c^T x + c0  s.t.  A x = b, lb <= x <= ub
 * (b is the constraint lower bound; the upper bound is not read).  The types
 * share their names with the LP solver's cp_al_types.h and the other QP
 * headers: include one of them per translation unit.
 */
#pragma once

#include <stdbool.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

/* reference: QP/PLANTED_CP_AL/include/cp_al_types.h:28-37 */
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

/* reference: QP/PLANTED_CP_AL/include/cp_al_types.h:39 */
typedef enum { NORM_TYPE_L2 = 0, NORM_TYPE_L_INF = 1 } norm_type_t;

/* reference: QP/PLANTED_CP_AL/include/cp_al_types.h:41-45 */
typedef struct {
    int *row_ptr;
    int *col_ind;
    double *val;
} CsrComponent;

/* reference: QP/PLANTED_CP_AL/include/cp_al_types.h:47-53 */
typedef enum {
    CP_AL_SPARSE_Q,
    CP_AL_DIAG_Q,
    CP_AL_LOW_RANK_PLUS_SPARSE_Q,
    CP_AL_LOW_RANK_Q,
    CP_AL_NON_Q
} quad_obj_type_t;

/* reference: QP/PLANTED_CP_AL/include/cp_al_types.h:55-79 */
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

/* reference: QP/PLANTED_CP_AL/include/cp_al_types.h:81-89 */
typedef struct {
    double artificial_restart_threshold;
    double sufficient_reduction_for_restart;
    double necessary_reduction_for_restart;
    double k_p;
    double k_i;
    double k_d;
    double i_smooth;
} restart_parameters_t;

/* reference: QP/PLANTED_CP_AL/include/cp_al_types.h:91-98 */
typedef struct {
    double eps_optimal_relative;
    double eps_feasible_relative;
    double eps_feas_polish_relative;
    double eps_infeasible;
    double time_sec_limit;
    int iteration_limit;
} termination_criteria_t;

/* reference: QP/PLANTED_CP_AL/include/cp_al_types.h:100-117 */
typedef struct {
    int l_inf_ruiz_iterations;
    bool has_pock_chambolle_alpha;
    double pock_chambolle_alpha;
    bool bound_objective_rescaling;
    bool use_al_qp;
    double sigma_init;
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
} cp_al_parameters_t;

/* reference: QP/PLANTED_CP_AL/include/cp_al_types.h:119-152 */
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
} cp_al_result_t;

/* reference: QP/PLANTED_CP_AL/include/cp_al_types.h:155-160 */
typedef enum {
    matrix_dense = 0,
    matrix_csr = 1,
    matrix_csc = 2,
    matrix_coo = 3
} matrix_format_t;

/* reference: QP/PLANTED_CP_AL/include/cp_al_types.h:163-197.  |x| <= zero_tolerance entries are
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
