// cp_al: the LP runner command line, linked against libcp_al.so.
//
// Reference: LP/CP_AL/src/cli.c:835-1092 (usage, getopt flags, outputs),
// :91-214 (anonymous matrix-structure features), :216-387 (feature auto
// policy), :408-469 (unified policy), :589-832 (two-phase structural tail),
// :471-586 (output files).  The policy layer configures the solver through the
// same CP_AL_* environment variables the library reads (the tail's positive-
// alpha pulse, the restart-candidate selection of the feature policy).
#include "cp_al.h"
#include "mps_parser.h"
#include "pdal_lp.h"

#include <getopt.h>
#include <libgen.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

namespace {

bool env_flag(const char *name)
{
    const char *v = getenv(name);
    return v && v[0] != '\0' && atoi(v) != 0;
}

void policy_set(const char *name, const char *value) { setenv(name, value, 1); }
void policy_clear(const char *name) { unsetenv(name); }

struct Features {  // cli.c:52-66
    double coeff_log_range = 0, row_logmax_std = 0, col_logmax_std = 0, row_logspan_mean = 0,
           col_logspan_mean = 0, row_span2_fraction = 0, col_span2_fraction = 0, row_nnz_cv = 0, col_nnz_cv = 0,
           equality_fraction = 0, two_sided_variable_fraction = 0, objective_density = 0;
};

double coef_var(double sum, double sum_sq, int count)
{
    if (count <= 0 || sum <= 0.0) return 0.0;
    const double mean = sum / count;
    return std::sqrt(std::fmax(sum_sq / count - mean * mean, 0.0)) / mean;
}

double stdev(double sum, double sum_sq, int count)
{
    if (count <= 0) return 0.0;
    const double mean = sum / count;
    return std::sqrt(std::fmax(sum_sq / count - mean * mean, 0.0));
}

Features matrix_features(const lp_problem_t *p)
{   // compute_feature_matrix_stats, cli.c:91-214
    Features f;
    const int m = p->num_constraints, n = p->num_variables, nnz = p->constraint_matrix_num_nonzeros;
    if (m <= 0 || n <= 0 || nnz <= 0) return f;
    std::vector<double> col_min(n, DBL_MAX), col_max(n, 0.0);
    std::vector<int> col_nnz(n, 0);
    double gmin = DBL_MAX, gmax = 0.0;
    double rlm = 0, rlm2 = 0, rls = 0, rn = 0, rn2 = 0;
    int nonempty_rows = 0, row_span2 = 0;
    for (int r = 0; r < m; ++r) {
        const int b = p->constraint_matrix_row_pointers[r], e = p->constraint_matrix_row_pointers[r + 1];
        rn += e - b;
        rn2 += (double)(e - b) * (e - b);
        double rmin = DBL_MAX, rmax = 0.0;
        for (int k = b; k < e; ++k) {
            const double v = std::fabs(p->constraint_matrix_values[k]);
            const int c = p->constraint_matrix_col_indices[k];
            if (!std::isfinite(v) || v <= 0.0 || c < 0 || c >= n) continue;
            gmin = std::fmin(gmin, v);
            gmax = std::fmax(gmax, v);
            rmin = std::fmin(rmin, v);
            rmax = std::fmax(rmax, v);
            col_min[c] = std::fmin(col_min[c], v);
            col_max[c] = std::fmax(col_max[c], v);
            col_nnz[c] += 1;
        }
        if (rmax > 0.0 && rmin < DBL_MAX) {
            const double lmax = std::log10(rmax), lspan = std::log10(rmax / rmin);
            rlm += lmax;
            rlm2 += lmax * lmax;
            rls += lspan;
            row_span2 += lspan >= 2.0 ? 1 : 0;
            nonempty_rows += 1;
        }
    }
    double clm = 0, clm2 = 0, cls = 0, cn = 0, cn2 = 0;
    int nonempty_cols = 0, col_span2 = 0;
    for (int c = 0; c < n; ++c) {
        cn += col_nnz[c];
        cn2 += (double)col_nnz[c] * col_nnz[c];
        if (col_max[c] > 0.0 && col_min[c] < DBL_MAX) {
            const double lmax = std::log10(col_max[c]), lspan = std::log10(col_max[c] / col_min[c]);
            clm += lmax;
            clm2 += lmax * lmax;
            cls += lspan;
            col_span2 += lspan >= 2.0 ? 1 : 0;
            nonempty_cols += 1;
        }
    }
    f.coeff_log_range = gmax > 0.0 && gmin < DBL_MAX ? std::log10(gmax / gmin) : 0.0;
    f.row_logmax_std = stdev(rlm, rlm2, nonempty_rows);
    f.col_logmax_std = stdev(clm, clm2, nonempty_cols);
    f.row_logspan_mean = nonempty_rows > 0 ? rls / nonempty_rows : 0.0;
    f.col_logspan_mean = nonempty_cols > 0 ? cls / nonempty_cols : 0.0;
    f.row_span2_fraction = nonempty_rows > 0 ? (double)row_span2 / nonempty_rows : 0.0;
    f.col_span2_fraction = nonempty_cols > 0 ? (double)col_span2 / nonempty_cols : 0.0;
    f.row_nnz_cv = coef_var(rn, rn2, m);
    f.col_nnz_cv = coef_var(cn, cn2, n);
    int eq = 0, two = 0, dense_obj = 0;
    for (int r = 0; r < m; ++r) {
        const double lo = p->constraint_lower_bound[r], hi = p->constraint_upper_bound[r];
        eq += std::isfinite(lo) && std::isfinite(hi) && lo == hi ? 1 : 0;
    }
    for (int c = 0; c < n; ++c) {
        two += std::isfinite(p->variable_lower_bound[c]) && std::isfinite(p->variable_upper_bound[c]) ? 1 : 0;
        dense_obj += std::fabs(p->objective_vector[c]) > 0.0 ? 1 : 0;
    }
    f.equality_fraction = (double)eq / m;
    f.two_sided_variable_fraction = (double)two / n;
    f.objective_density = (double)dense_obj / n;
    return f;
}

void apply_feature_auto_policy(const lp_problem_t *problem, cp_al_parameters_t *params, const Features *stats)
{   // cli.c:216-387
    if (!env_flag("CP_AL_FEATURE_AUTO_POLICY")) return;
    if (!stats) {
        fprintf(stderr, "[feature-auto-policy-error] reason=missing-precomputed-stats\n");
        return;
    }
    const char *variant = getenv("CP_AL_FEATURE_POLICY_VARIANT");
    if (!variant || !variant[0]) variant = "safe0";
    const bool sigma_probe = strcmp(variant, "probe") == 0;
    const bool sigma_always = strcmp(variant, "always_select") == 0 || strcmp(variant, "always_noselect") == 0;
    const bool sigma_trigger =
        strcmp(variant, "trigger") == 0 || strcmp(variant, "trigger_select") == 0 || sigma_probe;
    const bool restart_select = strcmp(variant, "always_select") == 0 || strcmp(variant, "trigger_select") == 0;

    int eval_freq = 200;
    const char *uef = getenv("CP_AL_UNIFIED_EVAL_FREQ");
    if (uef && uef[0] && atoi(uef) > 0) eval_freq = atoi(uef);
    params->termination_evaluation_frequency = eval_freq;
    params->reflection_coefficient = 1.0;
    policy_set("CP_AL_LAMBDA", "0");
    policy_set("CP_AL_LAMBDA_FACTOR", "2.01");
    policy_set("CP_AL_ADAPTIVE_LAMBDA", "0");
    policy_clear("CP_AL_LAMBDA_PROFILE");
    policy_clear("CP_AL_LAMBDA_GRID");

    const Features &s = *stats;
    const char *cr_override_raw = getenv("CP_AL_FEATURE_CR_ITERS");
    int cr_override = -1, cr = 0;
    if (cr_override_raw && cr_override_raw[0]) {
        cr_override = atoi(cr_override_raw) < 0 ? 0 : atoi(cr_override_raw);
        cr = cr_override;
    } else if (s.col_nnz_cv <= 0.19012745) {
        cr = 5;
    } else if (s.col_nnz_cv <= 0.39521255) {
        cr = 50;
    } else if (s.coeff_log_range <= 3.222854) {
        cr = 5;
    } else {
        cr = 20;
    }
    char buf[32];
    snprintf(buf, sizeof buf, "%d", cr);
    policy_set("CP_AL_CR_ITERS", buf);
    policy_set("CP_AL_CR_SCALING", cr > 0 ? "1" : "0");

    auto env_or = [](const char *k, const char *d) {
        const char *v = getenv(k);
        return (v && v[0]) ? v : d;
    };
    const char *sif = env_or("CP_AL_FEATURE_SIGMA_INIT_FACTOR", "6.0");
    const char *stm = env_or("CP_AL_FEATURE_SIGMA_TRIGGER_MIN_ITERS", "200");
    const char *stp = env_or("CP_AL_FEATURE_SIGMA_TRIGGER_PRIMAL", "1e-2");
    const char *std_ = env_or("CP_AL_FEATURE_SIGMA_TRIGGER_DUAL", "1e-6");
    const char *str = env_or("CP_AL_FEATURE_SIGMA_TRIGGER_RATIO", "1e6");
    policy_set("CP_AL_SIGMA_MODE", sigma_always ? "1" : "0");
    policy_set("CP_AL_SIGMA_ADAPTIVE", "1");
    policy_set("CP_AL_SIGMA_INIT_FACTOR", sif);
    policy_set("CP_AL_SIGMA_TRIGGER", sigma_trigger ? "1" : "0");
    policy_set("CP_AL_SIGMA_TRIGGER_MIN_ITERS", stm);
    policy_set("CP_AL_SIGMA_TRIGGER_PRIMAL", stp);
    policy_set("CP_AL_SIGMA_TRIGGER_DUAL", std_);
    policy_set("CP_AL_SIGMA_TRIGGER_RATIO", str);
    const char *def = env_or("CP_AL_FEATURE_DYNAMIC_EVAL_FAST", "50");
    policy_set("CP_AL_DYNAMIC_EVAL", sigma_trigger && !sigma_probe ? "1" : "0");
    policy_set("CP_AL_DYNAMIC_EVAL_FAST", def);
    policy_set("CP_AL_DYNAMIC_EVAL_MIN_ITERS", "200");
    policy_set("CP_AL_DYNAMIC_EVAL_SMALL_RES", "1e-6");
    policy_set("CP_AL_DYNAMIC_EVAL_LARGE_RES", "1e-4");
    policy_set("CP_AL_DYNAMIC_EVAL_RATIO", "1e4");
    policy_set("CP_AL_RESTART_SELECT", restart_select ? "3" : "0");
    policy_clear("CP_AL_RESTART_SELECT_PROFILE");
    policy_set("CP_AL_RESTART_SELECT_MARGIN", "0.9");
    policy_set("CP_AL_RESTART_SUFFICIENT", "0.2");
    policy_set("CP_AL_RESTART_NECESSARY", "0.8");
    policy_set("CP_AL_RESTART_ARTIFICIAL", "0.36");
    policy_clear("CP_AL_STEP_MARGIN");

    const double m = problem->num_constraints > 0 ? problem->num_constraints : 1;
    const double n = problem->num_variables > 0 ? problem->num_variables : 1;
    const double nnz = problem->constraint_matrix_num_nonzeros > 0 ? problem->constraint_matrix_num_nonzeros : 1;
    fprintf(stderr,
            "[feature-auto-policy] version=2 variant=%s eval_base=200 refl=1 rows=%d cols=%d nnz=%d "
            "n_over_m=%.6e nnz_per_col=%.6e sigma_pid=1 sigma_always=%d sigma_trigger=%d "
            "sigma_probe=%d sigma_init_factor=%s sigma_trigger_min_iters=%s "
            "restart_select=%d cr_override=%d cr_selected=%d coeff_log_range=%.6e row_logmax_std=%.6e "
            "col_logmax_std=%.6e row_logspan_mean=%.6e col_logspan_mean=%.6e row_span2_frac=%.6e "
            "col_span2_frac=%.6e row_nnz_cv=%.6e col_nnz_cv=%.6e equality_frac=%.6e two_sided_var_frac=%.6e "
            "objective_density=%.6e\n",
            variant, problem->num_constraints, problem->num_variables, problem->constraint_matrix_num_nonzeros,
            n / m, nnz / n, sigma_always ? 1 : 0, sigma_trigger ? 1 : 0, sigma_probe ? 1 : 0, sif, stm,
            restart_select ? 3 : 0, cr_override, cr, s.coeff_log_range, s.row_logmax_std, s.col_logmax_std,
            s.row_logspan_mean, s.col_logspan_mean, s.row_span2_fraction, s.col_span2_fraction, s.row_nnz_cv,
            s.col_nnz_cv, s.equality_fraction, s.two_sided_variable_fraction, s.objective_density);
}

void apply_unified_policy(const Features *stats)
{   // cli.c:408-469
    if (!env_flag("CP_AL_UNIFIED_POLICY")) return;
    if (!stats) {
        fprintf(stderr, "[unified-policy-error] reason=missing-precomputed-stats\n");
        return;
    }
    const Features &s = *stats;
    policy_set("CP_AL_FEATURE_AUTO_POLICY", "1");
    policy_set("CP_AL_FEATURE_POLICY_VARIANT", "safe0");
    const char *route;
    if (s.objective_density <= 0.01 && s.col_nnz_cv >= 5.0) {
        policy_set("CP_AL_FEATURE_CR_ITERS", "5");
        policy_set("CP_AL_UNIFIED_EVAL_FREQ", "100");
        policy_clear("CP_AL_STRUCTURAL_TAIL_CUT");
        policy_clear("CP_AL_TAIL_ONLY_PULSE");
        route = "lowobj_highcv_eval100_cr5";
    } else if (s.equality_fraction == 0.0) {
        policy_set("CP_AL_FEATURE_CR_ITERS", "0");
        policy_clear("CP_AL_STRUCTURAL_TAIL_CUT");
        policy_clear("CP_AL_TAIL_ONLY_PULSE");
        route = "zero_equality_cr0";
    } else {
        policy_clear("CP_AL_FEATURE_CR_ITERS");
        policy_set("CP_AL_STRUCTURAL_TAIL_CUT", "1");
        policy_set("CP_AL_TAIL_ONLY_PULSE", "1");
        route = "high_equality_tail_or_base";
    }
    fprintf(stderr,
            "[unified-policy] version=2 route=%s equality_frac=%.6e row_nnz_cv=%.6e col_nnz_cv=%.6e "
            "objective_density=%.6e row_logmax_std=%.6e\n",
            route, s.equality_fraction, s.row_nnz_cv, s.col_nnz_cv, s.objective_density, s.row_logmax_std);
}

std::string out_path(const char *dir, const std::string &name, const char *suffix)
{
    return std::string(dir) + "/" + name + suffix;
}

std::string instance_name(const char *filename)
{   // extract_instance_name: basename up to the first dot
    std::string s(filename);
    const size_t slash = s.find_last_of('/');
    if (slash != std::string::npos) s = s.substr(slash + 1);
    const size_t dot = s.find('.');
    if (dot != std::string::npos) s = s.substr(0, dot);
    return s;
}

void save_solution(const double *v, int n, const char *dir, const std::string &name, const char *suffix)
{   // cli.c:471-499, "%.10g" per line
    if (!v) return;
    const std::string path = out_path(dir, name, suffix);
    FILE *f = fopen(path.c_str(), "w");
    if (!f) {
        perror("Error opening solution file");
        return;
    }
    for (int i = 0; i < n; ++i) fprintf(f, "%.10g\n", v[i]);
    fclose(f);
}

const char *presolve_status_name(int status)
{   // get_presolve_status_str, presolve.c:16-31 (the library's status flags)
    switch (status) {
    case 0: return "UNCHANGED";
    case 1: return "REDUCED";
    case 2: return "INFEASIBLE";
    case 4: return "INFEASIBLE_OR_UNBOUNDED";
    default: return "UNKNOWN_STATUS";
    }
}

void save_summary(const cp_al_result_t *r, const char *dir, const std::string &name)
{   // save_solver_summary, cli.c:501-586 (same keys, parsed by the benchmark collectors)
    const std::string path = out_path(dir, name, "_summary.txt");
    FILE *f = fopen(path.c_str(), "w");
    if (!f) {
        perror("Error opening summary file");
        return;
    }
    fprintf(f, "Termination Reason: %s\n", termination_reason_to_string(r->termination_reason));
    if (r->presolve_time > 0.0) {  // cli.c:542-549
        fprintf(f, "Presolve Status: %s\n", presolve_status_name(r->presolve_status));
        fprintf(f, "Presolve Time (sec): %e\n", r->presolve_time);
        fprintf(f, "Reduced Rows: %d\n", r->num_reduced_constraints);
        fprintf(f, "Reduced Columns: %d\n", r->num_reduced_variables);
        fprintf(f, "Reduced Nonzeros: %d\n", r->num_reduced_nonzeros);
    }
    fprintf(f, "Precondition time (sec): %e\n", r->rescaling_time_sec);
    fprintf(f, "Runtime (sec): %e\n", r->cumulative_time_sec);
    fprintf(f, "Iterations Count: %d\n", r->total_count);
    fprintf(f, "Primal Objective Value: %e\n", r->primal_objective_value);
    fprintf(f, "Dual Objective Value: %e\n", r->dual_objective_value);
    fprintf(f, "Relative Primal Residual: %e\n", r->relative_primal_residual);
    fprintf(f, "Relative Dual Residual: %e\n", r->relative_dual_residual);
    fprintf(f, "Absolute Objective Gap: %e\n", r->objective_gap);
    fprintf(f, "Relative Objective Gap: %e\n", r->relative_objective_gap);
    fprintf(f, "Rows: %d\n", r->num_constraints);
    fprintf(f, "Columns: %d\n", r->num_variables);
    fprintf(f, "Nonzeros: %d\n", r->num_nonzeros);
    if (r->feasibility_polishing_time > 0.0) {
        fprintf(f, "Feasibility Polishing Time (sec): %e\n", r->feasibility_polishing_time);
        fprintf(f, "Feasibility Polishing Iteration Count: %d\n", r->feasibility_iteration);
    }
    fclose(f);
}

cp_al_result_t *structural_tail(lp_problem_t *problem, const cp_al_parameters_t *req, const Features *stats,
                                bool pulse_requested)
{   // run_structural_tail_cut, cli.c:607-832
    if (!stats) {
        fprintf(stderr, "[structural-tail-error] reason=missing-precomputed-stats\n");
        return solve_lp_problem(problem, req);
    }
    const Features &s = *stats;
    const bool homogeneous = s.objective_density >= 0.99 && s.row_logmax_std <= 0.05;
    const bool low_eq = homogeneous && s.equality_fraction <= 0.001;
    const bool high_eq = homogeneous && s.equality_fraction >= 0.99;
    const bool early_cut = low_eq || high_eq;
    const bool pulse = pulse_requested && early_cut;
    const double base_fraction = high_eq ? 0.05 : 0.75;
    const double limit = req->termination_criteria.time_sec_limit;
    if (!std::isfinite(limit) || limit <= 0.0) {
        fprintf(stderr, "[structural-tail-error] reason=invalid-time-limit value=%.12e\n", limit);
        return solve_lp_problem(problem, req);
    }
    if (!early_cut) {
        policy_set("CP_AL_ALPHA_DIRECT_PULSE", "0");
        policy_set("CP_AL_DETERMINISTIC_PULSE_FORCE", "0");
        fprintf(stderr,
                "[structural-tail-config] enabled=1 structural_early_cut=0 structural_low_equality=0 "
                "structural_high_equality=0 base_fraction=1.00 rescue_fraction=0.00 rescue_profile=none "
                "base_alpha=0 tail_pulse_requested=%d tail_pulse_enabled=0 tail_pulse_alpha=2.000000000000e-03 "
                "equality_fraction=%.12e objective_density=%.12e row_logmax_std=%.12e original_limit=%.12e "
                "base_limit=%.12e\n",
                pulse_requested ? 1 : 0, s.equality_fraction, s.objective_density, s.row_logmax_std, limit, limit);
        cp_al_result_t *r = solve_lp_problem(problem, req);
        if (!r) {
            fprintf(stderr, "[structural-tail-error] reason=null-single-phase-result\n");
            return nullptr;
        }
        fprintf(stderr, "[structural-tail-base] status=%s solve_sec=%.12e precondition_sec=%.12e iterations=%d\n",
                termination_reason_to_string(r->termination_reason), r->cumulative_time_sec, r->rescaling_time_sec,
                r->total_count);
        fprintf(stderr,
                "[structural-tail-final] phases=1 tail_started=0 status=%s structural_early_cut=0 "
                "base_fraction=1.00 rescue_profile=none base_solve_sec=%.12e rescue_solve_sec=0 "
                "aggregate_solve_sec=%.12e budget_excess=%.12e total_iterations=%d\n",
                termination_reason_to_string(r->termination_reason), r->cumulative_time_sec, r->cumulative_time_sec,
                std::fmax(r->cumulative_time_sec - limit, 0.0), r->total_count);
        return r;
    }
    cp_al_parameters_t base = *req;
    base.termination_criteria.time_sec_limit = base_fraction * limit;
    fprintf(stderr,
            "[structural-tail-config] enabled=1 structural_early_cut=1 structural_low_equality=%d "
            "structural_high_equality=%d base_fraction=%.2f rescue_fraction=%.2f rescue_profile=cr0 base_alpha=0 "
            "tail_pulse_requested=%d tail_pulse_enabled=%d tail_pulse_alpha=2.000000000000e-03 "
            "equality_fraction=%.12e objective_density=%.12e row_logmax_std=%.12e original_limit=%.12e "
            "base_limit=%.12e\n",
            low_eq ? 1 : 0, high_eq ? 1 : 0, base_fraction, 1.0 - base_fraction, pulse_requested ? 1 : 0,
            pulse ? 1 : 0, s.equality_fraction, s.objective_density, s.row_logmax_std, limit,
            base.termination_criteria.time_sec_limit);
    policy_set("CP_AL_ALPHA_DIRECT_PULSE", "0");
    policy_set("CP_AL_DETERMINISTIC_PULSE_FORCE", "0");
    cp_al_result_t *br = solve_lp_problem(problem, &base);
    if (!br) {
        fprintf(stderr, "[structural-tail-error] reason=null-base-result\n");
        return nullptr;
    }
    const double bt = br->cumulative_time_sec, bp = br->rescaling_time_sec, bpre = br->presolve_time,
                 bfp = br->feasibility_polishing_time;
    const int bi = br->total_count;
    fprintf(stderr, "[structural-tail-base] status=%s solve_sec=%.12e precondition_sec=%.12e iterations=%d\n",
            termination_reason_to_string(br->termination_reason), bt, bp, bi);
    if (br->termination_reason == TERMINATION_REASON_OPTIMAL) {
        fprintf(stderr,
                "[structural-tail-final] phases=1 tail_started=0 status=OPTIMAL structural_early_cut=1 "
                "base_fraction=%.2f rescue_profile=cr0 base_solve_sec=%.12e rescue_solve_sec=0 "
                "aggregate_solve_sec=%.12e budget_excess=%.12e total_iterations=%d\n",
                base_fraction, bt, bt, std::fmax(bt - limit, 0.0), bi);
        return br;
    }
    const double remaining = std::fmax(limit - bt, 0.0);
    if (remaining <= 0.0) {
        fprintf(stderr,
                "[structural-tail-final] phases=1 tail_started=0 status=%s structural_early_cut=1 "
                "base_fraction=%.2f rescue_profile=cr0 base_solve_sec=%.12e rescue_solve_sec=0 "
                "aggregate_solve_sec=%.12e budget_excess=%.12e total_iterations=%d reason=no-remaining-budget\n",
                termination_reason_to_string(br->termination_reason), base_fraction, bt, bt,
                std::fmax(bt - limit, 0.0), bi);
        return br;
    }
    // warm anchor = phase-1 result (structural_tail_set_warm_anchor, :589-605)
    set_start_values(problem, br->primal_solution, br->dual_solution);
    double sq = 0.0;
    for (int j = 0; j < problem->num_variables; ++j) {
        const double d = problem->primal_start[j] - br->primal_solution[j];
        sq += d * d;
    }
    for (int i = 0; i < problem->num_constraints; ++i) {
        const double d = problem->dual_start[i] - br->dual_solution[i];
        sq += d * d;
    }
    cp_al_parameters_t rescue = *req;
    double policy_time = 0.0;
    policy_set("CP_AL_FEATURE_CR_ITERS", "0");
    const clock_t pt = clock();
    apply_feature_auto_policy(problem, &rescue, stats);
    policy_time = (double)(clock() - pt) / CLOCKS_PER_SEC;
    rescue.termination_evaluation_frequency = 200;
    rescue.termination_criteria.time_sec_limit = remaining;
    policy_set("CP_AL_ALPHA_DIRECT_PULSE", pulse ? "1" : "0");
    policy_set("CP_AL_DETERMINISTIC_PULSE_FORCE", pulse ? "1" : "0");
    fprintf(stderr,
            "[structural-tail-start] profile=cr0 remaining_limit=%.12e anchor_copy_l2=%.12e policy_sec=%.12e "
            "tail_pulse_requested=%d tail_pulse_enabled=%d\n",
            remaining, std::sqrt(sq), policy_time, pulse_requested ? 1 : 0, pulse ? 1 : 0);
    cp_al_result_t *rr = solve_lp_problem(problem, &rescue);
    if (!rr) {
        fprintf(stderr, "[structural-tail-error] reason=null-rescue-result\n");
        return br;
    }
    fprintf(stderr, "[structural-tail-rescue] status=%s solve_sec=%.12e precondition_sec=%.12e iterations=%d\n",
            termination_reason_to_string(rr->termination_reason), rr->cumulative_time_sec, rr->rescaling_time_sec,
            rr->total_count);
    rr->cumulative_time_sec += bt;
    rr->rescaling_time_sec += bp;
    rr->presolve_time += bpre;
    rr->feasibility_polishing_time += bfp;
    rr->total_count += bi;
    fprintf(stderr,
            "[structural-tail-final] phases=2 tail_started=1 status=%s structural_early_cut=1 base_fraction=%.2f "
            "rescue_profile=cr0 base_solve_sec=%.12e rescue_solve_sec=%.12e aggregate_solve_sec=%.12e "
            "budget_excess=%.12e total_iterations=%d anchor_copy_l2=%.12e\n",
            termination_reason_to_string(rr->termination_reason), base_fraction, bt,
            rr->cumulative_time_sec - bt, rr->cumulative_time_sec, std::fmax(rr->cumulative_time_sec - limit, 0.0),
            rr->total_count, std::sqrt(sq));
    cp_al_result_free(br);
    return rr;
}

void usage(const char *prog)
{   // print_usage, cli.c:835-901
    fprintf(stderr, "Usage: %s [OPTIONS] <mps_file> <output_dir>\n\n", prog);
    fprintf(stderr, "Arguments:\n");
    fprintf(stderr, "  <mps_file>               Path to the input problem in MPS format (.mps or .mps.gz).\n");
    fprintf(stderr, "  <output_dir>             Directory where output files will be saved. It will contain:\n");
    fprintf(stderr, "                             - <basename>_summary.txt\n");
    fprintf(stderr, "                             - <basename>_primal_solution.txt\n");
    fprintf(stderr, "                             - <basename>_dual_solution.txt\n\n");
    fprintf(stderr, "Options:\n");
    fprintf(stderr, "  -h, --help                          Display this help message.\n");
    fprintf(stderr, "  -v, --verbose                       Enable verbose logging (default: false).\n");
    fprintf(stderr, "      --time_limit <seconds>          Time limit in seconds (default: 3600.0).\n");
    fprintf(stderr, "      --iter_limit <iterations>       Iteration limit (default: %d).\n", INT32_MAX);
    fprintf(stderr, "      --eps_opt <tolerance>           Relative optimality tolerance (default: 1e-4).\n");
    fprintf(stderr, "      --eps_feas <tolerance>          Relative feasibility tolerance (default: 1e-4).\n");
    fprintf(stderr, "      --l_inf_ruiz_iter <int>         Iterations for L-inf Ruiz rescaling (default: 10).\n");
    fprintf(stderr, "      --no_pock_chambolle             Disable Pock-Chambolle rescaling (default: enabled).\n");
    fprintf(stderr, "      --pock_chambolle_alpha <float>  Value for Pock-Chambolle alpha (default: 1.0).\n");
    fprintf(stderr, "      --no_bound_obj_rescaling        Disable bound objective rescaling (default: enabled).\n");
    fprintf(stderr, "      --eval_freq <int>               Termination evaluation frequency (default: 200).\n");
    fprintf(stderr, "      --sv_max_iter <int>             Max iterations for singular value estimation "
                    "(default: 5000).\n");
    fprintf(stderr, "      --sv_tol <float>                Tolerance for singular value estimation "
                    "(default: 1e-4).\n");
    fprintf(stderr, "  -f  --feasibility_polishing         Enable feasibility polishing (default: false).\n");
    fprintf(stderr, "      --eps_feas_polish <tolerance>   Relative feasibility polish tolerance (default: 1e-6).\n");
    fprintf(stderr, "      --opt_norm <norm_type>          Norm for optimality criteria: l2 or linf (default: l2).\n");
    fprintf(stderr, "      --no_presolve                   Disable presolve (default: enabled).\n");
    fprintf(stderr, "      --matrix_zero_tol <tolerance>.  Zero tolerance in constraint matrix.\n");
}

}  // namespace

int main(int argc, char *argv[])
{   // cli.c:903-1092
    cp_al_parameters_t params;
    set_default_parameters(&params);
    static struct option opts[] = {{"help", no_argument, 0, 'h'},
                                   {"verbose", no_argument, 0, 'v'},
                                   {"time_limit", required_argument, 0, 1001},
                                   {"iter_limit", required_argument, 0, 1002},
                                   {"eps_opt", required_argument, 0, 1003},
                                   {"eps_feas", required_argument, 0, 1004},
                                   {"eps_feas_polish", required_argument, 0, 1006},
                                   {"feasibility_polishing", no_argument, 0, 'f'},
                                   {"l_inf_ruiz_iter", required_argument, 0, 1007},
                                   {"pock_chambolle_alpha", required_argument, 0, 1008},
                                   {"no_pock_chambolle", no_argument, 0, 1009},
                                   {"no_bound_obj_rescaling", no_argument, 0, 1010},
                                   {"sv_max_iter", required_argument, 0, 1011},
                                   {"sv_tol", required_argument, 0, 1012},
                                   {"eval_freq", required_argument, 0, 1013},
                                   {"opt_norm", required_argument, 0, 1014},
                                   {"no_presolve", no_argument, 0, 1015},
                                   {"matrix_zero_tol", required_argument, 0, 1016},
                                   {0, 0, 0, 0}};
    int opt;
    while ((opt = getopt_long(argc, argv, "hvfp", opts, nullptr)) != -1) {
        switch (opt) {
        case 'h': usage(argv[0]); return 0;
        case 'v': params.verbose = true; break;
        case 1001: params.termination_criteria.time_sec_limit = atof(optarg); break;
        case 1002: params.termination_criteria.iteration_limit = atoi(optarg); break;
        case 1003: params.termination_criteria.eps_optimal_relative = atof(optarg); break;
        case 1004: params.termination_criteria.eps_feasible_relative = atof(optarg); break;
        case 1006: params.termination_criteria.eps_feas_polish_relative = atof(optarg); break;
        case 'f': params.feasibility_polishing = true; break;
        case 1007: params.l_inf_ruiz_iterations = atoi(optarg); break;
        case 1008: params.pock_chambolle_alpha = atof(optarg); break;
        case 1009: params.has_pock_chambolle_alpha = false; break;
        case 1010: params.bound_objective_rescaling = false; break;
        case 1011: params.sv_max_iter = atoi(optarg); break;
        case 1012: params.sv_tol = atof(optarg); break;
        case 1013: params.termination_evaluation_frequency = atoi(optarg); break;
        case 1014:
            if (strcmp(optarg, "l2") == 0) params.optimality_norm = NORM_TYPE_L2;
            else if (strcmp(optarg, "linf") == 0) params.optimality_norm = NORM_TYPE_L_INF;
            else {
                fprintf(stderr, "Error: opt_norm must be 'l2' or 'linf'\n");
                return 1;
            }
            break;
        case 1015: params.presolve = false; break;
        case 1016: params.matrix_zero_tol = atof(optarg); break;
        case '?': return 1;
        default: break;
        }
    }
    if (argc - optind != 2) {
        fprintf(stderr, "Error: You must specify an input file and an output directory.\n\n");
        usage(argv[0]);
        return 1;
    }
    const char *filename = argv[optind];
    const char *outdir = argv[optind + 1];
    const std::string name = instance_name(filename);
    lp_problem_t *problem = read_mps_file(filename);
    if (!problem) {
        fprintf(stderr, "Failed to read or parse the file.\n");
        return 1;
    }
    const bool unified = env_flag("CP_AL_UNIFIED_POLICY");
    const bool need_stats =
        unified || env_flag("CP_AL_FEATURE_AUTO_POLICY") || env_flag("CP_AL_STRUCTURAL_TAIL_CUT");
    Features stats;
    const clock_t t0 = clock();
    if (need_stats) stats = matrix_features(problem);
    apply_unified_policy(need_stats ? &stats : nullptr);
    const bool feature_policy = env_flag("CP_AL_FEATURE_AUTO_POLICY");
    const bool tail = env_flag("CP_AL_STRUCTURAL_TAIL_CUT");
    const bool tail_pulse = env_flag("CP_AL_TAIL_ONLY_PULSE");
    if (need_stats)
        fprintf(stderr, "[feature-matrix-stats] computed=1 feature_policy=%d structural_tail=%d\n",
                feature_policy ? 1 : 0, tail ? 1 : 0);
    apply_feature_auto_policy(problem, &params, need_stats ? &stats : nullptr);
    if (feature_policy)
        fprintf(stderr, "[feature-auto-policy-time] seconds=%.12g\n", (double)(clock() - t0) / CLOCKS_PER_SEC);
    if (env_flag("CP_AL_FEATURE_AUDIT_ONLY")) {
        lp_problem_free(problem);
        return 0;
    }
    cp_al_result_t *r = tail ? structural_tail(problem, &params, &stats, tail_pulse) : solve_lp_problem(problem, &params);
    if (!r) {
        fprintf(stderr, "Solver failed.\n");
    } else {
        save_summary(r, outdir, name);
        save_solution(r->primal_solution, problem->num_variables, outdir, name, "_primal_solution.txt");
        save_solution(r->dual_solution, problem->num_constraints, outdir, name, "_dual_solution.txt");
        cp_al_result_free(r);
    }
    lp_problem_free(problem);
    return 0;
}
