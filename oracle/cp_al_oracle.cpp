/*
 * CPU FP64 oracle for the CP_AL (LP) restarted reflected-Halpern solver.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path in
 * pdal_amd/ is compared against; nothing in the product links, loads or calls
 * it.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it (bench.py times it as the CPU baseline, "kind": "port").
 *
 * It exports the same C names as libcp_al.so (create_lp_problem,
 * solve_lp_problem, ...) so tests can drive both through one ctypes wrapper,
 * plus cp_al_oracle_reset_rng().  It is built as its own shared object with
 * hidden visibility and -Bsymbolic so its symbols never bind to the product's.
 *
 * Statement-by-statement restatement of the reference default path:
 *   optimize ................. LP/CP_AL/src/solver.cu:1175-1736
 *   initialize_solver_state .. solver.cu:1795-2273 (norms :1999-2067)
 *   iteration kernels ........ solver.cu:2340-2459, launch order :2523-2661
 *   fixed-point error ........ solver.cu:3611-3686
 *   perform_restart .......... solver.cu:2821-3101 (restart target of the candidate
 *                              selection; PID on the primal weight, or on sigma
 *                              in sigma mode :2857-2966)
 *   refresh_step_schedule .... solver.cu:3103-3126 (alpha and sigma mode)
 *   sigma parameters ......... solver.cu:3141-3164
 *   feasibility polishing .... solver.cu:3912-4456, utils.cu:1194-1440
 *   step size / primal weight  solver.cu:3538-3609
 *   create_result_from_state . solver.cu:3828-3909
 *   power iteration .......... LP/CP_AL/src/utils.cu:87-228 (global mt19937(1), :27-28)
 *   termination / restart .... utils.cu:275-359
 *   defaults ................. utils.cu:361-394
 *   residual ................. utils.cu:649-682, 812-917
 *   matrix filter (verbose) .. utils.cu:396-478
 *   conversions .............. utils.cu:1032-1192, cp_al.c:26-209
 *   rescaling ................ LP/CP_AL/src/preconditioner.cu:162-475, kernels :477-736
 *
 * Supported environment knobs (read where the reference reads them):
 *   CP_AL_LAMBDA, CP_AL_STEP_MARGIN, CP_AL_REFLECTION_COEFFICIENT,
 *   CP_AL_RESTART_SUFFICIENT / _NECESSARY / _ARTIFICIAL, CP_AL_CR_SCALING,
 *   CP_AL_CR_ITERS, and sigma mode: CP_AL_SIGMA_MODE, CP_AL_SIGMA,
 *   CP_AL_SIGMA_ADAPTIVE, CP_AL_SIGMA_INIT/MIN/MAX_FACTOR, CP_AL_LAMBDA_FACTOR,
 *   CP_AL_LAMBDA_PARAM, CP_AL_SIGMA_BALANCE_GAIN, CP_AL_SIGMA_I_SMOOTH,
 *   CP_AL_SIGMA_KP/KI/KD, CP_AL_SIGMA_UPDATE_MIN/MAX; the positive-alpha
 *   controllers (solver.cu:187-1057, used at :1202-1222, :1372-1494,
 *   :1631-1695): CP_AL_ALPHA_ROLLBACK_TRIAL / _TRACE / _FORCE_REJECT /
 *   _TEST_TRAJECTORY_LIMIT, CP_AL_ALPHA_DIRECT_PULSE / _TRACE /
 *   _TEST_TRAJECTORY_LIMIT, CP_AL_DETERMINISTIC_PULSE_FORCE (plus
 *   PDAL_ALPHA_TRIAL_MIN_SEC, a test knob for the trial's 2 s floor); restart
 *   candidate selection (solver.cu:2663-2819): CP_AL_RESTART_SELECT / _MIN_ITERS
 *   / _MIN_TIME / _THREEPOINT / _MARGIN; the sigma trigger (solver.cu:1496-1528,
 *   CP_AL_SIGMA_TRIGGER / _MIN_ITERS / _PRIMAL / _DUAL / _RATIO); the dynamic
 *   evaluation frequency (:1228-1260, 1533-1584, CP_AL_DYNAMIC_EVAL / _FAST /
 *   _AFTER_SEC / _AFTER_ITERS / _MIN_ITERS / _SMALL_RES / _LARGE_RES / _RATIO /
 *   _BOOTSTRAP_ITERS); adaptive lambda (:1059-1145, 3166-3536, CP_AL_ADAPTIVE_LAMBDA,
 *   CP_AL_LAMBDA_PROFILE, CP_AL_LAMBDA_GRID and the CP_AL_LAMBDA_* thresholds);
 *   the Curtis-Reid diagnostic and early stop (preconditioner.cu:171-252).
 *   PDAL_CONTROLLER_SECONDS_PER_ITERATION (a test knob shared with the HIP
 *   path) replaces the wall clock of the adaptive-lambda epochs and the
 *   dynamic evaluation's time gate by the iteration count times its value.
 *   With verbose: the dense trace (CP_AL_LOG_FREQ), the checkpoint lines and the
 *   progress lines (CP_AL_PROGRESS_LOG_FREQ), solver.cu:321-375.
 *
 * Deliberate differences from the reference, shared with the HIP path:
 *   - timing uses steady_clock wall time, not clock() CPU time;
 *   - presolve: PSLP is not vendored; the HIP path runs its own reductions
 *     (checked against HiGHS and the KKT conditions, not against this file),
 *     and this oracle ignores a request with a warning;
 *   - result->absolute_{primal,dual}_residual are filled in;
 *   - errors return NULL instead of exit().
 */
#include "cp_al.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <vector>

#define ORACLE_API extern "C" __attribute__((visibility("default")))

namespace {

// Per thread: a single-threaded caller sees the reference's process-global
// stream; the parallel CPU-baseline entry point gives every thread its own.
thread_local std::mt19937 g_gen(1);                              // utils.cu:27
thread_local std::normal_distribution<double> g_dist(0.0, 1.0);  // utils.cu:28

double env_double(const char *name, double def)  // solver.cu:377-391, utils.cu:33-46
{
    const char *raw = getenv(name);
    if (!raw || !raw[0]) return def;
    char *end = nullptr;
    const double v = strtod(raw, &end);
    if (end == raw || !std::isfinite(v)) return def;
    return v;
}

int env_int(const char *name, int def)  // solver.cu:393-407
{
    const char *raw = getenv(name);
    if (!raw || !raw[0]) return def;
    char *end = nullptr;
    const long v = strtol(raw, &end, 10);
    if (end == raw) return def;
    return (int)v;
}

struct Csr {
    int rows = 0, cols = 0;
    std::vector<int> rp, ci;
    std::vector<double> v;
};

void spmv(const Csr &A, const double *x, double *y)
{
    for (int i = 0; i < A.rows; ++i) {
        double s = 0.0;
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) s += A.v[k] * x[A.ci[k]];
        y[i] = s;
    }
}

double nrm2(const std::vector<double> &x)
{
    double s = 0.0;
    for (double a : x) s += a * a;
    return std::sqrt(s);
}

double dot(const std::vector<double> &a, const std::vector<double> &b)
{
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

double inf_norm(const std::vector<double> &x)
{
    double m = 0.0;
    for (double a : x) m = std::max(m, std::fabs(a));
    return m;
}

/* CSR -> CSR of the transpose; rows stay ascending inside each column
 * (the order cusparseCsr2cscEx2 produces, solver.cu:1859-1890). */
Csr transpose(const Csr &A)
{
    Csr T;
    T.rows = A.cols;
    T.cols = A.rows;
    T.rp.assign(A.cols + 1, 0);
    for (int k = 0; k < A.rp[A.rows]; ++k) T.rp[A.ci[k] + 1]++;
    for (int j = 0; j < A.cols; ++j) T.rp[j + 1] += T.rp[j];
    T.ci.resize(A.rp[A.rows]);
    T.v.resize(A.rp[A.rows]);
    std::vector<int> next(T.rp.begin(), T.rp.end() - 1);
    for (int i = 0; i < A.rows; ++i)
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const int p = next[A.ci[k]]++;
            T.ci[p] = i;
            T.v[p] = A.v[k];
        }
    return T;
}

struct State {
    int n = 0, m = 0;
    double obj_const = 0.0;
    Csr A, At;
    std::vector<double> c, lb, ub, l, u, l_fin, u_fin;
    std::vector<double> x0, x, xT, xH, xR, dual_prod, dual_slack, dual_res, dx;
    std::vector<double> y0, y, yT, yH, yR, prim_prod, primal_slack, primal_res, dy;
    std::vector<double> E, D;  // cumulative constraint / variable rescaling
    double b_scale = 1.0, c_scale = 1.0;
    double c_norm = 0.0, b_norm = 0.0;

    double step_size = 0.0, primal_weight = 1.0, al_lambda = 0.0, al_penalty = 0.0;
    double lambda_est = 1.0;
    // sigma mode (solver.cu:3141-3164)
    bool sigma_mode = false, sigma_adaptive = false;
    double al_sigma = 0.0, sigma_base = 0.0, sigma_min = 0.0, sigma_max = 0.0, sigma_best = 0.0;
    double sigma_err_sum = 0.0, sigma_last_err = 0.0, lambda_param = 0.0;
    double pw_error_sum = 0.0, pw_last_error = 0.0, best_pw = 1.0, best_gap = INFINITY;
    // adaptive lambda (solver.cu:3166-3536; internal_types.h:25 CP_AL_MAX_LAMBDA_TIERS = 16)
    bool adaptive_lambda = false, adaptive_lambda_metrics_ready = false, lambda_race = false;
    int lambda_num_tiers = 1, lambda_tier = 0, lambda_hold_remaining = 0, lambda_restart_count = 0;
    double lambda_values[16] = {};
    int lambda_hold_epochs = 2, lambda_min_restarts = 0, lambda_min_iterations = 0;
    int lambda_force_restart_iters = 0, lambda_last_restart_iter = 0;
    int lambda_race_next_tier = 1, lambda_race_best_tier = 0, lambda_race_exploit_remaining = 0;
    int lambda_race_exploit_epochs = 3, lambda_race_cooldown_remaining = 0, lambda_race_cooldown_epochs = 0;
    int lambda_race_cycle_best_tier = 0;
    double lambda_race_best_score = -INFINITY, lambda_race_cycle_best_score = -INFINITY;
    double lambda_race_bad_ratio = INFINITY, lambda_min_time_sec = 0.0;
    double lambda_feas_plateau_tol = 0.98, lambda_gap_plateau_tol = 0.95, lambda_fp_plateau_tol = 0.98;
    double lambda_feas_active_tol = 1e-6, lambda_gap_active_tol = 1e-6, lambda_fp_active_tol = 1e-8;
    double lambda_gap_dominance = 100.0;
    double lambda_epoch_feas_start = INFINITY, lambda_epoch_gap_start = INFINITY, lambda_epoch_fp_start = INFINITY;
    double lambda_epoch_time_start = 0.0;
    bool cr_early_stopped = false;
    int total_count = 0, inner_count = 0;
    double fp = 0.0, fp0 = 0.0, last_trial_fp = INFINITY;
    // restart candidate (choose_restart_candidate, solver.cu:2749-2819): 0 = T output,
    // 1 = the current (reflected Halpern) iterate, 2 = the plain Halpern iterate
    int restart_target = 0;
    double score_current = 0.0, score_average = INFINITY, score_halpern = INFINITY;

    double abs_p = 0, rel_p = 0, abs_d = 0, rel_d = 0, p_obj = 0, d_obj = 0, gap = 0, rel_gap = 0;
    termination_reason_t reason = TERMINATION_REASON_UNSPECIFIED;
    double rescaling_time = 0.0, solve_time = 0.0;
    double polish_time = 0.0;
    int polish_iters = 0;
};

/* ---------------- preconditioner.cu ---------------- */

void scale_problem(State &s, const std::vector<double> &e, const std::vector<double> &d,
                   const std::vector<double> &inv_e, const std::vector<double> &inv_d)
{   // preconditioner.cu:125-160 and kernels :477-531
    for (int j = 0; j < s.n; ++j) {
        s.c[j] *= inv_d[j];
        s.lb[j] *= d[j];
        s.ub[j] *= d[j];
        s.x0[j] *= d[j];
    }
    for (int i = 0; i < s.m; ++i) {
        s.l[i] *= inv_e[i];
        s.u[i] *= inv_e[i];
        s.y0[i] *= e[i];
    }
    for (int i = 0; i < s.m; ++i)
        for (int k = s.A.rp[i]; k < s.A.rp[i + 1]; ++k) s.A.v[k] *= inv_d[s.A.ci[k]] * inv_e[i];
    for (int j = 0; j < s.n; ++j)
        for (int k = s.At.rp[j]; k < s.At.rp[j + 1]; ++k) s.At.v[k] *= inv_d[j] * inv_e[s.At.ci[k]];
}

void row_absmax(const Csr &M, std::vector<double> &out)  // preconditioner.cu:533-551
{
    for (int i = 0; i < M.rows; ++i) {
        double mx = 0.0;
        for (int k = M.rp[i]; k < M.rp[i + 1]; ++k) {
            double v = std::fabs(M.v[k]);
            if (!std::isfinite(v)) v = 0.0;
            if (v > mx) mx = v;
        }
        out[i] = mx;
    }
}

double pow_fast(double v, double p)  // preconditioner.cu:652-661
{
    if (p == 2.0) return v * v;
    if (p == 1.0) return v;
    if (p == 0.5) return std::sqrt(v);
    return std::pow(v, p);
}

void row_powsum(const Csr &M, double degree, std::vector<double> &out)  // :663-681
{
    for (int i = 0; i < M.rows; ++i) {
        double acc = 0.0;
        for (int k = M.rp[i]; k < M.rp[i + 1]; ++k) {
            double v = std::fabs(M.v[k]);
            if (!std::isfinite(v)) v = 0.0;
            acc += pow_fast(v, degree);
        }
        out[i] = acc;
    }
}

void clamp_sqrt_accum(std::vector<double> &sf, std::vector<double> &inv, std::vector<double> &cum)
{   // preconditioner.cu:683-697, SCALING_EPSILON 1e-12
    for (size_t t = 0; t < sf.size(); ++t) {
        const double v = sf[t];
        const double q = (v < 1e-12) ? 1.0 : std::sqrt(v);
        cum[t] *= q;
        sf[t] = q;
        inv[t] = 1.0 / q;
    }
}

void cr_log_update(const Csr &M, const std::vector<double> &other, std::vector<double> &log_scale,
                   std::vector<double> *log_delta)
{   // preconditioner.cu:553-588 (log_delta feeds the diagnostic and the early stop)
    for (int r = 0; r < M.rows; ++r) {
        const int b = M.rp[r], e = M.rp[r + 1];
        if (e - b <= 0) {
            if (log_delta) (*log_delta)[r] = -log_scale[r];
            log_scale[r] = 0.0;
            continue;
        }
        double acc = 0.0;
        for (int k = b; k < e; ++k) {
            double v = std::fabs(M.v[k]);
            if (!std::isfinite(v) || v < 1e-300) v = 1e-300;
            acc += -std::log(v) - other[M.ci[k]];
        }
        const double updated = acc / (double)(e - b);
        if (log_delta) (*log_delta)[r] = updated - log_scale[r];
        log_scale[r] = updated;
    }
}

double cr_max_edge_change(const Csr &A, const std::vector<double> &row_delta, const std::vector<double> &col_delta)
{   // curtis_reid_edge_change_kernel (preconditioner.cu:590-609) + cub::DeviceReduce::Max (:222-228)
    double mx = 0.0;
    for (int r = 0; r < A.rows; ++r) {
        double row_max = 0.0;
        for (int k = A.rp[r]; k < A.rp[r + 1]; ++k) {
            const double ch = std::fabs(row_delta[r] + col_delta[A.ci[k]]);
            if (std::isfinite(ch) && ch > row_max) row_max = ch;
        }
        mx = r == 0 ? row_max : std::fmax(mx, row_max);
    }
    return mx;
}

void cr_exp_to_scale(std::vector<double> &ls, std::vector<double> &inv, std::vector<double> &cum)
{   // preconditioner.cu:611-625
    for (size_t i = 0; i < ls.size(); ++i) {
        double f = std::exp(ls[i]);
        f = std::fmin(std::fmax(f, 1e-30), 1e30);
        const double sc = 1.0 / f;
        ls[i] = sc;
        inv[i] = f;
        cum[i] *= sc;
    }
}

void rescale_problem(const cp_al_parameters_t *p, State &s)  // preconditioner.cu:379-475
{
    std::vector<double> e(s.m), d(s.n), inv_e(s.m), inv_d(s.n);
    s.E.assign(s.m, 1.0);
    s.D.assign(s.n, 1.0);

    const char *cr_raw = getenv("CP_AL_CR_SCALING");
    const bool cr_forced = cr_raw && cr_raw[0] != '\0' && atoi(cr_raw) != 0;
    int cr_iters = cr_forced ? 20 : 0;
    const char *cr_iters_raw = getenv("CP_AL_CR_ITERS");
    if (cr_iters > 0 && cr_iters_raw && cr_iters_raw[0] != '\0') cr_iters = atoi(cr_iters_raw);
    cr_iters = cr_iters < 0 ? 0 : cr_iters;
    s.cr_early_stopped = false;
    if (cr_iters > 0) {  // curtis_reid_rescaling, :162-268
        const bool emit_diagnostic = getenv("CP_AL_CR_DIAGNOSTIC") != NULL && atoi(getenv("CP_AL_CR_DIAGNOSTIC")) != 0;
        const bool use_early_stop =
            getenv("CP_AL_CR_EARLY_STOP") != NULL && atoi(getenv("CP_AL_CR_EARLY_STOP")) != 0 && cr_iters == 20;
        const char *thr_raw = getenv("CP_AL_CR_EARLY_STOP_THRESHOLD");
        const double early_stop_threshold = thr_raw != NULL ? atof(thr_raw) : 0.1;
        const bool measure = emit_diagnostic || use_early_stop;
        std::fill(e.begin(), e.end(), 0.0);
        std::fill(d.begin(), d.end(), 0.0);
        for (int it = 0; it < cr_iters; ++it) {
            cr_log_update(s.A, d, e, measure ? &inv_e : nullptr);
            cr_log_update(s.At, e, d, measure ? &inv_d : nullptr);
            if (measure) {
                const double max_edge_change = cr_max_edge_change(s.A, inv_e, inv_d);
                if (emit_diagnostic)
                    fprintf(stderr, "[cr-diagnostic] iteration=%d max_edge_log_change=%.12e\n", it + 1,
                            max_edge_change);
                if (use_early_stop && it + 1 == 5 && max_edge_change <= early_stop_threshold) {
                    s.cr_early_stopped = true;
                    printf("[cr-early-stop] selected=%d executed=%d max_edge_log_change=%.12e threshold=%.12e\n",
                           cr_iters, it + 1, max_edge_change, early_stop_threshold);
                    break;
                }
            }
        }
        cr_exp_to_scale(e, inv_e, s.E);
        cr_exp_to_scale(d, inv_d, s.D);
        scale_problem(s, e, d, inv_e, inv_d);
    }
    for (int it = 0; it < p->l_inf_ruiz_iterations; ++it) {  // ruiz_rescaling, :270-296
        row_absmax(s.A, e);
        clamp_sqrt_accum(e, inv_e, s.E);
        row_absmax(s.At, d);
        clamp_sqrt_accum(d, inv_d, s.D);
        scale_problem(s, e, d, inv_e, inv_d);
    }
    if (p->has_pock_chambolle_alpha) {  // pock_chambolle_rescaling, :298-323
        row_powsum(s.A, p->pock_chambolle_alpha, e);
        clamp_sqrt_accum(e, inv_e, s.E);
        row_powsum(s.At, 2.0 - p->pock_chambolle_alpha, d);
        clamp_sqrt_accum(d, inv_d, s.D);
        scale_problem(s, e, d, inv_e, inv_d);
    }
    s.b_scale = 1.0;
    s.c_scale = 1.0;
    if (p->bound_objective_rescaling) {  // bound_objective_rescaling, :325-377 + kernels :699-729
        double sq = 0.0;
        for (int i = 0; i < s.m; ++i) {
            const double L = s.l[i], U = s.u[i];
            const bool fL = std::isfinite(L), fU = std::isfinite(U);
            double acc = 0.0;
            if (fL && (!fU || std::fabs(L - U) > 1e-12)) acc += L * L;
            if (fU) acc += U * U;
            sq += acc;
        }
        const double bnd_norm = std::sqrt(sq);
        const double obj_norm = nrm2(s.c);
        const double cs = 1.0 / (bnd_norm + 1.0);
        const double os = 1.0 / (obj_norm + 1.0);
        for (int i = 0; i < s.m; ++i) {
            s.l[i] *= cs;
            s.u[i] *= cs;
            s.y0[i] *= os;
        }
        for (int j = 0; j < s.n; ++j) {
            s.lb[j] *= cs;
            s.ub[j] *= cs;
            s.c[j] *= os;
            s.x0[j] *= cs;
        }
        s.b_scale = cs;
        s.c_scale = os;
    }
}

/* ---------------- utils.cu ---------------- */

double estimate_max_singular_value(const State &s, int max_iter, double tol)  // utils.cu:87-228
{
    const int m = s.m;
    std::vector<double> v(m), nv(m), w(s.n);
    for (int i = 0; i < m; ++i) v[i] = g_dist(g_gen);
    double sigma_sq = 1.0;
    for (int it = 0; it < max_iter; ++it) {
        nv = v;
        const double inv = 1.0 / nrm2(nv);
        for (double &a : nv) a *= inv;
        spmv(s.At, nv.data(), w.data());
        spmv(s.A, w.data(), v.data());
        sigma_sq = dot(nv, v);
        const double neg = -sigma_sq;
        for (int i = 0; i < m; ++i) nv[i] = neg * nv[i] + v[i];
        if (nrm2(nv) < tol) break;
    }
    return std::sqrt(sigma_sq);
}

void compute_residual(State &s, norm_type_t norm)  // utils.cu:812-917
{
    spmv(s.A, s.xT.data(), s.prim_prod.data());
    spmv(s.At, s.yT.data(), s.dual_prod.data());
    for (int i = 0; i < s.m; ++i) {  // compute_residual_kernel, :649-682
        const double cl = std::fmax(s.l[i], std::fmin(s.prim_prod[i], s.u[i]));
        s.primal_res[i] = (s.prim_prod[i] - cl) * s.E[i];
        s.primal_slack[i] = std::fmax(s.yT[i], 0.0) * s.l_fin[i] + std::fmin(s.yT[i], 0.0) * s.u_fin[i];
    }
    for (int j = 0; j < s.n; ++j)
        s.dual_res[j] = (s.c[j] - s.dual_prod[j] - s.dual_slack[j]) * s.D[j];

    if (norm == NORM_TYPE_L_INF) {
        s.abs_p = inf_norm(s.primal_res);
        s.abs_d = inf_norm(s.dual_res);
    } else {
        s.abs_p = nrm2(s.primal_res) / s.b_scale;
        s.abs_d = nrm2(s.dual_res);
    }
    s.abs_d /= s.c_scale;
    const double bc = s.b_scale * s.c_scale;
    s.p_obj = dot(s.c, s.xT) / bc + s.obj_const;
    const double base = dot(s.dual_slack, s.xT);
    double sum = 0.0;
    for (double a : s.primal_slack) sum += a;
    s.d_obj = (base + sum) / bc + s.obj_const;
    s.rel_p = s.abs_p / (1.0 + s.b_norm);
    s.rel_d = s.abs_d / (1.0 + s.c_norm);
    s.gap = std::fabs(s.p_obj - s.d_obj);
    s.rel_gap = s.gap / (1.0 + std::fabs(s.p_obj) + std::fabs(s.d_obj));
}

/* ---------------- solver.cu ---------------- */

void refresh_step_schedule(State &s)  // solver.cu:3103-3126
{
    const double margin = env_double("CP_AL_STEP_MARGIN", 0.998 * 0.998);
    if (s.sigma_mode) {  // :3106-3126: tau = 1/(sigma lambda_param), dual step = penalty = sigma
        const double sigma = std::fmax(s.al_sigma, 0.0);
        if (sigma <= 0.0) {
            s.al_penalty = 0.0;
            s.lambda_param = 0.0;
            s.step_size = std::sqrt(std::fmax(margin / std::fmax(s.lambda_est, 1e-12), 1e-24));
            s.al_lambda = 0.0;
            return;
        }
        if (!std::isfinite(s.lambda_param) || s.lambda_param <= 0.0)
            s.lambda_param = env_double("CP_AL_LAMBDA_PARAM",
                                        env_double("CP_AL_LAMBDA_FACTOR", 2.01) * std::fmax(s.lambda_est, 1e-12));
        s.lambda_param = std::fmax(s.lambda_param, (2.0 + 1e-12) * std::fmax(s.lambda_est, 1e-12));
        const double tau = 1.0 / (sigma * s.lambda_param);
        const double rho = sigma;
        s.al_penalty = sigma;
        s.step_size = std::sqrt(std::fmax(tau * rho, 1e-24));
        s.primal_weight = std::sqrt(std::fmax(rho / tau, 1e-24));
        s.al_lambda = 0.0;
        return;
    }
    s.al_lambda = std::fmin(std::fmax(s.al_lambda, 0.0), 0.95);
    const double eta_sq = margin * (1.0 - s.al_lambda) / std::fmax(s.lambda_est, 1e-12);
    s.step_size = std::sqrt(std::fmax(eta_sq, 1e-24));
    const double dual_step = s.step_size * std::fmax(s.primal_weight, 1e-12);
    s.al_penalty = dual_step * s.al_lambda / std::fmax(1.0 - s.al_lambda, 1e-12);
}

/* adaptive lambda (solver.cu:1059-1145, 3166-3536) */

double controller_seconds(const State &s)
{   // cumulative_time_sec as the adaptive-lambda epochs and the dynamic evaluation
    // read it (utils.cu:301); PDAL_CONTROLLER_SECONDS_PER_ITERATION > 0 makes it
    // total_count times that value (a test knob shared with the HIP path)
    const double per_iter = env_double("PDAL_CONTROLLER_SECONDS_PER_ITERATION", 0.0);
    return per_iter > 0.0 ? per_iter * (double)s.total_count : s.solve_time;
}

int parse_lambda_grid_env(const char *raw, double *values, int max_values)  // :1059-1083
{
    if (!raw || !raw[0] || max_values <= 0) return 0;
    char buffer[512];
    strncpy(buffer, raw, sizeof(buffer) - 1);
    buffer[sizeof(buffer) - 1] = '\0';
    int count = 0;
    char *token = strtok(buffer, ",;: \t");
    while (token && count < max_values) {
        char *end = NULL;
        const double value = strtod(token, &end);
        if (end != token && std::isfinite(value)) values[count++] = std::fmin(std::fmax(value, 0.0), 0.95);
        token = strtok(NULL, ",;: \t");
    }
    return count;
}

void set_lambda_grid(State &s, const double *values, int count)  // :1085-1145
{
    double grid[16];
    int n = 0;
    if (count <= 0) {
        grid[n++] = 0.0;
    } else {
        bool has_zero = false;
        for (int i = 0; i < count && n < 16; ++i) {
            const double value = std::fmin(std::fmax(values[i], 0.0), 0.95);
            if (value <= 1e-15) has_zero = true;
            grid[n++] = value;
        }
        if (!has_zero && n < 16) grid[n++] = 0.0;
    }
    for (int i = 1; i < n; ++i) {
        const double key = grid[i];
        int j = i - 1;
        while (j >= 0 && grid[j] > key) {
            grid[j + 1] = grid[j];
            --j;
        }
        grid[j + 1] = key;
    }
    int unique_count = 0;
    for (int i = 0; i < n; ++i)
        if (unique_count == 0 || std::fabs(grid[i] - s.lambda_values[unique_count - 1]) > 1e-15)
            s.lambda_values[unique_count++] = grid[i];
    if (unique_count == 0) s.lambda_values[unique_count++] = 0.0;
    s.lambda_num_tiers = unique_count;
    for (int i = unique_count; i < 16; ++i) s.lambda_values[i] = s.lambda_values[unique_count - 1];
}

void initialize_adaptive_lambda(State &s)  // :3166-3262
{
    const char *lambda_profile = getenv("CP_AL_LAMBDA_PROFILE");
    const bool use_grid4_race_profile =
        lambda_profile && (strcmp(lambda_profile, "grid4_race") == 0 || strcmp(lambda_profile, "race4") == 0);
    const bool use_mp_race_profile =
        lambda_profile && (strcmp(lambda_profile, "mp_race") == 0 || strcmp(lambda_profile, "safe_race") == 0);
    const bool use_mp_auto_profile =
        lambda_profile && (strcmp(lambda_profile, "mp_auto") == 0 || strcmp(lambda_profile, "safe_auto") == 0);
    const bool use_small_grid_profile = use_grid4_race_profile || use_mp_race_profile || use_mp_auto_profile;

    s.adaptive_lambda = use_small_grid_profile || env_int("CP_AL_ADAPTIVE_LAMBDA", 0) != 0;
    s.adaptive_lambda_metrics_ready = false;
    s.lambda_tier = 0;
    s.lambda_hold_remaining = 0;
    s.lambda_restart_count = 0;

    double lambda_grid[16];
    int lambda_grid_count = parse_lambda_grid_env(getenv("CP_AL_LAMBDA_GRID"), lambda_grid, 16);
    if (lambda_grid_count == 0 && (use_mp_race_profile || use_mp_auto_profile)) {
        const double default_grid[] = {0.0, 0.005, 0.01, 0.02, 0.035, 0.05, 0.075, 0.10, 0.15, 0.20, 0.30, 0.40};
        lambda_grid_count = 12;
        for (int i = 0; i < lambda_grid_count; ++i) lambda_grid[i] = default_grid[i];
    }
    if (lambda_grid_count == 0) {
        lambda_grid[0] = env_double("CP_AL_LAMBDA_BASE", 0.0);
        lambda_grid[1] = env_double("CP_AL_LAMBDA_LOW", use_small_grid_profile ? 0.02 : 0.10);
        lambda_grid[2] = env_double("CP_AL_LAMBDA_MID", use_small_grid_profile ? 0.10 : 0.10);
        lambda_grid[3] = env_double("CP_AL_LAMBDA_HIGH", use_small_grid_profile ? 0.20 : 0.10);
        lambda_grid_count = 4;
    }
    set_lambda_grid(s, lambda_grid, lambda_grid_count);

    s.lambda_hold_epochs = env_int("CP_AL_LAMBDA_HOLD_EPOCHS", use_mp_auto_profile ? 1 : 2);
    s.lambda_min_restarts = env_int("CP_AL_LAMBDA_MIN_RESTARTS", use_mp_race_profile || use_mp_auto_profile
                                                                      ? 2
                                                                      : (use_grid4_race_profile ? 1 : 0));
    s.lambda_min_iterations =
        env_int("CP_AL_LAMBDA_MIN_ITERS", use_mp_race_profile || use_mp_auto_profile ? 10000 : 0);
    s.lambda_force_restart_iters = env_int("CP_AL_LAMBDA_FORCE_RESTART_ITERS", use_mp_race_profile ? 20000 : 0);
    s.lambda_last_restart_iter = 0;
    s.lambda_race = use_grid4_race_profile || use_mp_race_profile || env_int("CP_AL_LAMBDA_RACE", 0) != 0;
    s.lambda_race_next_tier = 1;
    s.lambda_race_best_tier = 0;
    s.lambda_race_exploit_remaining = 0;
    s.lambda_race_exploit_epochs = env_int("CP_AL_LAMBDA_RACE_EXPLOIT_EPOCHS",
                                           use_mp_race_profile ? 2 : (use_grid4_race_profile ? 9999 : 3));
    s.lambda_race_cooldown_remaining = 0;
    s.lambda_race_cooldown_epochs = env_int("CP_AL_LAMBDA_RACE_COOLDOWN_EPOCHS",
                                            use_mp_race_profile ? 3 : (use_grid4_race_profile ? 1000000000 : 0));
    s.lambda_race_best_score = -INFINITY;
    s.lambda_race_cycle_best_tier = 0;
    s.lambda_race_cycle_best_score = -INFINITY;
    s.lambda_race_bad_ratio = env_double("CP_AL_LAMBDA_RACE_BAD_RATIO",
                                         use_mp_race_profile ? 3.0 : (use_grid4_race_profile ? 10.0 : INFINITY));
    s.lambda_min_time_sec = env_double("CP_AL_LAMBDA_MIN_TIME", 0.0);
    s.lambda_feas_plateau_tol = env_double("CP_AL_LAMBDA_FEAS_PLATEAU", 0.98);
    s.lambda_gap_plateau_tol = env_double("CP_AL_LAMBDA_GAP_PLATEAU", 0.95);
    s.lambda_fp_plateau_tol = env_double("CP_AL_LAMBDA_FP_PLATEAU", 0.98);
    s.lambda_feas_active_tol = env_double("CP_AL_LAMBDA_FEAS_ACTIVE", 1e-6);
    s.lambda_gap_active_tol = env_double("CP_AL_LAMBDA_GAP_ACTIVE", 1e-6);
    s.lambda_fp_active_tol = env_double("CP_AL_LAMBDA_FP_ACTIVE", 1e-8);
    s.lambda_gap_dominance = env_double("CP_AL_LAMBDA_GAP_DOMINANCE", 100.0);

    s.lambda_hold_epochs = s.lambda_hold_epochs < 0 ? 0 : s.lambda_hold_epochs;
    s.lambda_min_restarts = s.lambda_min_restarts < 0 ? 0 : s.lambda_min_restarts;
    s.lambda_min_iterations = s.lambda_min_iterations < 0 ? 0 : s.lambda_min_iterations;
    s.lambda_force_restart_iters = s.lambda_force_restart_iters < 0 ? 0 : s.lambda_force_restart_iters;
    s.lambda_race_exploit_epochs = s.lambda_race_exploit_epochs < 0 ? 0 : s.lambda_race_exploit_epochs;
    s.lambda_race_cooldown_epochs = s.lambda_race_cooldown_epochs < 0 ? 0 : s.lambda_race_cooldown_epochs;
    s.lambda_race_bad_ratio = std::fmax(s.lambda_race_bad_ratio, 1.0);
    s.lambda_min_time_sec = std::fmax(s.lambda_min_time_sec, 0.0);

    if (s.adaptive_lambda) s.al_lambda = s.lambda_values[0];  // lambda_base
    s.lambda_epoch_feas_start = INFINITY;
    s.lambda_epoch_gap_start = INFINITY;
    s.lambda_epoch_fp_start = INFINITY;
    s.lambda_epoch_time_start = controller_seconds(s);
}

double lambda_from_tier(const State &s, int tier)  // :3264-3276
{
    if (s.lambda_num_tiers > 0) {
        if (tier < 0) tier = 0;
        if (tier >= s.lambda_num_tiers) tier = s.lambda_num_tiers - 1;
        return s.lambda_values[tier];
    }
    return 0.0;
}

void update_adaptive_lambda_on_restart(State &s, bool verbose)  // :3278-3536
{
    if (!s.adaptive_lambda) return;
    const double eps = 1e-300;
    const double now = controller_seconds(s);
    const double feas_end = std::fmax(s.rel_p, s.rel_d);
    const double gap_end = std::fmax(std::fabs(s.rel_gap), 0.0);
    const double fp_end = (std::isfinite(s.fp) && s.fp > 0.0) ? s.fp : INFINITY;

    s.lambda_restart_count += 1;
    if (!s.adaptive_lambda_metrics_ready) {
        s.lambda_epoch_feas_start = feas_end;
        s.lambda_epoch_gap_start = gap_end;
        s.lambda_epoch_fp_start = fp_end;
        s.lambda_epoch_time_start = now;
        s.adaptive_lambda_metrics_ready = true;
        if (verbose)
            printf("lambda-adapt init restart %d lambda %.3e feas %.3e gap %.3e fp %.3e\n", s.lambda_restart_count,
                   s.al_lambda, feas_end, gap_end, fp_end);
        return;
    }

    const double feas_start = s.lambda_epoch_feas_start;
    const double gap_start = s.lambda_epoch_gap_start;
    const double fp_start = s.lambda_epoch_fp_start;
    const double feas_ratio = std::fmax(feas_end, eps) / std::fmax(feas_start, eps);
    const double gap_ratio = std::fmax(gap_end, eps) / std::fmax(gap_start, eps);
    const double fp_ratio = std::fmax(fp_end, eps) / std::fmax(fp_start, eps);
    const double epoch_time = std::fmax(now - s.lambda_epoch_time_start, 1e-9);
    const bool mature_enough = s.lambda_restart_count >= s.lambda_min_restarts &&
                               s.total_count >= s.lambda_min_iterations && now >= s.lambda_min_time_sec;

    if (s.lambda_race) {
        const int old_tier = s.lambda_tier;
        const double old_lambda = s.al_lambda;
        double score = 0.45 * std::log(std::fmax(feas_start, eps) / std::fmax(feas_end, eps)) +
                       0.35 * std::log(std::fmax(fp_start, eps) / std::fmax(fp_end, eps)) +
                       0.20 * std::log(std::fmax(gap_start, eps) / std::fmax(gap_end, eps));
        score /= epoch_time;
        const bool bad_positive_epoch =
            old_tier > 0 && (feas_ratio > s.lambda_race_bad_ratio || gap_ratio > s.lambda_race_bad_ratio ||
                             fp_ratio > s.lambda_race_bad_ratio);
        if (!bad_positive_epoch && std::isfinite(score) && score > s.lambda_race_best_score) {
            s.lambda_race_best_score = score;
            s.lambda_race_best_tier = old_tier;
        }
        if (!bad_positive_epoch && std::isfinite(score) && score > s.lambda_race_cycle_best_score) {
            s.lambda_race_cycle_best_score = score;
            s.lambda_race_cycle_best_tier = old_tier;
        }
        const double worst_ratio = std::fmax(feas_ratio, std::fmax(gap_ratio, fp_ratio));
        const bool moderate_instability = worst_ratio < 3.0;
        const bool feasibility_only_lag = score > 5.0 && s.lambda_restart_count <= 10 && feas_ratio > 1.20 &&
                                          feas_ratio < 2.00 && gap_ratio < 1.20 && fp_ratio < 1.20;
        const bool feasibility_gain_tradeoff = score > 5.0 && feas_ratio < 0.50 &&
                                               std::fmax(gap_ratio, fp_ratio) > 3.00 &&
                                               worst_ratio < s.lambda_race_bad_ratio;
        const bool compact_plateau = score < -10.0 && moderate_instability && feas_ratio < 1.10;
        const bool race_plateau =
            mature_enough &&
            ((std::fmax(feas_start, feas_end) > s.lambda_feas_active_tol && feas_ratio > s.lambda_feas_plateau_tol) ||
             (std::fmax(gap_start, gap_end) > s.lambda_gap_active_tol && gap_ratio > s.lambda_gap_plateau_tol) ||
             (std::fmax(fp_start, fp_end) > s.lambda_fp_active_tol && fp_ratio > s.lambda_fp_plateau_tol));
        const bool start_race =
            old_tier > 0 || race_plateau || feasibility_only_lag || feasibility_gain_tradeoff || compact_plateau;

        int requested_tier = 0;
        if (bad_positive_epoch) {
            requested_tier = 0;
            s.lambda_race_next_tier = 1;
            s.lambda_race_exploit_remaining = 0;
            s.lambda_race_cooldown_remaining = s.lambda_race_cooldown_epochs;
            s.lambda_race_cycle_best_tier = 0;
            s.lambda_race_cycle_best_score = -INFINITY;
        } else if (s.lambda_race_cooldown_remaining > 0) {
            requested_tier = 0;
            s.lambda_race_cooldown_remaining -= 1;
            s.lambda_race_next_tier = 1;
            s.lambda_race_exploit_remaining = 0;
            s.lambda_race_cycle_best_tier = 0;
            s.lambda_race_cycle_best_score = std::isfinite(score) ? score : -INFINITY;
        } else if (!start_race) {
            requested_tier = 0;
            s.lambda_race_next_tier = 1;
            s.lambda_race_exploit_remaining = 0;
            s.lambda_race_cycle_best_tier = 0;
            s.lambda_race_cycle_best_score = std::isfinite(score) ? score : -INFINITY;
        } else if (!mature_enough) {
            requested_tier = 0;
            s.lambda_race_cycle_best_tier = 0;
            s.lambda_race_cycle_best_score = std::isfinite(score) ? score : -INFINITY;
        } else if (s.lambda_race_next_tier < s.lambda_num_tiers) {
            requested_tier = s.lambda_race_next_tier;
            if (s.lambda_race_next_tier == 1) {
                s.lambda_race_cycle_best_tier = old_tier;
                s.lambda_race_cycle_best_score = std::isfinite(score) ? score : -INFINITY;
            }
            s.lambda_race_next_tier += 1;
            s.lambda_race_exploit_remaining = s.lambda_race_exploit_epochs;
        } else if (s.lambda_race_exploit_remaining > 0) {
            requested_tier = s.lambda_race_cycle_best_tier;
            s.lambda_race_exploit_remaining -= 1;
        } else {
            s.lambda_race_next_tier = 1;
            requested_tier = s.lambda_race_cycle_best_tier;
            s.lambda_race_cycle_best_tier = requested_tier;
            s.lambda_race_cycle_best_score = std::isfinite(score) ? score : -INFINITY;
        }

        s.lambda_tier = requested_tier;
        s.lambda_hold_remaining = 0;
        s.al_lambda = lambda_from_tier(s, s.lambda_tier);
        s.lambda_epoch_feas_start = feas_end;
        s.lambda_epoch_gap_start = gap_end;
        s.lambda_epoch_fp_start = fp_end;
        s.lambda_epoch_time_start = now;
        if (verbose || old_tier != s.lambda_tier || std::fabs(old_lambda - s.al_lambda) > 1e-15)
            printf("lambda-race restart %d score %.3e best(tier=%d score=%.3e) cycle(tier=%d score=%.3e) "
                   "tier %d->%d lambda %.3e->%.3e next=%d exploit=%d cool=%d bad=%d start=%d mature=%d "
                   "ratios(feas %.3e gap %.3e fp %.3e)\n",
                   s.lambda_restart_count, score, s.lambda_race_best_tier, s.lambda_race_best_score,
                   s.lambda_race_cycle_best_tier, s.lambda_race_cycle_best_score, old_tier, s.lambda_tier, old_lambda,
                   s.al_lambda, s.lambda_race_next_tier, s.lambda_race_exploit_remaining,
                   s.lambda_race_cooldown_remaining, (int)bad_positive_epoch, (int)start_race, (int)mature_enough,
                   feas_ratio, gap_ratio, fp_ratio);
        return;
    }

    const bool feas_plateau =
        std::fmax(feas_start, feas_end) > s.lambda_feas_active_tol && feas_ratio > s.lambda_feas_plateau_tol;
    const bool gap_plateau =
        std::fmax(gap_start, gap_end) > s.lambda_gap_active_tol && gap_ratio > s.lambda_gap_plateau_tol;
    const bool fp_plateau =
        std::fmax(fp_start, fp_end) > s.lambda_fp_active_tol && fp_ratio > s.lambda_fp_plateau_tol;
    const bool gap_dominates = gap_end > s.lambda_gap_dominance * std::fmax(feas_end, eps);
    const int low_tier = s.lambda_num_tiers > 1 ? 1 : 0;
    const int mid_tier = s.lambda_num_tiers > 2 ? s.lambda_num_tiers / 2 : low_tier;
    const int high_tier = s.lambda_num_tiers > 0 ? s.lambda_num_tiers - 1 : 0;
    int requested_tier = 0;
    if (!mature_enough) requested_tier = 0;
    else if (gap_plateau && gap_dominates) requested_tier = high_tier;
    else if (feas_plateau || (gap_plateau && fp_plateau)) requested_tier = mid_tier;
    else if (gap_plateau || fp_plateau) requested_tier = low_tier;

    const int old_tier = s.lambda_tier;
    const double old_lambda = s.al_lambda;
    if (requested_tier > 0) {
        s.lambda_tier = requested_tier;
        s.lambda_hold_remaining = s.lambda_hold_epochs;
    } else if (s.lambda_hold_remaining > 0 && s.lambda_tier > 0) {
        s.lambda_hold_remaining -= 1;
    } else {
        s.lambda_tier = 0;
        s.lambda_hold_remaining = 0;
    }
    s.al_lambda = lambda_from_tier(s, s.lambda_tier);
    s.lambda_epoch_feas_start = feas_end;
    s.lambda_epoch_gap_start = gap_end;
    s.lambda_epoch_fp_start = fp_end;
    s.lambda_epoch_time_start = now;
    if (verbose || old_tier != s.lambda_tier || std::fabs(old_lambda - s.al_lambda) > 1e-15)
        printf("lambda-adapt restart %d ratios(feas %.3e gap %.3e fp %.3e) "
               "plateau(feas %d gap %d fp %d dom %d mature %d) tier %d->%d hold %d lambda %.3e->%.3e\n",
               s.lambda_restart_count, feas_ratio, gap_ratio, fp_ratio, (int)feas_plateau, (int)gap_plateau,
               (int)fp_plateau, (int)gap_dominates, (int)mature_enough, old_tier, s.lambda_tier,
               s.lambda_hold_remaining, old_lambda, s.al_lambda);
}

void iteration(State &s, int k_offset, double rho, bool major)
{
    const double tau = s.step_size / s.primal_weight;  // sync_step_sizes_to_gpu, solver.cu:1738-1749
    const double sigma = s.step_size * s.primal_weight;
    const double beta = s.al_penalty;
    const int k = s.inner_count + k_offset;
    const double w = (double)k / (double)(k + 1);

    // compute_next_primal_solution, solver.cu:2523-2607
    const double *dual_src = s.y.data();
    if (beta > 0.0) {
        spmv(s.A, s.x.data(), s.prim_prod.data());
        for (int i = 0; i < s.m; ++i) {  // compute_box_al_predictor_kernel, :2459-2479
            const double z = std::fmax(s.l[i], std::fmin(s.prim_prod[i] - s.y[i] / beta, s.u[i]));
            s.yR[i] = s.y[i] - beta * (s.prim_prod[i] - z);
        }
        dual_src = s.yR.data();
    }
    spmv(s.At, dual_src, s.dual_prod.data());
    for (int j = 0; j < s.n; ++j) {  // kernels :2340-2397
        const double xj = s.x[j];
        const double temp = xj - tau * (s.c[j] - s.dual_prod[j]);
        const double proj = std::fmax(s.lb[j], std::fmin(temp, s.ub[j]));
        if (major) {
            s.xT[j] = proj;
            s.dual_slack[j] = (proj - temp) / tau;
            s.xH[j] = w * proj + (1.0 - w) * s.x0[j];
        }
        s.xR[j] = 2.0 * proj - xj;
        const double refl = rho * s.xR[j] + (1.0 - rho) * xj;
        s.x[j] = w * refl + (1.0 - w) * s.x0[j];
    }

    // compute_next_dual_solution, solver.cu:2609-2661
    spmv(s.A, s.xR.data(), s.prim_prod.data());
    const double prox = sigma + beta;
    for (int i = 0; i < s.m; ++i) {  // kernels :2399-2457
        const double yi = s.y[i];
        const double pp = s.prim_prod[i];
        const double zb = std::fmax(s.l[i], std::fmin(pp - yi / prox, s.u[i]));
        const double yn = yi - sigma * (pp - zb);
        if (major) {
            s.yT[i] = yn;
            s.yH[i] = w * yn + (1.0 - w) * s.y0[i];
            s.yR[i] = 2.0 * yn - yi;
        }
        const double refl = rho * (2.0 * yn - yi) + (1.0 - rho) * yi;
        s.y[i] = w * refl + (1.0 - w) * s.y0[i];
    }
}

void compute_fixed_point_error(State &s)  // solver.cu:3611-3686
{
    for (int j = 0; j < s.n; ++j) s.dx[j] = s.xR[j] - s.xT[j];
    for (int i = 0; i < s.m; ++i) s.dy[i] = s.yR[i] - s.yT[i];
    spmv(s.At, s.dy.data(), s.dual_prod.data());
    spmv(s.A, s.dx.data(), s.prim_prod.data());
    const double dn = nrm2(s.dy), pn = nrm2(s.dx), an = nrm2(s.prim_prod);
    if (s.sigma_mode && s.al_sigma > 0.0 && s.lambda_param > 0.0) {  // :3664-3671
        const double val = s.al_sigma * (s.lambda_param * pn * pn - an * an) + dn * dn / std::fmax(s.al_sigma, 1e-16);
        s.fp = std::sqrt(std::fmax(val, 0.0));
        return;
    }
    const double movement = pn * pn * s.primal_weight + dn * dn / s.primal_weight;
    const double cross = dot(s.dual_prod, s.dx);
    const double interaction = 2 * s.step_size * cross;
    const double aug = s.step_size * s.al_penalty * an * an;
    s.fp = std::sqrt(std::fmax(movement + interaction - aug, 0.0));
}

bool should_restart(State &s, const restart_parameters_t &rp, int K)  // utils.cu:319-359
{
    const double suff = env_double("CP_AL_RESTART_SUFFICIENT", rp.sufficient_reduction_for_restart);
    const double nec = env_double("CP_AL_RESTART_NECESSARY", rp.necessary_reduction_for_restart);
    const double art = env_double("CP_AL_RESTART_ARTIFICIAL", rp.artificial_restart_threshold);
    bool r = false;
    if (s.total_count == K) {
        r = true;
    } else if (s.total_count > K) {
        if (s.fp <= suff * s.fp0) r = true;
        if (s.fp <= nec * s.fp0 && s.fp > s.last_trial_fp) r = true;
        if (s.inner_count >= art * s.total_count) r = true;
    }
    s.last_trial_fp = s.fp;
    return r;
}

void perform_restart(State &s, const cp_al_parameters_t *p)  // solver.cu:2821-3101
{
    const int target = env_int("CP_AL_RESTART_SELECT", 0) != 0 ? s.restart_target : 0;  // :2823-2834
    const std::vector<double> &rx = target == 1 ? s.x : target == 2 ? s.xH : s.xT;
    const std::vector<double> &ry = target == 1 ? s.y : target == 2 ? s.yH : s.yT;
    double pd = 0.0, dd = 0.0;
    for (int j = 0; j < s.n; ++j) {
        const double t = rx[j] - s.x0[j];
        pd += t * t;
    }
    for (int i = 0; i < s.m; ++i) {
        const double t = ry[i] - s.y0[i];
        dd += t * t;
    }
    pd = std::sqrt(pd);
    dd = std::sqrt(dd);
    const double ratio = s.rel_d / std::fmax(s.rel_p, 1e-300);
    if (s.sigma_mode) {  // solver.cu:2857-2966
        if (s.sigma_adaptive) {
            for (int j = 0; j < s.n; ++j) s.dx[j] = rx[j] - s.x0[j];
            spmv(s.A, s.dx.data(), s.prim_prod.data());
            const double ax = nrm2(s.prim_prod);
            const double dx_metric_sq = s.lambda_param * pd * pd - ax * ax;
            if (dd > 1e-16 && dd < 1e12 && dx_metric_sq > 1e-24 && dx_metric_sq < 1e24 && std::isfinite(s.al_sigma) &&
                s.al_sigma > 1e-16 && ratio > 1e-8 && ratio < 1e8) {
                const restart_parameters_t &rp = p->restart_params;
                const double gain = env_double("CP_AL_SIGMA_BALANCE_GAIN", 0.15);
                const double smooth = env_double("CP_AL_SIGMA_I_SMOOTH", rp.i_smooth);
                const double kp = env_double("CP_AL_SIGMA_KP", rp.k_p);
                const double ki = env_double("CP_AL_SIGMA_KI", rp.k_i);
                const double kd = env_double("CP_AL_SIGMA_KD", rp.k_d);
                const double rmin = env_double("CP_AL_SIGMA_UPDATE_MIN", 0.60);
                const double rmax = env_double("CP_AL_SIGMA_UPDATE_MAX", 1.70);
                double err = std::log(std::fmax(dd, 1e-300)) - 0.5 * std::log(std::fmax(dx_metric_sq, 1e-300)) -
                             std::log(std::fmax(s.al_sigma, 1e-300));
                double bal = 0.5 * std::log(std::fmax(ratio, 1e-300));
                bal = std::fmax(-2.0, std::fmin(bal, 2.0));
                err += gain * bal;
                s.sigma_err_sum = smooth * s.sigma_err_sum + err;
                const double de = err - s.sigma_last_err;
                double r = std::exp(kp * err + ki * s.sigma_err_sum + kd * de);
                r = std::fmin(std::fmax(r, rmin), rmax);
                s.al_sigma *= r;
                s.sigma_last_err = err;
            } else {
                s.al_sigma = s.sigma_best;
                s.sigma_err_sum = 0.0;
                s.sigma_last_err = 0.0;
            }
        }
        if (!std::isfinite(s.al_sigma) || s.al_sigma <= 1e-16) {
            s.al_sigma = s.sigma_best;
            s.sigma_err_sum = 0.0;
            s.sigma_last_err = 0.0;
        }
        s.al_sigma = std::fmin(std::fmax(s.al_sigma, s.sigma_min), s.sigma_max);
        const double g = std::fabs(std::log10(std::fmax(s.rel_d, 1e-300) / std::fmax(s.rel_p, 1e-300)));
        if (std::isfinite(g) && g < s.best_gap) {
            s.best_gap = g;
            s.sigma_best = s.al_sigma;
        }
        s.pw_error_sum = 0.0;
        s.pw_last_error = 0.0;
    } else if (pd > 1e-16 && dd > 1e-16 && pd < 1e12 && dd < 1e12 && ratio > 1e-8 && ratio < 1e8) {
        const double err = std::log(dd) - std::log(pd) - std::log(s.primal_weight);
        s.pw_error_sum *= p->restart_params.i_smooth;
        s.pw_error_sum += err;
        const double de = err - s.pw_last_error;
        s.primal_weight *= std::exp(p->restart_params.k_p * err + p->restart_params.k_i * s.pw_error_sum +
                                    p->restart_params.k_d * de);
        s.pw_last_error = err;
    } else {
        s.primal_weight = s.best_pw;
        s.pw_error_sum = 0.0;
        s.pw_last_error = 0.0;
    }
    if (!s.sigma_mode) {
        const double g = std::fabs(std::log10(std::fmax(s.rel_d, 1e-300) / std::fmax(s.rel_p, 1e-300)));
        if (g < s.best_gap) {
            s.best_gap = g;
            s.best_pw = s.primal_weight;
        }
        update_adaptive_lambda_on_restart(s, p->verbose);  // :2996
    }
    refresh_step_schedule(s);
    const int select_mode = env_int("CP_AL_RESTART_SELECT", 0);
    if (select_mode != 0 && (p->verbose || env_int("CP_AL_RESTART_SELECT_TRACE", 0) != 0))  // :3000-3015
        fprintf(stderr,
                "[restart-select] iter=%d target=%s score_current=%.12e score_reflected_halpern=%.12e "
                "score_halpern=%.12e candidate=%.12e margin=%.6e mode=%d\n",
                s.total_count, target == 1 ? "reflected_halpern" : target == 2 ? "halpern" : "current",
                s.score_current, s.score_average, s.score_halpern, s.fp,
                env_double("CP_AL_RESTART_SELECT_MARGIN", 1.0), select_mode);
    if (target == 1) {  // :3019-3043: anchor = T output = the current iterate, reflected = old T output
        s.xR = s.xT;
        s.yR = s.yT;
        s.x0 = s.x;
        s.xT = s.x;
        s.y0 = s.y;
        s.yT = s.y;
    } else if (target == 2) {  // :3044-3076
        s.xR = s.xT;
        s.yR = s.yT;
        s.x0 = s.xH;
        s.x = s.xH;
        s.xT = s.xH;
        s.y0 = s.yH;
        s.y = s.yH;
        s.yT = s.yH;
    } else {
        s.x0 = s.xT;
        s.x = s.xT;
        s.y0 = s.yT;
        s.y = s.yT;
    }
    s.lambda_last_restart_iter = s.total_count;  // :3098
    s.inner_count = 0;
    s.last_trial_fp = INFINITY;
}

/* ---------------- restart candidate (solver.cu:2663-2819) ---------------- */

double restart_delta_score(State &s, const std::vector<double> &cx, const std::vector<double> &rx,
                           const std::vector<double> &cy, const std::vector<double> &ry)
{   // compute_restart_delta_score, solver.cu:2663-2747: the fixed-point metric of (candidate - reference)
    for (int j = 0; j < s.n; ++j) s.dx[j] = cx[j] - rx[j];
    for (int i = 0; i < s.m; ++i) s.dy[i] = cy[i] - ry[i];
    spmv(s.At, s.dy.data(), s.dual_prod.data());
    spmv(s.A, s.dx.data(), s.prim_prod.data());
    const double pn = nrm2(s.dx), dn = nrm2(s.dy), an = nrm2(s.prim_prod);
    if (s.sigma_mode && s.al_sigma > 0.0 && s.lambda_param > 0.0) {
        const double val = s.al_sigma * (s.lambda_param * pn * pn - an * an) + dn * dn / std::fmax(s.al_sigma, 1e-16);
        return std::sqrt(std::fmax(val, 0.0));
    }
    const double cross = dot(s.dual_prod, s.dx);
    const double movement = pn * pn * s.primal_weight + dn * dn / s.primal_weight;
    return std::sqrt(std::fmax(movement + 2 * s.step_size * cross - s.step_size * s.al_penalty * an * an, 0.0));
}

void choose_restart_candidate(State &s)  // solver.cu:2749-2819
{
    const int mode = env_int("CP_AL_RESTART_SELECT", 0);
    s.restart_target = 0;
    s.score_current = s.fp;
    s.score_average = INFINITY;
    s.score_halpern = INFINITY;
    if (mode == 0) return;
    if (s.total_count < env_int("CP_AL_RESTART_SELECT_MIN_ITERS", 0) ||
        s.solve_time < env_double("CP_AL_RESTART_SELECT_MIN_TIME", 0.0))
        return;
    s.score_average = restart_delta_score(s, s.xT, s.x, s.yT, s.y);
    const bool three = mode == 3 || env_int("CP_AL_RESTART_SELECT_THREEPOINT", 0) != 0;
    if (three) s.score_halpern = restart_delta_score(s, s.xT, s.xH, s.yT, s.yH);
    const double margin = env_double("CP_AL_RESTART_SELECT_MARGIN", 1.0);
    if (mode == 2) {
        s.restart_target = 1;
        s.fp = s.score_average;
    } else if (std::isfinite(s.score_current)) {
        int best = 0;
        double score = s.score_current;
        if (std::isfinite(s.score_average) && s.score_average < score) {
            best = 1;
            score = s.score_average;
        }
        if (three && std::isfinite(s.score_halpern) && s.score_halpern < score) {
            best = 2;
            score = s.score_halpern;
        }
        if (best != 0 && score <= margin * s.score_current) {
            s.restart_target = best;
            s.fp = score;
        } else {
            s.fp = s.score_current;
        }
    } else {
        s.fp = s.score_current;
    }
}

/* ---------------- feasibility polishing (solver.cu:3912-4456) ---------------- */

const char *reason_name(termination_reason_t r);

struct PolishOut {
    termination_reason_t reason;
    int total;
    double time;
};

// primal_feasibility_polish / dual_feasibility_polish (solver.cu:3980-4127) on
// the phase state ps; ori is the main state.
PolishOut polish_loop(const cp_al_parameters_t *p, State &ps, const State &ori, bool primal,
                      std::chrono::steady_clock::time_point t_start)
{
    const auto t0 = std::chrono::steady_clock::now();
    const int K = p->termination_evaluation_frequency;
    const double rho = p->reflection_coefficient;
    const termination_criteria_t &tc = p->termination_criteria;
    const bool linf = p->optimality_norm == NORM_TYPE_L_INF;
    auto phase_fp = [&]() {  // compute_primal/dual_fixed_point_error, solver.cu:4421-4456
        if (primal) {
            for (int j = 0; j < ps.n; ++j) ps.dx[j] = ps.xR[j] - ps.xT[j];
            const double pn = nrm2(ps.dx);
            ps.fp = pn * pn * ps.primal_weight;
        } else {
            for (int i = 0; i < ps.m; ++i) ps.dy[i] = ps.yR[i] - ps.yT[i];
            const double dn = nrm2(ps.dy);
            ps.fp = dn * dn / ps.primal_weight;
        }
    };
    PolishOut out{TERMINATION_REASON_UNSPECIFIED, 0, 0.0};
    bool do_restart = false;
    while (ps.reason == TERMINATION_REASON_UNSPECIFIED) {
        iteration(ps, 1, rho, true);
        if (do_restart) {
            phase_fp();
            ps.fp0 = ps.fp;
            do_restart = false;
        }
        for (int i = 2; i <= K; ++i) iteration(ps, i, rho, i == K);
        phase_fp();
        const double bc = ps.b_scale * ps.c_scale;
        if (primal) {  // compute_primal_feas_polish_residual, utils.cu:1341-1392
            spmv(ps.A, ps.xT.data(), ps.prim_prod.data());
            for (int i = 0; i < ps.m; ++i) {
                const double cl = std::fmax(ps.l[i], std::fmin(ps.prim_prod[i], ps.u[i]));
                ps.primal_res[i] = (ps.prim_prod[i] - cl) * ps.E[i];
            }
            ps.abs_p = (linf ? inf_norm(ps.primal_res) : nrm2(ps.primal_res)) / ps.b_scale;
            ps.rel_p = ps.abs_p / (1.0 + ps.b_norm);
            ps.p_obj = dot(ori.c, ps.xT) / bc + ps.obj_const;
        } else {  // compute_dual_feas_polish_residual, utils.cu:1394-1440
            spmv(ps.At, ps.yT.data(), ps.dual_prod.data());
            for (int j = 0; j < ps.n; ++j) ps.dual_res[j] = (ps.c[j] - ps.dual_prod[j] - ps.dual_slack[j]) * ps.D[j];
            double contrib = 0.0;
            for (int i = 0; i < ps.m; ++i)
                contrib += std::fmax(ps.yT[i], 0.0) * ori.l_fin[i] + std::fmin(ps.yT[i], 0.0) * ori.u_fin[i];
            ps.abs_d = (linf ? inf_norm(ps.dual_res) : nrm2(ps.dual_res)) / ps.c_scale;
            ps.rel_d = ps.abs_d / (1.0 + ps.c_norm);
            ps.d_obj = (dot(ps.dual_slack, ori.xT) + contrib) / bc + ps.obj_const;
        }
        ps.inner_count += K;
        ps.total_count += K;
        out.time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        // check_feas_polishing_termination_criteria, utils.cu:1194-1227
        if ((primal ? ps.rel_p : ps.rel_d) <= tc.eps_feas_polish_relative)
            ps.reason = TERMINATION_REASON_FEAS_POLISH_SUCCESS;
        else if (ps.total_count >= tc.iteration_limit)
            ps.reason = TERMINATION_REASON_ITERATION_LIMIT;
        else if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count() >=
                 tc.time_sec_limit)
            ps.reason = TERMINATION_REASON_TIME_LIMIT;
        do_restart = should_restart(ps, p->restart_params, K);
        if (do_restart) {  // perform_primal_restart / perform_dual_restart, solver.cu:4391-4419
            if (primal) {
                ps.x0 = ps.xT;
                ps.x = ps.xT;
            } else {
                ps.y0 = ps.yT;
                ps.y = ps.yT;
            }
            ps.inner_count = 0;
            ps.last_trial_fp = INFINITY;
        }
    }
    out.reason = ps.reason;
    out.total = ps.total_count;
    return out;
}

void reset_phase(State &ps, double pw0)
{   // scalar resets of initialize_primal/dual_feas_polish_state (solver.cu:4172-4190, 4318-4336)
    ps.pw_error_sum = 0.0;
    ps.pw_last_error = 0.0;
    ps.primal_weight = pw0;
    ps.best_pw = pw0;
    ps.fp = ps.fp0 = ps.last_trial_fp = INFINITY;
    ps.total_count = 0;
    ps.inner_count = 0;
    ps.reason = TERMINATION_REASON_UNSPECIFIED;
    ps.best_gap = INFINITY;
}

void feasibility_polish(const cp_al_parameters_t *p, State &s, std::chrono::steady_clock::time_point t_start)
{   // solver.cu:3912-3978
    const auto t0 = std::chrono::steady_clock::now();
    const termination_criteria_t &tc = p->termination_criteria;
    if (s.rel_p < tc.eps_feas_polish_relative && s.rel_d < tc.eps_feas_polish_relative) {
        printf("Skipping feasibility polishing as the solution is already sufficiently feasible.\n");
        fflush(stdout);
        return;
    }
    const double pw0 = p->bound_objective_rescaling ? 1.0 : (s.c_norm + 1.0) / (s.b_norm + 1.0);

    // primal phase (initialize_primal_feas_polish_state, solver.cu:4129-4190): c = 0, constant 0,
    // dual iterates and slacks zero, primal iterates from the main state
    State ps = s;
    std::fill(ps.c.begin(), ps.c.end(), 0.0);
    ps.obj_const = 0.0;
    for (auto *v : {&ps.y0, &ps.y, &ps.yT, &ps.yH, &ps.yR, &ps.primal_slack, &ps.primal_res, &ps.dy})
        std::fill(v->begin(), v->end(), 0.0);
    for (auto *v : {&ps.dual_prod, &ps.dual_slack, &ps.dual_res, &ps.dx}) std::fill(v->begin(), v->end(), 0.0);
    reset_phase(ps, pw0);
    ps.rel_d = ps.abs_d = ps.rel_gap = ps.gap = 0.0;
    const PolishOut pr = polish_loop(p, ps, s, true, t_start);
    if (pr.reason == TERMINATION_REASON_FEAS_POLISH_SUCCESS) {
        s.xT = ps.xT;
        s.abs_p = ps.abs_p;
        s.rel_p = ps.rel_p;
        s.p_obj = ps.p_obj;
    }
    s.polish_iters += ps.total_count - 1;

    // dual phase (initialize_dual_feas_polish_state, solver.cu:4209-4337): finite bounds 0,
    // primal iterates zero, dual iterates and slack from the main state
    State ds = s;
    for (auto *v : {&ds.l, &ds.u, &ds.lb, &ds.ub})
        for (double &b : *v)
            if (std::isfinite(b)) b = 0.0;
    std::fill(ds.l_fin.begin(), ds.l_fin.end(), 0.0);
    std::fill(ds.u_fin.begin(), ds.u_fin.end(), 0.0);
    for (auto *v : {&ds.x0, &ds.x, &ds.xT, &ds.xH, &ds.xR, &ds.dual_res, &ds.dx}) std::fill(v->begin(), v->end(), 0.0);
    for (auto *v : {&ds.prim_prod, &ds.primal_slack, &ds.primal_res, &ds.dy}) std::fill(v->begin(), v->end(), 0.0);
    reset_phase(ds, pw0);
    ds.rel_p = ds.abs_p = ds.rel_gap = ds.gap = 0.0;
    const PolishOut du = polish_loop(p, ds, s, false, t_start);
    if (du.reason == TERMINATION_REASON_FEAS_POLISH_SUCCESS) {
        s.yT = ds.yT;
        s.abs_d = ds.abs_d;
        s.rel_d = ds.rel_d;
        s.d_obj = ds.d_obj;
    }
    s.polish_iters += ds.total_count - 1;
    s.gap = std::fabs(s.p_obj - s.d_obj);
    s.rel_gap = s.gap / (1.0 + std::fabs(s.p_obj) + std::fabs(s.d_obj));
    // cp_al_feas_polish_final_log, utils.cu:1247-1268
    if (p->verbose) printf("---------------------------------------------------------------------------------------\n");
    printf("Feasibility Polishing Summary\n");
    printf("  Primal Status        : %s\n", reason_name(pr.reason));
    printf("  Primal Iterations    : %d\n", pr.total);
    printf("  Primal Time Usage    : %.3g sec\n", pr.time);
    printf("  Dual Status          : %s\n", reason_name(du.reason));
    printf("  Dual Iterations      : %d\n", du.total);
    printf("  Dual Time Usage      : %.3g sec\n", du.time);
    printf("  Primal Residual      : %.3e\n", ps.rel_p);
    printf("  Dual Residual        : %.3e\n", ds.rel_d);
    printf("  Primal Dual Gap      : %.3e\n",
           std::fabs(ps.p_obj - ds.d_obj) / (1.0 + std::fabs(ps.p_obj) + std::fabs(ds.d_obj)));
    fflush(stdout);
    s.polish_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

const char *reason_name(termination_reason_t r)
{
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

void filter_matrix(lp_problem_t *pr, const cp_al_parameters_t *p)  // utils.cu:396-478
{
    const int rows = pr->num_constraints, nnz = pr->constraint_matrix_num_nonzeros;
    if (rows == 0 || nnz == 0) return;
    int kept = 0;
    for (int k = 0; k < nnz; ++k)
        if (std::fabs(pr->constraint_matrix_values[k]) > p->matrix_zero_tol) ++kept;
    if (kept == nnz) return;
    int *rp = (int *)malloc((size_t)(rows + 1) * sizeof(int));
    int *ci = (int *)malloc((size_t)std::max(kept, 1) * sizeof(int));
    double *v = (double *)malloc((size_t)std::max(kept, 1) * sizeof(double));
    int pos = 0;
    rp[0] = 0;
    for (int i = 0; i < rows; ++i) {
        for (int k = pr->constraint_matrix_row_pointers[i]; k < pr->constraint_matrix_row_pointers[i + 1]; ++k) {
            const double a = pr->constraint_matrix_values[k];
            if (std::fabs(a) <= p->matrix_zero_tol) continue;
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
}

void init_state(State &s, const lp_problem_t *w, const cp_al_parameters_t *p)  // solver.cu:1795-2273
{
    s.n = w->num_variables;
    s.m = w->num_constraints;
    s.obj_const = w->objective_constant;
    const int nnz = w->constraint_matrix_num_nonzeros;
    s.A.rows = s.m;
    s.A.cols = s.n;
    s.A.rp.assign(w->constraint_matrix_row_pointers, w->constraint_matrix_row_pointers + s.m + 1);
    s.A.ci.assign(w->constraint_matrix_col_indices, w->constraint_matrix_col_indices + nnz);
    s.A.v.assign(w->constraint_matrix_values, w->constraint_matrix_values + nnz);
    s.At = transpose(s.A);
    s.c.assign(w->objective_vector, w->objective_vector + s.n);
    s.lb.assign(w->variable_lower_bound, w->variable_lower_bound + s.n);
    s.ub.assign(w->variable_upper_bound, w->variable_upper_bound + s.n);
    s.l.assign(w->constraint_lower_bound, w->constraint_lower_bound + s.m);
    s.u.assign(w->constraint_upper_bound, w->constraint_upper_bound + s.m);
    for (auto *v : {&s.x0, &s.x, &s.xT, &s.xH, &s.xR, &s.dual_prod, &s.dual_slack, &s.dual_res, &s.dx})
        v->assign(s.n, 0.0);
    for (auto *v : {&s.y0, &s.y, &s.yT, &s.yH, &s.yR, &s.prim_prod, &s.primal_slack, &s.primal_res, &s.dy})
        v->assign(s.m, 0.0);
    if (w->primal_start) s.x0.assign(w->primal_start, w->primal_start + s.n);
    if (w->dual_start) s.y0.assign(w->dual_start, w->dual_start + s.m);

    const auto t0 = std::chrono::steady_clock::now();
    rescale_problem(p, s);
    s.rescaling_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    s.x = s.x0;
    s.y = s.y0;
    s.xT = s.x0;
    s.xH = s.x0;
    s.yT = s.y0;
    s.yH = s.y0;
    s.l_fin.resize(s.m);
    s.u_fin.resize(s.m);
    for (int i = 0; i < s.m; ++i) {  // fill_finite_bounds_kernel, :2323-2338
        s.l_fin[i] = std::isfinite(s.l[i]) ? s.l[i] : 0.0;
        s.u_fin[i] = std::isfinite(s.u[i]) ? s.u[i] : 0.0;
    }
    // norms of the unscaled working problem, :1999-2067
    const bool linf = p->optimality_norm == NORM_TYPE_L_INF;
    double sq = 0.0, mx = 0.0;
    for (int j = 0; j < s.n; ++j) {
        const double cj = w->objective_vector[j];
        if (linf) mx = std::max(mx, std::fabs(cj));
        else sq += cj * cj;
    }
    s.c_norm = linf ? mx : std::sqrt(sq);
    sq = 0.0;
    mx = 0.0;
    for (int i = 0; i < s.m; ++i) {
        const double lo = w->constraint_lower_bound[i], hi = w->constraint_upper_bound[i];
        if (linf) {
            if (std::isfinite(lo) && lo != hi) mx = std::max(mx, std::fabs(lo));
            if (std::isfinite(hi)) mx = std::max(mx, std::fabs(hi));
        } else {
            if (std::isfinite(lo) && lo != hi) sq += lo * lo;
            if (std::isfinite(hi)) sq += hi * hi;
        }
    }
    s.b_norm = linf ? mx : std::sqrt(sq);
}

/* ---------------- positive-alpha controllers (solver.cu:187-1057) ---------------- */

double metric_of(double v) { return std::fmax(std::fabs(v), 1e-300); }  // alpha_trial_metric, :409-412

double log_rate(double start, double end, int iters)  // alpha_trial_log_rate, :414-422
{
    if (iters <= 0 || !std::isfinite(start) || !std::isfinite(end)) return NAN;
    const double r = std::log(metric_of(start) / metric_of(end)) / (double)iters;
    return std::isfinite(r) ? r : NAN;
}

double predict_ratio(double rate, int iters)  // alpha_trial_predict_ratio, :424-428
{
    return std::exp(std::fmax(-700.0, std::fmin(700.0, -rate * (double)iters)));
}

struct ZeroEpoch {  // alpha_zero_epoch_record_t, :198-208
    bool valid = false;
    int iterations = 0;
    double seconds = 0, fp_rate = 0, p_rate = 0, d_rate = 0, g_rate = 0, kkt_rate = 0;
};

struct AlphaTrial {  // alpha_trial_controller_t, :210-241; the snapshot is a copy of the whole state
    bool enabled = false, trace = false, force_reject = false, active = false, epoch_valid = false;
    int started = 0, accepted = 0, rejected = 0, positive_epochs = 0, positive_charged = 0, count = 0;
    ZeroEpoch hist[2];
    int ep_total = 0;
    double ep_time = 0, ep_p = 0, ep_d = 0, ep_g = 0;
    int t_total = 0, t_charged = 0, max_iters = 0;
    double t_time = 0, t_fp = 0, t_p = 0, t_d = 0, t_g = 0, max_sec = 0;
    bool snap_valid = false;
    State snap;
};

enum { PULSE_IDLE = 0, PULSE_POSITIVE = 1, PULSE_DONE = 2 };

struct DirectPulse {  // alpha_direct_pulse_controller_t, :250-274
    bool enabled = false, trace = false, force = false;
    int phase = PULSE_IDLE, started = 0, completed = 0, aborted = 0, gate_eval = 0, gate_pass = 0, gate_rej = 0;
    int pulse_iters = 200, gate_anchor = 0, gate_prev_k = 0, gate_recent_k = 0, anchor_traj = 0, anchor_charged = 0;
    double gate_dfp = 0, pos_fp = 0, pos_p = 0, pos_d = 0, pos_g = 0;
    termination_reason_t pos_reason = TERMINATION_REASON_UNSPECIFIED;
};

double snapshot_restore(AlphaTrial &t, State &s)  // alpha_trial_snapshot_restore, :502-585
{   // every scalar and the ten iterate vectors of the snapshot; the elapsed time stays charged.
    // Returns the largest 2-norm change of the ten vectors.
    if (!t.snap_valid) return INFINITY;
    auto l2 = [](const std::vector<double> &a, const std::vector<double> &b) {
        double q = 0.0;
        for (size_t k = 0; k < a.size(); ++k) q += (a[k] - b[k]) * (a[k] - b[k]);
        return std::sqrt(q);
    };
    const std::vector<double> *now[10] = {&s.x0, &s.x, &s.xT, &s.xH, &s.xR, &s.y0, &s.y, &s.yT, &s.yH, &s.yR};
    const std::vector<double> *old[10] = {&t.snap.x0, &t.snap.x, &t.snap.xT, &t.snap.xH, &t.snap.xR,
                                          &t.snap.y0, &t.snap.y, &t.snap.yT, &t.snap.yH, &t.snap.yR};
    double diff = 0.0;
    for (int k = 0; k < 10; ++k) diff = std::fmax(diff, l2(*old[k], *now[k]));
    const double elapsed = s.solve_time;
    s = t.snap;
    s.solve_time = elapsed;
    return diff;
}

void trial_init(AlphaTrial &t)  // alpha_trial_controller_initialize, :587-603
{
    t.enabled = env_int("CP_AL_ALPHA_ROLLBACK_TRIAL", 0) != 0;
    t.trace = env_int("CP_AL_ALPHA_ROLLBACK_TRACE", 0) != 0;
    t.force_reject = env_int("CP_AL_ALPHA_ROLLBACK_FORCE_REJECT", 0) != 0;
    if (t.enabled)
        fprintf(stderr,
                "[alpha-trial-config] enabled=1 alpha=2.000000000000e-03 history=2 fp_gain=0.10 kkt_gain=0.05 "
                "component_slack=0.05 force_reject=%d\n",
                t.force_reject ? 1 : 0);
}

void set_epoch_start(AlphaTrial &t, const State &s)  // alpha_trial_set_epoch_start, :605-614
{
    t.epoch_valid = true;
    t.ep_total = s.total_count;
    t.ep_time = s.solve_time;
    t.ep_p = metric_of(s.rel_p);
    t.ep_d = metric_of(s.rel_d);
    t.ep_g = metric_of(s.rel_gap);
}

void record_zero_epoch(AlphaTrial &t, const State &s)  // alpha_trial_record_zero_epoch, :616-678
{
    if (!t.epoch_valid || t.active || s.al_lambda > 0.0) return;
    ZeroEpoch r;
    r.iterations = s.total_count - t.ep_total;
    r.seconds = s.solve_time - t.ep_time;
    const double k0 = std::fmax(t.ep_p, std::fmax(t.ep_d, t.ep_g));
    const double k1 = std::fmax(metric_of(s.rel_p), std::fmax(metric_of(s.rel_d), metric_of(s.rel_gap)));
    r.fp_rate = log_rate(s.fp0, s.fp, r.iterations);
    r.p_rate = log_rate(t.ep_p, s.rel_p, r.iterations);
    r.d_rate = log_rate(t.ep_d, s.rel_d, r.iterations);
    r.g_rate = log_rate(t.ep_g, s.rel_gap, r.iterations);
    r.kkt_rate = log_rate(k0, k1, r.iterations);
    r.valid = r.iterations > 0 && r.seconds >= 0.0 && std::isfinite(r.fp_rate) && std::isfinite(r.p_rate) &&
              std::isfinite(r.d_rate) && std::isfinite(r.g_rate) && std::isfinite(r.kkt_rate);
    if (!r.valid) return;
    if (t.count < 2) {
        t.hist[t.count++] = r;
    } else {
        t.hist[0] = t.hist[1];
        t.hist[1] = r;
    }
    if (t.trace)
        fprintf(stderr,
                "[alpha-zero-epoch] iter=%d K=%d sec=%.12e rates(fp=%.12e p=%.12e d=%.12e g=%.12e max=%.12e) "
                "history=%d\n",
                s.total_count, r.iterations, r.seconds, r.fp_rate, r.p_rate, r.d_rate, r.g_rate, r.kkt_rate, t.count);
}

bool trial_eligible(const AlphaTrial &t, const State &s, const cp_al_parameters_t *p, int charged)  // :680-703
{
    // The 2 s floor on the elapsed time is the reference's; PDAL_ALPHA_TRIAL_MIN_SEC lowers
    // it so that a test reaches the trial at a deterministic point.
    const double min_sec = env_double("PDAL_ALPHA_TRIAL_MIN_SEC", 2.0);
    if (!t.enabled || t.started != 0 || t.count < 2 || charged < 10000 || s.solve_time < min_sec ||
        s.reason != TERMINATION_REASON_UNSPECIFIED)
        return false;
    const int hmax = std::max(t.hist[0].iterations, t.hist[1].iterations);
    const double hsec = std::fmax(2.0, std::fmax(t.hist[0].seconds, t.hist[1].seconds));
    const int rem_it = p->termination_criteria.iteration_limit - charged;
    const double rem_sec = p->termination_criteria.time_sec_limit - s.solve_time;
    return hmax > 0 && rem_it >= 3 * hmax && rem_sec >= 3.0 * hsec;
}

void trial_start(AlphaTrial &t, State &s, int charged)  // alpha_trial_start, :704-761
{
    t.snap = s;
    t.snap_valid = true;
    t.started = 1;
    t.active = true;
    t.t_total = s.total_count;
    t.t_charged = charged;
    t.t_time = s.solve_time;
    t.t_p = metric_of(s.rel_p);
    t.t_d = metric_of(s.rel_d);
    t.t_g = metric_of(s.rel_gap);
    const int hmax = std::max(t.hist[0].iterations, t.hist[1].iterations);
    t.max_iters = 2 * hmax;
    t.max_sec = 2.0 * std::fmax(2.0, std::fmax(t.hist[0].seconds, t.hist[1].seconds));
    s.sigma_mode = false;
    s.adaptive_lambda = false;
    s.al_lambda = 0.002;
    refresh_step_schedule(s);
    compute_fixed_point_error(s);
    t.t_fp = metric_of(s.fp);
    if (!std::isfinite(t.t_fp) || t.t_fp <= 0.0) {
        t.active = false;
        t.rejected = 1;
        const double d = snapshot_restore(t, s);
        fprintf(stderr, "[alpha-trial-restore] reason=invalid-start charged_iter=%d trajectory_iter=%d "
                        "max_vector_l2=%.12e\n", charged, s.total_count, d);
        t.snap_valid = false;
        return;
    }
    fprintf(stderr,
            "[alpha-trial-start] charged_iter=%d trajectory_iter=%d solve_sec=%.12e alpha=2.000000000000e-03 "
            "fp=%.12e primal=%.12e dual=%.12e gap=%.12e max_K=%d max_sec=%.12e\n",
            charged, s.total_count, s.solve_time, t.t_fp, t.t_p, t.t_d, t.t_g, t.max_iters, t.max_sec);
}

bool trial_cap(const AlphaTrial &t, const State &s, int charged)  // alpha_trial_safety_cap_reached, :763-775
{
    if (!t.active) return false;
    return charged - t.t_charged >= t.max_iters || s.solve_time - t.t_time >= t.max_sec;
}

bool trial_accepts(const AlphaTrial &t, const State &s, int charged, const char **why)  // :777-861
{
    const int k = s.total_count - t.t_total;
    if (k <= 0) {
        *why = "empty-epoch";
        return false;
    }
    const double fr = 0.5 * (t.hist[0].fp_rate + t.hist[1].fp_rate);
    const double pr = 0.5 * (t.hist[0].p_rate + t.hist[1].p_rate);
    const double dr = 0.5 * (t.hist[0].d_rate + t.hist[1].d_rate);
    const double gr = 0.5 * (t.hist[0].g_rate + t.hist[1].g_rate);
    const double kr = 0.5 * (t.hist[0].kkt_rate + t.hist[1].kkt_rate);
    const double fp_pred = predict_ratio(fr, k);
    const double p_pred = t.t_p * predict_ratio(pr, k);
    const double d_pred = t.t_d * predict_ratio(dr, k);
    const double g_pred = t.t_g * predict_ratio(gr, k);
    const double k_pred = std::fmax(t.t_p, std::fmax(t.t_d, t.t_g)) * predict_ratio(kr, k);
    const double fp_obs = metric_of(s.fp) / t.t_fp;
    const double p_obs = metric_of(s.rel_p), d_obs = metric_of(s.rel_d), g_obs = metric_of(s.rel_gap);
    const double k_obs = std::fmax(p_obs, std::fmax(d_obs, g_obs));
    const bool finite = std::isfinite(fp_pred) && std::isfinite(p_pred) && std::isfinite(d_pred) &&
                        std::isfinite(g_pred) && std::isfinite(k_pred) && std::isfinite(fp_obs) &&
                        std::isfinite(p_obs) && std::isfinite(d_obs) && std::isfinite(g_obs) && std::isfinite(k_obs);
    bool ok = finite && fp_obs <= 0.90 * fp_pred && k_obs <= 0.95 * k_pred && p_obs <= 1.05 * p_pred &&
              d_obs <= 1.05 * d_pred && g_obs <= 1.05 * g_pred;
    if (t.force_reject) {
        ok = false;
        *why = "forced-reject";
    } else {
        *why = !finite ? "nonfinite" : !ok ? "acceptance-gate" : "accepted";
    }
    fprintf(stderr,
            "[alpha-trial-decision] reason=%s accept=%d charged_iter=%d trajectory_iter=%d K=%d "
            "fp_ratio_obs=%.12e fp_ratio_pred=%.12e kkt_obs=%.12e kkt_pred=%.12e p_obs=%.12e p_pred=%.12e "
            "d_obs=%.12e d_pred=%.12e g_obs=%.12e g_pred=%.12e\n",
            *why, ok ? 1 : 0, charged, s.total_count, k, fp_obs, fp_pred, k_obs, k_pred, p_obs, p_pred, d_obs, d_pred,
            g_obs, g_pred);
    return ok;
}

void pulse_init(DirectPulse &d)  // alpha_direct_pulse_initialize, :884-904
{
    d.enabled = env_int("CP_AL_ALPHA_DIRECT_PULSE", 0) != 0;
    d.trace = env_int("CP_AL_ALPHA_DIRECT_PULSE_TRACE", 0) != 0;
    d.force = env_int("CP_AL_DETERMINISTIC_PULSE_FORCE", 0) != 0;
    if (d.enabled)
        fprintf(stderr,
                "[alpha-direct-pulse-config] enabled=1 alpha=2.000000000000e-03 pulse_iterations=200 "
                "history_epochs=2 deterministic_anchor_iter_ge=10000 wall_time_features=0 tail_only=1 "
                "force_pulse=%d\n",
                d.force ? 1 : 0);
}

bool pulse_eligible(DirectPulse &d, const AlphaTrial &h, const State &s, const cp_al_parameters_t *p, int charged,
                    int K)  // alpha_direct_pulse_is_eligible, :906-958
{
    if (!d.enabled || d.started != 0 || d.gate_eval != 0 || d.phase != PULSE_IDLE || h.count < 2 ||
        charged < 10000 || s.reason != TERMINATION_REASON_UNSPECIFIED || K != 200)
        return false;
    for (int i = 0; i < 2; ++i)
        if (!h.hist[i].valid || h.hist[i].iterations <= 0 || !std::isfinite(h.hist[i].fp_rate)) return false;
    if (p->termination_criteria.iteration_limit - charged < 400) return false;
    d.gate_eval = 1;
    d.gate_anchor = charged;
    d.gate_prev_k = h.hist[0].iterations;
    d.gate_recent_k = h.hist[1].iterations;
    d.gate_dfp = h.hist[1].fp_rate - h.hist[0].fp_rate;
    const bool pass = d.force;
    d.gate_pass = pass ? 1 : 0;
    d.gate_rej = pass ? 0 : 1;
    fprintf(stderr,
            "[alpha-deterministic-tail-anchor] charged_iter=%d trajectory_iter=%d previous_K=%d recent_K=%d "
            "previous_fp_rate=%.12e recent_fp_rate=%.12e delta_fp_rate=%.12e pass=%d force_pulse=%d\n",
            charged, s.total_count, h.hist[0].iterations, h.hist[1].iterations, h.hist[0].fp_rate,
            h.hist[1].fp_rate, d.gate_dfp, pass ? 1 : 0, d.force ? 1 : 0);
    return pass;
}

void pulse_start(DirectPulse &d, State &s, int charged, int K)  // alpha_direct_pulse_start, :960-982
{
    d.started = 1;
    d.phase = PULSE_POSITIVE;
    d.pulse_iters = K;
    d.anchor_traj = s.total_count;
    d.anchor_charged = charged;
    s.sigma_mode = false;
    s.adaptive_lambda = false;
    s.al_lambda = 0.002;
    refresh_step_schedule(s);
    fprintf(stderr, "[alpha-direct-pulse-start] charged_iter=%d trajectory_iter=%d alpha=2.000000000000e-03 "
                    "pulse_iterations=%d\n", charged, s.total_count, d.pulse_iters);
}

void pulse_finish(DirectPulse &d, State &s, int charged)  // alpha_direct_pulse_finish_positive, :984-1025
{
    if (!d.enabled || d.phase != PULSE_POSITIVE) return;
    d.pos_p = metric_of(s.rel_p);
    d.pos_d = metric_of(s.rel_d);
    d.pos_g = metric_of(s.rel_gap);
    d.pos_reason = s.reason;
    s.reason = TERMINATION_REASON_UNSPECIFIED;
    s.sigma_mode = false;
    s.adaptive_lambda = false;
    s.al_lambda = 0.0;
    refresh_step_schedule(s);
    compute_fixed_point_error(s);
    d.pos_fp = metric_of(s.fp);
    choose_restart_candidate(s);
    s.reason = d.pos_reason;
    d.completed = 1;
    d.aborted = d.pos_reason != TERMINATION_REASON_UNSPECIFIED && d.pos_reason != TERMINATION_REASON_OPTIMAL;
    d.phase = PULSE_DONE;
    fprintf(stderr,
            "[alpha-direct-pulse-complete] charged_iter=%d trajectory_iter=%d positive_iters=%d fp0map=%.12e "
            "primal=%.12e dual=%.12e gap=%.12e term=%d\n",
            charged, s.total_count, d.pulse_iters, d.pos_fp, d.pos_p, d.pos_d, d.pos_g, (int)d.pos_reason);
}

void alpha_final_logs(const AlphaTrial &t, const DirectPulse &d, const State &s, int charged)  // :863-882, :1027-1057
{
    if (t.enabled)
        fprintf(stderr,
                "[alpha-trial-final] charged_iter=%d trajectory_iter=%d started=%d accepted=%d rejected=%d "
                "positive_epochs=%d positive_charged_iters=%d final_alpha=%.12e final_penalty=%.12e\n",
                charged, s.total_count, t.started, t.accepted, t.rejected, t.positive_epochs, t.positive_charged,
                s.al_lambda, s.al_penalty);
    if (d.enabled)
        fprintf(stderr,
                "[alpha-direct-pulse-final] charged_iter=%d trajectory_iter=%d started=%d completed=%d aborted=%d "
                "gate_evaluated=%d gate_passed=%d gate_rejected=%d gate_anchor_charged_iter=%d gate_previous_K=%d "
                "gate_recent_K=%d gate_delta_fp_rate=%.12e positive_charged_iters=%d final_alpha=%.12e "
                "final_penalty=%.12e\n",
                charged, s.total_count, d.started, d.completed, d.aborted, d.gate_eval, d.gate_pass, d.gate_rej,
                d.gate_anchor, d.gate_prev_k, d.gate_recent_k, d.gate_dfp, d.completed ? d.pulse_iters : 0,
                s.al_lambda, s.al_penalty);
}

}  // namespace

/* ---------------- public C API (cp_al.c, utils.cu conversions) ---------------- */

ORACLE_API void cp_al_oracle_reset_rng(void)
{
    g_gen.seed(1);
    g_dist.reset();
}

ORACLE_API void set_default_parameters(cp_al_parameters_t *p)  // utils.cu:361-394
{
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

static void fill_or_copy(double **dst, int n, const double *src, double fill)  // utils.cu:1032-1040
{
    *dst = (double *)malloc((size_t)std::max(n, 1) * sizeof(double));
    if (src) memcpy(*dst, src, (size_t)n * sizeof(double));
    else
        for (int i = 0; i < n; ++i) (*dst)[i] = fill;
}

ORACLE_API lp_problem_t *create_lp_problem(const double *c, const matrix_desc_t *A, const double *con_lb,
                                           const double *con_ub, const double *var_lb, const double *var_ub,
                                           const double *obj_const)
{   // cp_al.c:26-116, conversions utils.cu:1042-1192
    if (!A) return nullptr;
    lp_problem_t *p = (lp_problem_t *)calloc(1, sizeof(lp_problem_t));
    const int m = A->m, n = A->n;
    p->num_variables = n;
    p->num_constraints = m;
    int *rp = (int *)calloc((size_t)m + 1, sizeof(int));
    int *ci = nullptr;
    double *v = nullptr;
    int nnz = 0;
    switch (A->fmt) {
    case matrix_dense: {
        for (long i = 0; i < (long)m * n; ++i)
            if (std::fabs(A->data.dense.A[i]) > 0.0) ++nnz;
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
        ci = (int *)malloc((size_t)std::max(nnz, 1) * sizeof(int));
        v = (double *)malloc((size_t)std::max(nnz, 1) * sizeof(double));
        memcpy(rp, A->data.csr.row_ptr, (size_t)(m + 1) * sizeof(int));
        memcpy(ci, A->data.csr.col_ind, (size_t)nnz * sizeof(int));
        memcpy(v, A->data.csr.vals, (size_t)nnz * sizeof(double));
        break;
    case matrix_csc:
    case matrix_coo: {
        const bool csc = A->fmt == matrix_csc;
        nnz = csc ? A->data.csc.nnz : A->data.coo.nnz;
        // row index of each stored entry, then a stable counting sort by row
        std::vector<int> rows(nnz), cols(nnz);
        if (csc) {
            for (int j = 0; j < n; ++j)
                for (int k = A->data.csc.col_ptr[j]; k < A->data.csc.col_ptr[j + 1]; ++k) {
                    rows[k] = A->data.csc.row_ind[k];
                    cols[k] = j;
                }
        } else {
            for (int k = 0; k < nnz; ++k) {
                rows[k] = A->data.coo.row_ind[k];
                cols[k] = A->data.coo.col_ind[k];
            }
        }
        for (int k = 0; k < nnz; ++k)
            if (rows[k] < 0 || rows[k] >= m || cols[k] < 0 || cols[k] >= n) {
                fprintf(stderr, "[interface] %s: index out of range\n", csc ? "CSC" : "COO");
                free(rp);
                free(p);
                return nullptr;
            }
        const double *vals = csc ? A->data.csc.vals : A->data.coo.vals;
        for (int k = 0; k < nnz; ++k) rp[rows[k] + 1]++;
        for (int i = 0; i < m; ++i) rp[i + 1] += rp[i];
        ci = (int *)malloc((size_t)std::max(nnz, 1) * sizeof(int));
        v = (double *)malloc((size_t)std::max(nnz, 1) * sizeof(double));
        std::vector<int> next(rp, rp + m);
        for (int k = 0; k < nnz; ++k) {
            const int pos = next[rows[k]]++;
            ci[pos] = cols[k];
            v[pos] = vals[k];
        }
        break;
    }
    default:
        fprintf(stderr, "[interface] unsupported matrix format %d.\n", (int)A->fmt);
        free(rp);
        free(p);
        return nullptr;
    }
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

ORACLE_API void set_start_values(lp_problem_t *p, const double *primal, const double *dual)  // cp_al.c:146-178
{
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

// utils.cu:230-253; the LP runner prints it (an oracle-backed runner links this library)
ORACLE_API const char *termination_reason_to_string(termination_reason_t r) { return reason_name(r); }

ORACLE_API void cp_al_result_free(cp_al_result_t *r)
{
    if (!r) return;
    free(r->primal_solution);
    free(r->dual_solution);
    free(r->reduced_cost);
    free(r);
}

ORACLE_API void lp_problem_free(lp_problem_t *p)
{
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
    free(p);
}

void initialize_sigma_parameters(State &s)  // solver.cu:3141-3164
{
    const bool sigma_env = getenv("CP_AL_SIGMA") && getenv("CP_AL_SIGMA")[0] != '\0';
    s.lambda_param = env_double("CP_AL_LAMBDA_PARAM",
                                env_double("CP_AL_LAMBDA_FACTOR", 2.01) * std::fmax(s.lambda_est, 1e-12));
    s.lambda_param = std::fmax(s.lambda_param, (2.0 + 1e-12) * std::fmax(s.lambda_est, 1e-12));
    s.sigma_base = 1.0 / std::sqrt(std::fmax(s.lambda_param, 1e-16));
    s.sigma_min = env_double("CP_AL_SIGMA_MIN_FACTOR", 1e-4) * s.sigma_base;
    s.sigma_max = env_double("CP_AL_SIGMA_MAX_FACTOR", 100.0) * s.sigma_base;
    s.al_sigma = env_double("CP_AL_SIGMA", env_double("CP_AL_SIGMA_INIT_FACTOR", 6.0) * s.sigma_base);
    s.al_sigma = std::fmin(std::fmax(s.al_sigma, s.sigma_min), s.sigma_max);
    s.sigma_best = s.al_sigma;
    s.sigma_err_sum = 0.0;
    s.sigma_last_err = 0.0;
    s.sigma_adaptive = env_int("CP_AL_SIGMA_ADAPTIVE", sigma_env ? 0 : 1) != 0;
    s.sigma_mode = true;
}

ORACLE_API cp_al_result_t *optimize(const cp_al_parameters_t *params, lp_problem_t *prob)
{   // solver.cu:1175-1736
    if (!params || !prob) return nullptr;
    if (params->termination_evaluation_frequency < 3) {
        fprintf(stderr, "Error: termination_evaluation_frequency must be >= 3 (got %d).\n",
                params->termination_evaluation_frequency);
        return nullptr;
    }
    if (params->verbose) filter_matrix(prob, params);

    State s;
    init_state(s, prob, params);

    // initialize_step_size_and_primal_weight, solver.cu:3538-3609
    s.lambda_est = 1.0;
    if (s.A.rp[s.m] != 0) {
        const double sv = estimate_max_singular_value(s, params->sv_max_iter, params->sv_tol);
        s.lambda_est = sv * sv;
    }
    s.primal_weight = params->bound_objective_rescaling ? 1.0 : (s.c_norm + 1.0) / (s.b_norm + 1.0);
    const bool sigma_env = getenv("CP_AL_SIGMA") && getenv("CP_AL_SIGMA")[0] != '\0';
    s.sigma_mode = sigma_env || env_int("CP_AL_SIGMA_MODE", 0) != 0;
    if (s.sigma_mode) initialize_sigma_parameters(s);
    s.al_lambda = env_double("CP_AL_LAMBDA", 0.05);
    initialize_adaptive_lambda(s);
    if (s.sigma_mode) s.adaptive_lambda = false;
    refresh_step_schedule(s);
    s.best_pw = s.primal_weight;
    if (params->verbose) {  // solver.cu:3587-3608
        fprintf(stderr,
                "[box-cpal] pw_init=%.3e al_lambda=%.3e sigma_mode=%d sigma_adapt=%d al_sigma=%.3e sigma_base=%.3e "
                "lambda_param=%.3e eta=%.3e refl=%.3e al_penalty=%.3e lambdaA=%.3e adaptive=%d lambda_grid=",
                s.primal_weight, s.al_lambda, (int)s.sigma_mode, (int)s.sigma_adaptive, s.al_sigma, s.sigma_base,
                s.lambda_param, s.step_size, params->reflection_coefficient, s.al_penalty, s.lambda_est,
                (int)s.adaptive_lambda);
        for (int i = 0; i < s.lambda_num_tiers; ++i) fprintf(stderr, "%s%.3e", i == 0 ? "" : "/", s.lambda_values[i]);
        fprintf(stderr, "\n");
    }

    AlphaTrial trial;
    trial_init(trial);
    DirectPulse pulse;
    pulse_init(pulse);
    if (trial.enabled && pulse.enabled) {  // solver.cu:1216-1222 (an error return instead of exit)
        fprintf(stderr, "[alpha-controller-error] rollback trial and direct pulse cannot be enabled together\n");
        return nullptr;
    }
    const int trial_test_limit = env_int("CP_AL_ALPHA_ROLLBACK_TEST_TRAJECTORY_LIMIT", 0);
    const int pulse_test_limit = env_int("CP_AL_ALPHA_DIRECT_PULSE_TEST_TRAJECTORY_LIMIT", 0);

    const auto t_start = std::chrono::steady_clock::now();
    int charged = s.total_count;
    bool do_restart = false;
    const double rho = params->reflection_coefficient;
    const auto &tc = params->termination_criteria;
    // dynamic evaluation (solver.cu:1228-1260)
    const int base_eval_frequency = params->termination_evaluation_frequency;
    int eval_frequency = base_eval_frequency;
    int dynamic_eval_fast = env_int("CP_AL_DYNAMIC_EVAL_FAST", 50);
    dynamic_eval_fast = dynamic_eval_fast < 3 ? 3 : dynamic_eval_fast;
    dynamic_eval_fast = dynamic_eval_fast > eval_frequency ? eval_frequency : dynamic_eval_fast;
    const bool use_residual_dynamic_eval = env_int("CP_AL_DYNAMIC_EVAL", 0) != 0;
    const double dynamic_eval_after_sec = env_double("CP_AL_DYNAMIC_EVAL_AFTER_SEC", -1.0);
    const bool use_time_dynamic_eval = dynamic_eval_after_sec > 0.0;
    const int dynamic_eval_after_iters = env_int("CP_AL_DYNAMIC_EVAL_AFTER_ITERS", -1);
    const bool use_iteration_dynamic_eval = dynamic_eval_after_iters > 0;
    const bool use_dynamic_eval = (use_residual_dynamic_eval || use_time_dynamic_eval || use_iteration_dynamic_eval) &&
                                  dynamic_eval_fast < eval_frequency;
    const int dynamic_eval_min_iters =
        std::max(env_int("CP_AL_DYNAMIC_EVAL_MIN_ITERS", eval_frequency), eval_frequency);
    const double dynamic_eval_small_res = env_double("CP_AL_DYNAMIC_EVAL_SMALL_RES", 1e-6);
    const double dynamic_eval_large_res = env_double("CP_AL_DYNAMIC_EVAL_LARGE_RES", 1e-4);
    const double dynamic_eval_ratio = env_double("CP_AL_DYNAMIC_EVAL_RATIO", 1e4);
    const int dynamic_eval_bootstrap_iters = std::max(env_int("CP_AL_DYNAMIC_EVAL_BOOTSTRAP_ITERS", 0), 0);
    bool dynamic_eval_latched = false;
    if (use_dynamic_eval && dynamic_eval_bootstrap_iters > 0) eval_frequency = dynamic_eval_fast;
    auto log_freq = [](const char *name) {  // get_trace_log_frequency / get_progress_log_frequency, :321-375
        const char *raw = getenv(name);
        const int v = (raw && raw[0]) ? atoi(raw) : 0;
        return v > 0 ? v : 0;
    };
    const int trace_freq = params->verbose ? log_freq("CP_AL_LOG_FREQ") : 0;
    const int progress_freq = params->verbose ? log_freq("CP_AL_PROGRESS_LOG_FREQ") : 0;
    while (charged < tc.iteration_limit) {
        const int K = eval_frequency;
        if (do_restart) {
            compute_fixed_point_error(s);
            s.fp0 = s.fp;
            do_restart = false;
        }
        for (int i = 1; i <= K; ++i) {
            iteration(s, i, rho, i == 1 || i == K);
            const long long it = (long long)s.total_count + i;
            if (trace_freq > 0 && it % trace_freq == 0 && it % K != 0) {  // should_emit_trace, emit_trace_fp
                compute_fixed_point_error(s);
                if (std::isfinite(s.fp) && s.fp > 0.0) printf("trace iter %lld fp %.12e\n", it, s.fp);
            }
        }
        compute_fixed_point_error(s);
        choose_restart_candidate(s);  // solver.cu:1351
        compute_residual(s, params->optimality_norm);
        s.inner_count += K;
        s.total_count += K;
        charged += K;
        if (params->verbose) {  // emit_checkpoint_fp, emit_progress_metrics (solver.cu:1356-1363, 329-375)
            if (std::isfinite(s.fp) && s.fp > 0.0) printf("checkpoint iter %lld fp %.12e\n", (long long)s.total_count, s.fp);
            if (progress_freq > 0 && s.total_count % progress_freq == 0) {
                const double bal = std::fabs(std::log10(std::fmax(s.rel_d, 1e-300) / std::fmax(s.rel_p, 1e-300)));
                printf("progress iter %lld time %.6e rpr %.12e rdr %.12e rgap %.12e fp %.12e pw %.12e pw_best %.12e "
                       "bal %.12e\n",
                       (long long)s.total_count, s.solve_time, s.rel_p, s.rel_d, s.rel_gap,
                       (std::isfinite(s.fp) && s.fp > 0.0) ? s.fp : NAN, s.primal_weight, s.best_pw, bal);
            }
        }

        // check_termination_criteria, utils.cu:299-317, and the controllers' test limits (:1372-1391)
        s.solve_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        if (s.rel_d < tc.eps_feasible_relative && s.rel_p < tc.eps_feasible_relative &&
            s.rel_gap < tc.eps_optimal_relative)
            s.reason = TERMINATION_REASON_OPTIMAL;
        else if (s.total_count >= tc.iteration_limit)
            s.reason = TERMINATION_REASON_ITERATION_LIMIT;
        else if (s.solve_time >= tc.time_sec_limit)
            s.reason = TERMINATION_REASON_TIME_LIMIT;
        if (s.reason == TERMINATION_REASON_UNSPECIFIED && charged >= tc.iteration_limit)
            s.reason = TERMINATION_REASON_ITERATION_LIMIT;
        if (s.reason == TERMINATION_REASON_UNSPECIFIED && trial.enabled && trial_test_limit > 0 && !trial.active &&
            trial.rejected > 0 && s.total_count >= trial_test_limit)
            s.reason = TERMINATION_REASON_ITERATION_LIMIT;
        if (s.reason == TERMINATION_REASON_UNSPECIFIED && pulse.enabled && pulse_test_limit > 0 &&
            pulse.phase == PULSE_DONE && s.total_count >= pulse_test_limit)
            s.reason = TERMINATION_REASON_ITERATION_LIMIT;
        if (pulse.phase == PULSE_POSITIVE) {  // :1416-1425
            pulse_finish(pulse, s, charged);
            if (s.reason == TERMINATION_REASON_UNSPECIFIED && pulse_test_limit > 0 &&
                s.total_count >= pulse_test_limit)
                s.reason = TERMINATION_REASON_ITERATION_LIMIT;
        }
        if (s.reason != TERMINATION_REASON_UNSPECIFIED) {  // :1426-1474
            if (trial.active) {
                const int pos = charged - trial.t_charged;
                trial.positive_charged += pos;
                trial.positive_epochs += pos > 0 ? 1 : 0;
                if (s.reason == TERMINATION_REASON_OPTIMAL) {
                    trial.accepted += 1;
                    trial.active = false;
                    s.al_lambda = 0.0;
                    refresh_step_schedule(s);
                    fprintf(stderr, "[alpha-trial-decision] reason=terminal-optimal accept=1 charged_iter=%d "
                                    "trajectory_iter=%d K=%d\n", charged, s.total_count, s.total_count - trial.t_total);
                } else {
                    const termination_reason_t why = s.reason;
                    const double d = snapshot_restore(trial, s);
                    s.reason = why;
                    trial.rejected += 1;
                    trial.active = false;
                    fprintf(stderr, "[alpha-trial-restore] reason=charged-termination charged_iter=%d "
                                    "trajectory_iter=%d max_vector_l2=%.12e\n", charged, s.total_count, d);
                }
                trial.snap_valid = false;
            }
            break;
        }
        if (trial_cap(trial, s, charged)) {  // :1476-1494
            const int pos = charged - trial.t_charged;
            trial.positive_charged += pos;
            trial.positive_epochs += pos > 0 ? 1 : 0;
            const double d = snapshot_restore(trial, s);
            trial.rejected += 1;
            trial.active = false;
            set_epoch_start(trial, s);
            fprintf(stderr, "[alpha-trial-restore] reason=safety-cap charged_iter=%d trajectory_iter=%d "
                            "positive_iters=%d max_vector_l2=%.12e\n", charged, s.total_count, pos, d);
            trial.snap_valid = false;
            do_restart = true;
            continue;
        }

        const bool sigma_trigger_enabled = !s.sigma_mode && env_int("CP_AL_SIGMA_TRIGGER", 0) != 0;
        bool sigma_trigger_restart = false;
        if (sigma_trigger_enabled && s.total_count >= env_int("CP_AL_SIGMA_TRIGGER_MIN_ITERS", 200)) {  // :1496-1528
            const double rpr = std::fmax(s.rel_p, 1e-300);
            const double rdr = std::fmax(s.rel_d, 1e-300);
            const double trigger_primal = env_double("CP_AL_SIGMA_TRIGGER_PRIMAL", 1e-2);
            const double trigger_dual = env_double("CP_AL_SIGMA_TRIGGER_DUAL", 1e-6);
            const double trigger_ratio = env_double("CP_AL_SIGMA_TRIGGER_RATIO", 1e6);
            if (rpr >= trigger_primal && rdr <= trigger_dual && rpr / rdr >= trigger_ratio) {
                initialize_sigma_parameters(s);
                s.adaptive_lambda = false;
                sigma_trigger_restart = true;
                if (params->verbose)
                    printf("sigma-trigger iter %d rpr %.12e rdr %.12e ratio %.12e sigma %.12e base %.12e\n",
                           s.total_count, s.rel_p, s.rel_d, rpr / rdr, s.al_sigma, s.sigma_base);
            }
        }
        if (use_dynamic_eval && !s.sigma_mode &&
            (s.total_count >= dynamic_eval_min_iters ||
             (dynamic_eval_bootstrap_iters > 0 && s.total_count >= dynamic_eval_bootstrap_iters))) {  // :1533-1584
            const double rpr = std::fmax(s.rel_p, 1e-300);
            const double rdr = std::fmax(s.rel_d, 1e-300);
            const double min_res = std::fmin(rpr, rdr);
            const double max_res = std::fmax(rpr, rdr);
            const double residual_ratio = max_res / min_res;
            const bool one_sided_residual = min_res <= dynamic_eval_small_res && max_res >= dynamic_eval_large_res;
            const bool ratio_lag = residual_ratio >= dynamic_eval_ratio && max_res >= dynamic_eval_large_res;
            const double now = controller_seconds(s);
            const bool time_gate = use_time_dynamic_eval && !s.cr_early_stopped && now >= dynamic_eval_after_sec;
            const bool iteration_gate =
                use_iteration_dynamic_eval && !s.cr_early_stopped && s.total_count >= dynamic_eval_after_iters;
            if (time_gate || iteration_gate || (use_residual_dynamic_eval && (one_sided_residual || ratio_lag))) {
                dynamic_eval_latched = true;
                if (eval_frequency != dynamic_eval_fast) {
                    eval_frequency = dynamic_eval_fast;
                    if (params->verbose || use_time_dynamic_eval || use_iteration_dynamic_eval)
                        printf("[dynamic-eval] iter=%d solve_sec=%.12e eval_frequency=%d->%d relative_primal=%.12e "
                               "relative_dual=%.12e residual_ratio=%.12e reason=%s\n",
                               s.total_count, now, K, eval_frequency, s.rel_p, s.rel_d, residual_ratio,
                               time_gate ? "time" : iteration_gate ? "iteration" : "residual");
                }
            } else if (!dynamic_eval_latched && dynamic_eval_bootstrap_iters > 0 &&
                       eval_frequency == dynamic_eval_fast && s.total_count >= dynamic_eval_bootstrap_iters) {
                eval_frequency = base_eval_frequency;
                if (params->verbose)
                    printf("dynamic-eval iter %d eval_freq %d->%d rpr %.12e rdr %.12e ratio %.12e\n", s.total_count,
                           K, eval_frequency, s.rel_p, s.rel_d, residual_ratio);
            }
        }
        do_restart = sigma_trigger_restart || should_restart(s, params->restart_params, K);
        if (!do_restart && s.adaptive_lambda && s.lambda_race && s.lambda_force_restart_iters > 0 &&
            s.total_count > 0 && s.total_count - s.lambda_last_restart_iter >= s.lambda_force_restart_iters) {
            do_restart = true;  // :1590-1600
            if (params->verbose) printf("lambda-race artificial restart iter %d fp %.12e\n", s.total_count, s.fp);
        }
        if (do_restart) {  // :1631-1695
            if (params->verbose) printf("restart iter %d fp %.12e\n", s.total_count, s.fp);
            if (trial.active) {
                const int pos = charged - trial.t_charged;
                trial.positive_charged += pos;
                trial.positive_epochs += pos > 0 ? 1 : 0;
                const char *why = nullptr;
                if (!trial_accepts(trial, s, charged, &why)) {
                    const double d = snapshot_restore(trial, s);
                    trial.rejected += 1;
                    trial.active = false;
                    set_epoch_start(trial, s);
                    fprintf(stderr, "[alpha-trial-restore] reason=%s charged_iter=%d trajectory_iter=%d "
                                    "positive_iters=%d max_vector_l2=%.12e\n", why, charged, s.total_count, pos, d);
                    trial.snap_valid = false;
                    do_restart = true;
                    continue;
                }
                trial.accepted += 1;
                trial.active = false;
                s.al_lambda = 0.0;
                refresh_step_schedule(s);
                trial.snap_valid = false;
            } else {
                record_zero_epoch(trial, s);
            }
            perform_restart(s, params);
            set_epoch_start(trial, s);
            if (trial_eligible(trial, s, params, charged)) trial_start(trial, s, charged);
            if (pulse_eligible(pulse, trial, s, params, charged, K)) pulse_start(pulse, s, charged, K);
        }
    }
    if (s.reason == TERMINATION_REASON_UNSPECIFIED) {
        s.reason = TERMINATION_REASON_ITERATION_LIMIT;
        compute_residual(s, params->optimality_norm);
    }
    alpha_final_logs(trial, pulse, s, charged);
    if (params->feasibility_polishing && s.reason != TERMINATION_REASON_DUAL_INFEASIBLE &&
        s.reason != TERMINATION_REASON_PRIMAL_INFEASIBLE)  // solver.cu:1715-1719
        feasibility_polish(params, s, t_start);

    // create_result_from_state, solver.cu:3828-3909
    cp_al_result_t *r = (cp_al_result_t *)calloc(1, sizeof(cp_al_result_t));
    spmv(s.At, s.yT.data(), s.dual_prod.data());
    r->primal_solution = (double *)malloc((size_t)std::max(s.n, 1) * sizeof(double));
    r->dual_solution = (double *)malloc((size_t)std::max(s.m, 1) * sizeof(double));
    r->reduced_cost = (double *)malloc((size_t)std::max(s.n, 1) * sizeof(double));
    for (int j = 0; j < s.n; ++j) {
        double rc = (s.c[j] - s.dual_prod[j]) * s.D[j] / s.c_scale;
        if (!std::isfinite(s.lb[j])) rc = std::fmin(rc, 0.0);
        if (!std::isfinite(s.ub[j])) rc = std::fmax(rc, 0.0);
        r->reduced_cost[j] = rc;
        r->primal_solution[j] = s.xT[j] / s.D[j] / s.b_scale;
    }
    for (int i = 0; i < s.m; ++i) r->dual_solution[i] = s.yT[i] / s.E[i] / s.c_scale;
    r->num_variables = prob->num_variables;
    r->num_constraints = prob->num_constraints;
    r->num_nonzeros = prob->constraint_matrix_num_nonzeros;
    r->total_count = charged;
    r->rescaling_time_sec = s.rescaling_time;
    r->cumulative_time_sec = s.solve_time;
    r->absolute_primal_residual = s.abs_p;
    r->relative_primal_residual = s.rel_p;
    r->absolute_dual_residual = s.abs_d;
    r->relative_dual_residual = s.rel_d;
    r->primal_objective_value = s.p_obj;
    r->dual_objective_value = s.d_obj;
    r->objective_gap = s.gap;
    r->relative_objective_gap = s.rel_gap;
    r->termination_reason = s.reason;
    r->feasibility_polishing_time = s.polish_time;
    r->feasibility_iteration = s.polish_iters;
    return r;
}

ORACLE_API cp_al_result_t *solve_lp_problem(lp_problem_t *prob, const cp_al_parameters_t *params)
{   // cp_al.c:180-209
    if (!prob) {
        fprintf(stderr, "[interface] solve_lp_problem: invalid arguments.\n");
        return nullptr;
    }
    cp_al_parameters_t local;
    if (params) local = *params;
    else set_default_parameters(&local);
    return optimize(&local, prob);
}

/* ---------------- test hooks (pin individual oracle functions) ---------------- */

// sigma_max(A) by the oracle's power iteration (utils.cu:87-228), generator reset first.
// CPU baseline (no counterpart in the reference): solve problems[0..count)
// with `params` on `nthreads` threads (OpenMP when built with -fopenmp; one at
// a time otherwise), each solve starting from a freshly seeded power-iteration
// generator, as a batch of independent reference solves would.  out[i] takes
// problem i's result (free with cp_al_result_free).
ORACLE_API void cp_al_oracle_solve_many(lp_problem_t **problems, int count, const cp_al_parameters_t *params,
                                        int nthreads, cp_al_result_t **out)
{
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(dynamic, 1)
#endif
    for (int i = 0; i < count; ++i) {
        g_gen.seed(1);
        g_dist.reset();
        out[i] = optimize(params, problems[i]);
    }
    (void)nthreads;
}

// Threads the parallel entry point can use (1 without OpenMP).
ORACLE_API int cp_al_oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

ORACLE_API double cp_al_oracle_sigma_max(int m, int n, const int *rp, const int *ci, const double *v, int max_iter,
                                         double tol)
{
    State s;
    s.m = m;
    s.n = n;
    s.A.rows = m;
    s.A.cols = n;
    s.A.rp.assign(rp, rp + m + 1);
    s.A.ci.assign(ci, ci + rp[m]);
    s.A.v.assign(v, v + rp[m]);
    s.At = transpose(s.A);
    cp_al_oracle_reset_rng();
    return estimate_max_singular_value(s, max_iter, tol);
}

// The oracle's rescaling (preconditioner.cu:379-475) of a problem; outputs the
// scaled matrix values (CSR order of the input), cumulative E and D, and the
// bound / objective scales.
ORACLE_API void cp_al_oracle_rescale(const lp_problem_t *w, const cp_al_parameters_t *p, double *vals_out,
                                     double *E_out, double *D_out, double *scales_out)
{
    State s;
    s.n = w->num_variables;
    s.m = w->num_constraints;
    const int nnz = w->constraint_matrix_num_nonzeros;
    s.A.rows = s.m;
    s.A.cols = s.n;
    s.A.rp.assign(w->constraint_matrix_row_pointers, w->constraint_matrix_row_pointers + s.m + 1);
    s.A.ci.assign(w->constraint_matrix_col_indices, w->constraint_matrix_col_indices + nnz);
    s.A.v.assign(w->constraint_matrix_values, w->constraint_matrix_values + nnz);
    s.At = transpose(s.A);
    s.c.assign(w->objective_vector, w->objective_vector + s.n);
    s.lb.assign(w->variable_lower_bound, w->variable_lower_bound + s.n);
    s.ub.assign(w->variable_upper_bound, w->variable_upper_bound + s.n);
    s.l.assign(w->constraint_lower_bound, w->constraint_lower_bound + s.m);
    s.u.assign(w->constraint_upper_bound, w->constraint_upper_bound + s.m);
    s.x0.assign(s.n, 0.0);
    s.y0.assign(s.m, 0.0);
    rescale_problem(p, s);
    std::copy(s.A.v.begin(), s.A.v.end(), vals_out);
    std::copy(s.E.begin(), s.E.end(), E_out);
    std::copy(s.D.begin(), s.D.end(), D_out);
    scales_out[0] = s.b_scale;
    scales_out[1] = s.c_scale;
}
