// CP_AL optimize(): host driver of the MI355X plain path.
//
// Reference: LP/CP_AL/src/solver.cu:1175-1736 (main loop), :1795-2273 (state
// set-up), :3538-3609 (step size and primal weight), :2821-3126 (restart and
// step schedule), :3611-3686 (fixed-point error), :3828-3909 (result);
// LP/CP_AL/src/utils.cu:87-228 (power iteration), :275-359 (termination and
// restart tests), :812-917 (residual); LP/CP_AL/src/preconditioner.cu (scaling).
//
// MI355X design: one HIP stream; a block of K = eval_freq iterations plus the
// evaluation pass is captured once into a hipGraph and replayed every block.
// Each iteration is two launches (three with the AL predictor): the At.y SpMV
// fused with the primal update and the A.x_bar SpMV fused with the dual update.
// The evaluation pass is two dual-RHS SpMVs with fused deterministic
// reductions; the host reads 13 doubles per block (one synchronisation per
// eval_freq iterations) and runs the scalar restart / PID / termination logic,
// which is a handful of flops.
#include "cp_al.h"
#include "cp_al_internal.hpp"

#include "../common/csr_host.hpp"
#include "../common/hip_util.hpp"
#include "alpha_tiers.hpp"
#include "cp_al_kernels.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

namespace pdal {
namespace lp {

namespace {

std::mt19937 g_gen(1);                              // utils.cu:27 (process-global)
std::normal_distribution<double> g_dist(0.0, 1.0);  // utils.cu:28

struct Eval {
    double cross, pn2, dres2, cx, lx, rx2, dresmax;
    double ax2, dn2, pres2, contrib, ry2, presmax;
};

class Solver {
  public:
    Solver(const cp_al_parameters_t *p, const lp_problem_t *w) : P_(*p), W_(w) {}
    ~Solver()
    {
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        if (polish_graph_) (void)hipGraphExecDestroy(polish_graph_);
        if (stream_) (void)hipStreamDestroy(stream_);
    }
    cp_al_result_t *run(const lp_problem_t *original);

    // Session use: set up once, then solve repeatedly from the start point.
    void setup();
    cp_al_result_t *solve_once(const lp_problem_t *original);
    void time_kernels(int iters, double *out);
    double setup_seconds() const { return setup_time_; }
    void reset_iterates();

  private:
    // set-up
    void upload();
    void rescale();
    void autotune_lanes();
    void setup_sliced();
    double power_iteration();
    void init_steps();
    // iteration
    void launch_iteration(int k_offset, bool major);
    void launch_primal(const PrimalArgs &pa, int k_offset, bool major);
    void launch_dual(const DualArgs &da, int k_offset, bool major);
    void launch_predictor(const PredictorArgs &a);
    void launch_eval();
    void launch_block(int K);
    void launch_block_traced(int K, int trace_freq);
    void record_graph(int K);
    void sync_scalars();
    Eval read_eval();
    void fixed_point_error(const Eval &e);
    void residual(const Eval &e);
    bool should_restart(int K);
    void perform_restart(const Eval &e);
    void refresh_step_schedule();
    void init_sigma();
    void set_sigma_parameters();
    void init_alpha();
    double controller_clock() const;
    double delta_normal_motion(const double *rx);
    cp_al_result_t *make_result(const lp_problem_t *original, int charged);
    void display() const;
    // feasibility polishing (solver.cu:3912-4456)
    struct PolishOut {
        termination_reason_t reason;
        int total;
        double abs, rel, obj, time;
    };
    void feasibility_polish(std::chrono::steady_clock::time_point t_start);
    PolishOut polish_phase(bool primal, double pw0, std::chrono::steady_clock::time_point t_start);
    void launch_polish_block(int K);

    cp_al_parameters_t P_;
    const lp_problem_t *W_;
    hipStream_t stream_ = nullptr;
    int n_ = 0, m_ = 0;
    long nnz_ = 0;
    double obj_const_ = 0.0;
    DevCsr A_, At_;
    // column-sliced copies of A and At for the iteration passes when their
    // gathered vector is far larger than an XCD's L2 (sliced_csr.hpp)
    DevSliced sA_, sAt_;
    SlicedPlan planA_, planAt_;
    int sliced_mode_ = -1;  // PDAL_SPMV_SLICED: 0 never, 1 always, unset: chosen by timing
    int slice_unroll_ = 4;  // PDAL_SPMV_SLICE_UNROLL: chunks of 64 nonzeros in flight (2, 4, 8)
    DBuf<double> c_, lb_, ub_, l_, u_, E_, D_;
    DBuf<double> x0_, x_, xT_, xH_, xR_, dslack_;
    DBuf<double> y0_, y_, yT_, yH_, yR_;
    DBuf<double> x_start_, y_start_;  // scaled start point, kept for repeated solves
    DBuf<double> parts_at_, parts_a_, evalq_;
    DBuf<IterScalars> sc_;
    int eval_grid_at_ = 1, eval_grid_a_ = 1;
    PinnedBuf<double> h_eval_{16};
    PinnedBuf<IterScalars> h_sc_{1};
    hipGraphExec_t graph_exec_ = nullptr;
    int graph_K_ = 0;
    bool graph_pred_ = false;
    bool use_graph_ = true;

    double b_scale_ = 1.0, c_scale_ = 1.0, c_norm_ = 0.0, b_norm_ = 0.0;
    double step_size_ = 0.0, pw_ = 1.0, al_lambda_ = 0.0, al_penalty_ = 0.0, lambda_est_ = 1.0;
    // sigma mode (solver.cu:3141-3164): steps from sigma and lambda_param, PID on sigma at restarts
    bool sigma_mode_ = false, sigma_adaptive_ = false;
    double al_sigma_ = 0.0, sigma_base_ = 0.0, sigma_min_ = 0.0, sigma_max_ = 0.0, sigma_best_ = 0.0;
    double sigma_err_sum_ = 0.0, sigma_last_err_ = 0.0, lambda_param_ = 0.0;
    AlphaTiers tiers_;    // adaptive alpha tiers (solver.cu:3166-3536)
    bool cr_early_stopped_ = false;  // Curtis-Reid stopped at its fifth pass (preconditioner.cu:244-252)
    double pw_err_sum_ = 0.0, pw_last_err_ = 0.0, best_pw_ = 1.0, best_gap_ = INFINITY;
    int total_ = 0, inner_ = 0;
    double fp_ = 0.0, fp0_ = 0.0, last_trial_fp_ = INFINITY;
    double abs_p_ = 0, rel_p_ = 0, abs_d_ = 0, rel_d_ = 0, p_obj_ = 0, d_obj_ = 0, gap_ = 0, rel_gap_ = 0;
    termination_reason_t reason_ = TERMINATION_REASON_UNSPECIFIED;
    double rescale_time_ = 0.0, solve_time_ = 0.0, setup_time_ = 0.0;
    double pw_init_ = 1.0;
    Eval last_eval_{};
    // Restart candidate (choose_restart_candidate, solver.cu:2749-2819): 0 = the
    // T output, 1 = the current (reflected Halpern) iterate, 2 = the Halpern point.
    int restart_target_ = 0;
    double score_current_ = 0.0, score_average_ = INFINITY, score_halpern_ = INFINITY;
    bool fp_stale_ = false;  // a restart moved x_R or x_T after the last evaluation
    DBuf<double> scoreq_;
    PinnedBuf<double> h_score_{4};
    double delta_score(const double *cx, const double *rx, const double *cy, const double *ry);
    void choose_restart_candidate();
    void refresh_fp();
    // Positive-alpha controllers (solver.cu:187-1057): the zero-alpha epoch
    // history, the rollback trial with its snapshot, the deterministic pulse.
    struct Scalars {
        double step_size, pw, al_lambda, al_penalty, lambda_param, al_sigma, sigma_best, sigma_err_sum,
            sigma_last_err, pw_err_sum, pw_last_err, best_pw, best_gap, fp, fp0, last_trial_fp, abs_p, rel_p, abs_d,
            rel_d, p_obj, d_obj, gap, rel_gap, score_current, score_average, score_halpern;
        bool sigma_mode, fp_stale;
        AlphaTiers tiers;
        int total, inner, restart_target;
        termination_reason_t reason;
        Eval last_eval;
    };
    Scalars save_scalars() const;
    void load_scalars(const Scalars &v);
    struct ZeroEpoch {
        bool valid = false;
        int iterations = 0;
        double seconds = 0, fp_rate = 0, p_rate = 0, d_rate = 0, g_rate = 0, kkt_rate = 0;
    };
    struct Trial {  // alpha_trial_controller_t, solver.cu:210-241
        bool enabled = false, trace = false, force_reject = false, active = false, epoch_valid = false;
        int started = 0, accepted = 0, rejected = 0, positive_epochs = 0, positive_charged = 0, count = 0;
        ZeroEpoch hist[2];
        int ep_total = 0, t_total = 0, t_charged = 0, max_iters = 0;
        double ep_time = 0, ep_p = 0, ep_d = 0, ep_g = 0, t_time = 0, t_fp = 0, t_p = 0, t_d = 0, t_g = 0,
               max_sec = 0;
        bool snap_valid = false;
        Scalars snap{};
    } trial_;
    enum { kPulseIdle = 0, kPulsePositive = 1, kPulseDone = 2 };
    struct Pulse {  // alpha_direct_pulse_controller_t, solver.cu:250-274
        bool enabled = false, trace = false, force = false;
        int phase = 0, started = 0, completed = 0, aborted = 0, gate_eval = 0, gate_pass = 0, gate_rej = 0;
        int pulse_iters = 200, gate_anchor = 0, gate_prev_k = 0, gate_recent_k = 0;
        double gate_dfp = 0, pos_fp = 0, pos_p = 0, pos_d = 0, pos_g = 0;
        termination_reason_t pos_reason = TERMINATION_REASON_UNSPECIFIED;
    } pulse_;
    DBuf<double> snap_x_[5], snap_y_[5];  // x0, x, x_T, x_H, x_R and the y counterparts
    void trial_snapshot();
    double trial_restore();
    void set_epoch_start();
    void record_zero_epoch();
    bool trial_eligible(int charged) const;
    void trial_start(int charged);
    bool trial_accepts(int charged, const char **why) const;
    bool pulse_eligible(int charged, int K);
    void pulse_start(int charged, int K);
    void pulse_finish(int charged);
    // Problem data the iteration kernels read and the vectors the evaluation
    // reads; the feasibility-polishing phases point them at their own copies.
    struct View {
        const double *c, *lb, *ub, *l, *u;
    };
    View iter_view_{}, eval_view_{};
    const double *eval_x_ = nullptr;  // x_T of the objective products (the polished x in the dual phase)
    hipGraphExec_t polish_graph_ = nullptr;
    double polish_time_ = 0.0;
    int polish_iters_ = 0;
};

void Solver::upload()
{
    const lp_problem_t *w = W_;
    n_ = w->num_variables;
    m_ = w->num_constraints;
    nnz_ = w->constraint_matrix_num_nonzeros;
    obj_const_ = w->objective_constant;
    PDAL_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    use_graph_ = !(getenv("PDAL_NO_GRAPH") && atoi(getenv("PDAL_NO_GRAPH")) != 0);

    HostCsr T = transpose_csr(m_, n_, w->constraint_matrix_row_pointers, w->constraint_matrix_col_indices,
                              w->constraint_matrix_values);
    auto put_csr = [&](DevCsr &D, int rows, const int *rp, const int *ci, const double *v, const char *env) {
        D.rows = rows;
        D.nnz = (int)nnz_;
        D.rp.alloc((size_t)rows + 1);
        D.ci.alloc(nnz_);
        D.v.alloc(nnz_);
        D.rp.upload(rp, (size_t)rows + 1, stream_);
        D.ci.upload(ci, nnz_, stream_);
        D.v.upload(v, nnz_, stream_);
        D.G = pick_group(env, nnz_, rows);
    };
    put_csr(A_, m_, w->constraint_matrix_row_pointers, w->constraint_matrix_col_indices,
            w->constraint_matrix_values, "PDAL_SPMV_G_A");
    put_csr(At_, n_, T.rp.data(), T.ci.data(), T.v.data(), "PDAL_SPMV_G_AT");

    // Plan the column-sliced copies while the host CSR is at hand; the device
    // copies are made after rescaling (setup_sliced).  Automatic only when a
    // gathered vector exceeds 8 MB (twice an XCD's L2) on a large matrix.
    sliced_mode_ = getenv("PDAL_SPMV_SLICED") ? (atoi(getenv("PDAL_SPMV_SLICED")) != 0 ? 1 : 0) : -1;
    const long slice_cols = std::max(1, env_int("PDAL_SPMV_SLICE_COLS", (int)kSliceDefaultCols));
    const int rpw_env = env_int("PDAL_SPMV_SLICE_RPW", 0);
    slice_unroll_ = env_int("PDAL_SPMV_SLICE_UNROLL", 4);
    auto want = [&](int cols) {
        if (sliced_mode_ == 0 || nnz_ == 0) return false;
        if (sliced_mode_ == 1) return true;
        return (long)cols * (long)sizeof(double) > (8L << 20) && nnz_ >= (1L << 20);
    };
    if (want(m_) && !build_sliced_plan(n_, m_, T.rp.data(), T.ci.data(), slice_cols,
                                       sliced_rows_per_wave(n_, m_, rpw_env), planAt_))
        planAt_ = SlicedPlan{};
    if (want(n_) && !build_sliced_plan(m_, n_, w->constraint_matrix_row_pointers, w->constraint_matrix_col_indices,
                                       slice_cols, sliced_rows_per_wave(m_, n_, rpw_env), planA_))
        planA_ = SlicedPlan{};

    for (auto *b : {&c_, &lb_, &ub_, &D_, &x0_, &x_, &xT_, &xH_, &xR_, &dslack_}) b->alloc(n_);
    for (auto *b : {&l_, &u_, &E_, &y0_, &y_, &yT_, &yH_, &yR_}) b->alloc(m_);
    c_.upload(w->objective_vector, n_, stream_);
    lb_.upload(w->variable_lower_bound, n_, stream_);
    ub_.upload(w->variable_upper_bound, n_, stream_);
    l_.upload(w->constraint_lower_bound, m_, stream_);
    u_.upload(w->constraint_upper_bound, m_, stream_);
    for (auto *b : {&x0_, &x_, &xT_, &xH_, &xR_, &dslack_})
        PDAL_HIP(hipMemsetAsync(b->get(), 0, (size_t)n_ * sizeof(double), stream_));
    for (auto *b : {&y0_, &y_, &yT_, &yH_, &yR_})
        PDAL_HIP(hipMemsetAsync(b->get(), 0, (size_t)m_ * sizeof(double), stream_));
    if (w->primal_start) x0_.upload(w->primal_start, n_, stream_);
    if (w->dual_start) y0_.upload(w->dual_start, m_, stream_);

    // evaluation partials: fixed grids so reductions are deterministic
    eval_grid_at_ = std::max(1, std::min(2048, blocks_for((long)n_ * At_.G)));
    eval_grid_a_ = std::max(1, std::min(2048, blocks_for((long)m_ * A_.G)));
    parts_at_.alloc((size_t)eval_grid_at_ * kEvalStride);
    parts_a_.alloc((size_t)eval_grid_a_ * kEvalStride);
    evalq_.alloc(16);
    scoreq_.alloc(4);
    sc_.alloc(1);
    PDAL_HIP(hipMemsetAsync(sc_.get(), 0, sizeof(IterScalars), stream_));
    PDAL_HIP(hipStreamSynchronize(stream_));  // host CSR copies are released below
}

void Solver::rescale()
{   // rescale_problem, preconditioner.cu:379-475
    const auto t0 = std::chrono::steady_clock::now();
    DBuf<double> e(m_), d(n_), inv_e(m_), inv_d(n_);
    fill_kernel<<<blocks_for(m_), kBlock, 0, stream_>>>(E_.get(), 1.0, m_);
    fill_kernel<<<blocks_for(n_), kBlock, 0, stream_>>>(D_.get(), 1.0, n_);
    PDAL_LAUNCHED();
    auto scale_problem = [&]() {
        scale_vars_kernel<<<blocks_for(n_), kBlock, 0, stream_>>>(c_.get(), lb_.get(), ub_.get(), x0_.get(), d.get(),
                                                                 inv_d.get(), n_);
        scale_cons_kernel<<<blocks_for(m_), kBlock, 0, stream_>>>(l_.get(), u_.get(), y0_.get(), e.get(),
                                                                 inv_e.get(), m_);
        scale_rows_kernel<<<blocks_for(m_), kBlock, 0, stream_>>>(A_.rp.get(), A_.ci.get(), A_.v.get(), m_,
                                                                 inv_e.get(), inv_d.get());
        scale_rows_kernel<<<blocks_for(n_), kBlock, 0, stream_>>>(At_.rp.get(), At_.ci.get(), At_.v.get(), n_,
                                                                 inv_d.get(), inv_e.get());
        PDAL_LAUNCHED();
    };

    const char *cr_raw = getenv("CP_AL_CR_SCALING");
    const bool cr_forced = cr_raw && cr_raw[0] != '\0' && atoi(cr_raw) != 0;
    int cr_iters = cr_forced ? 20 : 0;
    const char *cr_iters_raw = getenv("CP_AL_CR_ITERS");
    if (cr_iters > 0 && cr_iters_raw && cr_iters_raw[0] != '\0') cr_iters = atoi(cr_iters_raw);
    cr_iters = std::max(cr_iters, 0);
    cr_early_stopped_ = false;
    if (cr_iters > 0) {  // curtis_reid_rescaling, :162-268
        if (P_.verbose) printf("  Curtis-Reid scaling (%d iterations)\n", cr_iters);
        // the edge-change diagnostic and the early stop after five passes (:171-178, :217-252)
        const char *diag_raw = getenv("CP_AL_CR_DIAGNOSTIC");
        const char *early_raw = getenv("CP_AL_CR_EARLY_STOP");
        const char *thr_raw = getenv("CP_AL_CR_EARLY_STOP_THRESHOLD");
        const bool diag = diag_raw && atoi(diag_raw) != 0;
        const bool early = early_raw && atoi(early_raw) != 0 && cr_iters == 20;
        const double thr = thr_raw ? atof(thr_raw) : 0.1;
        const bool measure = diag || early;
        const int egrid = std::max(1, std::min(1024, blocks_for(m_)));
        DBuf<double> eparts(measure ? egrid : 0), emax(measure ? 1 : 0);
        PDAL_HIP(hipMemsetAsync(e.get(), 0, (size_t)m_ * sizeof(double), stream_));
        PDAL_HIP(hipMemsetAsync(d.get(), 0, (size_t)n_ * sizeof(double), stream_));
        for (int it = 0; it < cr_iters; ++it) {
            cr_log_update_kernel<<<blocks_for(m_), kBlock, 0, stream_>>>(A_.view(), d.get(), e.get(),
                                                                         measure ? inv_e.get() : nullptr);
            cr_log_update_kernel<<<blocks_for(n_), kBlock, 0, stream_>>>(At_.view(), e.get(), d.get(),
                                                                         measure ? inv_d.get() : nullptr);
            PDAL_LAUNCHED();
            if (!measure) continue;
            cr_edge_change_kernel<<<egrid, kBlock, 0, stream_>>>(A_.view(), inv_e.get(), inv_d.get(), eparts.get());
            reduce_partials_kernel<<<1, kBlock, 0, stream_>>>(eparts.get(), egrid, 1, 0, 1, emax.get());
            PDAL_LAUNCHED();
            double mx = 0.0;
            emax.download(&mx, 1, stream_);
            PDAL_HIP(hipStreamSynchronize(stream_));
            if (diag) fprintf(stderr, "[cr-diagnostic] iteration=%d max_edge_log_change=%.12e\n", it + 1, mx);
            if (early && it + 1 == 5 && mx <= thr) {
                cr_early_stopped_ = true;
                printf("[cr-early-stop] selected=%d executed=%d max_edge_log_change=%.12e threshold=%.12e\n",
                       cr_iters, it + 1, mx, thr);
                break;
            }
        }
        cr_exp_kernel<<<blocks_for(m_), kBlock, 0, stream_>>>(e.get(), inv_e.get(), E_.get(), m_);
        cr_exp_kernel<<<blocks_for(n_), kBlock, 0, stream_>>>(d.get(), inv_d.get(), D_.get(), n_);
        PDAL_LAUNCHED();
        scale_problem();
    }
    if (P_.l_inf_ruiz_iterations > 0 && P_.verbose)
        printf("  Ruiz scaling (%d iterations)\n", P_.l_inf_ruiz_iterations);
    for (int it = 0; it < P_.l_inf_ruiz_iterations; ++it) {  // ruiz_rescaling, :270-296
        row_absmax_kernel<<<blocks_for(m_), kBlock, 0, stream_>>>(A_.view(), e.get());
        clamp_sqrt_accum_kernel<<<blocks_for(m_), kBlock, 0, stream_>>>(e.get(), inv_e.get(), E_.get(), m_);
        row_absmax_kernel<<<blocks_for(n_), kBlock, 0, stream_>>>(At_.view(), d.get());
        clamp_sqrt_accum_kernel<<<blocks_for(n_), kBlock, 0, stream_>>>(d.get(), inv_d.get(), D_.get(), n_);
        PDAL_LAUNCHED();
        scale_problem();
    }
    if (P_.has_pock_chambolle_alpha) {  // pock_chambolle_rescaling, :298-323
        if (P_.verbose) printf("  Pock-Chambolle scaling (alpha=%.4f)\n", P_.pock_chambolle_alpha);
        row_powsum_kernel<<<blocks_for(m_), kBlock, 0, stream_>>>(A_.view(), P_.pock_chambolle_alpha, e.get());
        clamp_sqrt_accum_kernel<<<blocks_for(m_), kBlock, 0, stream_>>>(e.get(), inv_e.get(), E_.get(), m_);
        row_powsum_kernel<<<blocks_for(n_), kBlock, 0, stream_>>>(At_.view(), 2.0 - P_.pock_chambolle_alpha,
                                                                  d.get());
        clamp_sqrt_accum_kernel<<<blocks_for(n_), kBlock, 0, stream_>>>(d.get(), inv_d.get(), D_.get(), n_);
        PDAL_LAUNCHED();
        scale_problem();
    }
    b_scale_ = 1.0;
    c_scale_ = 1.0;
    if (P_.bound_objective_rescaling) {  // bound_objective_rescaling, :325-377
        if (P_.verbose) printf("  Bound-objective scaling\n");
        const int grid = std::max(1, std::min(1024, blocks_for(std::max(m_, n_))));
        DBuf<double> parts((size_t)grid * 2), out(2);
        bound_contrib_partials_kernel<<<grid, kBlock, 0, stream_>>>(l_.get(), u_.get(), m_, c_.get(), n_,
                                                                    parts.get());
        reduce_partials_kernel<<<1, kBlock, 0, stream_>>>(parts.get(), grid, 2, 2, 2, out.get());
        PDAL_LAUNCHED();
        double h[2];
        out.download(h, 2, stream_);
        PDAL_HIP(hipStreamSynchronize(stream_));
        const double cs = 1.0 / (std::sqrt(h[0]) + 1.0);
        const double os = 1.0 / (std::sqrt(h[1]) + 1.0);
        bound_obj_scale_kernel<<<blocks_for((long)m_ + n_), kBlock, 0, stream_>>>(
            l_.get(), u_.get(), y0_.get(), m_, lb_.get(), ub_.get(), c_.get(), x0_.get(), n_, cs, os);
        PDAL_LAUNCHED();
        b_scale_ = cs;
        c_scale_ = os;
    }
    PDAL_HIP(hipStreamSynchronize(stream_));
    rescale_time_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

double Solver::power_iteration()
{   // estimate_maximum_singular_value, utils.cu:87-228
    const int max_iter = P_.sv_max_iter;
    std::vector<double> h(m_);
    for (int i = 0; i < m_; ++i) h[i] = g_dist(g_gen);
    DBuf<double> v(m_), vh(m_), w(n_);
    v.upload(h.data(), m_, stream_);
    const int gridA = std::max(1, std::min(2048, blocks_for((long)m_ * A_.G)));
    const int gridR = std::max(1, std::min(2048, blocks_for(m_)));
    DBuf<double> partsA(gridA), partsR((size_t)gridR * 2), parts0(gridR);
    DBuf<PowerState> ps(1);
    PowerState init{};
    ps.upload(&init, 1, stream_);
    // ||v0||
    sumsq_partials_kernel<<<gridR, kBlock, 0, stream_>>>(v.get(), m_, parts0.get());
    DBuf<double> tmp(1);
    reduce_partials_kernel<<<1, kBlock, 0, stream_>>>(parts0.get(), gridR, 1, 1, 1, tmp.get());
    PDAL_LAUNCHED();
    double ss = 0.0;
    tmp.download(&ss, 1, stream_);
    PDAL_HIP(hipStreamSynchronize(stream_));
    init.vnorm = std::sqrt(ss);
    ps.upload(&init, 1, stream_);

    auto one_iter = [&]() {
        power_normalize_kernel<<<blocks_for(m_), kBlock, 0, stream_>>>(v.get(), vh.get(), m_, ps.get(), max_iter);
        dispatch_group(At_.G, [&](auto g) {
            constexpr int G = decltype(g)::value;
            spmv_gated_kernel<G><<<blocks_for((long)n_ * G), kBlock, 0, stream_>>>(At_.view(), vh.get(), w.get(),
                                                                                  ps.get(), max_iter);
        });
        dispatch_group(A_.G, [&](auto g) {
            constexpr int G = decltype(g)::value;
            power_av_dot_kernel<G><<<gridA, kBlock, 0, stream_>>>(A_.view(), w.get(), vh.get(), v.get(),
                                                                 partsA.get(), ps.get(), max_iter);
        });
        power_finish_dot_kernel<<<1, kBlock, 0, stream_>>>(partsA.get(), gridA, ps.get(), max_iter);
        power_resid_kernel<<<gridR, kBlock, 0, stream_>>>(v.get(), vh.get(), m_, partsR.get(), ps.get(), max_iter);
        power_finish_resid_kernel<<<1, kBlock, 0, stream_>>>(partsR.get(), gridR, ps.get(), max_iter, P_.sv_tol);
        PDAL_LAUNCHED();
    };
    constexpr int kChunk = 32;
    hipGraphExec_t exec = nullptr;
    if (use_graph_) {
        hipGraph_t graph;
        PDAL_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < kChunk; ++i) one_iter();
        PDAL_HIP(hipStreamEndCapture(stream_, &graph));
        PDAL_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        PDAL_HIP(hipGraphDestroy(graph));
    }
    PowerState st{};
    for (int done_iters = 0; done_iters < max_iter; done_iters += kChunk) {
        if (exec) PDAL_HIP(hipGraphLaunch(exec, stream_));
        else
            for (int i = 0; i < kChunk; ++i) one_iter();
        ps.download(&st, 1, stream_);
        PDAL_HIP(hipStreamSynchronize(stream_));
        if (st.done || st.iter >= max_iter) break;
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    return std::sqrt(st.sigma_sq);
}

// Rows-per-wavefront autotuner: time a plain SpMV of each matrix for every
// lane count in {1,..,32} (hipEvents, 3 launches after a warm-up) and keep the
// fastest.  An explicit PDAL_SPMV_G_A / PDAL_SPMV_G_AT setting wins.
void Solver::autotune_lanes()
{
    if (nnz_ == 0 || (long)nnz_ < env_int("PDAL_AUTOTUNE_MIN_NNZ", 1 << 18)) return;
    DBuf<double> xa(std::max(n_, m_)), ya(std::max(n_, m_));
    PDAL_HIP(hipMemsetAsync(xa.get(), 0, (size_t)std::max(n_, m_) * sizeof(double), stream_));
    hipEvent_t e0, e1;
    PDAL_HIP(hipEventCreate(&e0));
    PDAL_HIP(hipEventCreate(&e1));
    auto tune = [&](DevCsr &M, const char *env) {
        if (getenv(env)) return;
        float best = 1e30f;
        int best_g = M.G;
        for (int G : {1, 2, 4, 8, 16, 32}) {
            auto launch = [&]() {
                dispatch_group(G, [&](auto g) {
                    constexpr int GG = decltype(g)::value;
                    spmv_kernel<GG><<<blocks_for((long)M.rows * GG), kBlock, 0, stream_>>>(M.view(), xa.get(), ya.get());
                });
            };
            launch();
            PDAL_HIP(hipEventRecord(e0, stream_));
            for (int r = 0; r < 3; ++r) launch();
            PDAL_HIP(hipEventRecord(e1, stream_));
            PDAL_LAUNCHED();
            PDAL_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            PDAL_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) {
                best = ms;
                best_g = G;
            }
        }
        M.G = best_g;
    };
    tune(At_, "PDAL_SPMV_G_AT");
    tune(A_, "PDAL_SPMV_G_A");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    eval_grid_at_ = std::max(1, std::min(2048, blocks_for((long)n_ * At_.G)));
    eval_grid_a_ = std::max(1, std::min(2048, blocks_for((long)m_ * A_.G)));
    parts_at_.alloc((size_t)eval_grid_at_ * kEvalStride);
    parts_a_.alloc((size_t)eval_grid_a_ * kEvalStride);
}

void Solver::setup_sliced()
{   // device copies of the planned sliced layouts (values after rescaling);
    // in automatic mode each is kept only if its plain SpMV beats the best
    // lane count of the CSR kernel on this matrix
    auto make = [&](DevSliced &S, SlicedPlan &P, const DevCsr &M) {
        S = DevSliced{};
        if (P.groups == 0) return;
        S.rows = P.rows;
        S.slices = P.slices;
        S.rpw = P.rpw;
        S.groups = P.groups;
        S.colbits = P.colbits;
        S.slice_cols = P.slice_cols;
        const long nnz = (long)P.perm.size();
        DBuf<int> perm(nnz);
        perm.upload(P.perm.data(), nnz, stream_);
        S.ci.alloc(nnz);
        S.v.alloc(nnz);
        S.seg.alloc(P.seg.size());
        S.ci.upload(P.ci.data(), nnz, stream_);
        S.seg.upload(P.seg.data(), P.seg.size(), stream_);
        sliced_permute_kernel<<<std::max(1, std::min(8192, blocks_for(nnz))), kBlock, 0, stream_>>>(
            perm.get(), nnz, M.v.get(), S.v.get());
        PDAL_LAUNCHED();
        PDAL_HIP(hipStreamSynchronize(stream_));
        P = SlicedPlan{};
        S.ready = true;
    };
    make(sAt_, planAt_, At_);
    make(sA_, planA_, A_);
    if (sliced_mode_ == 1) return;
    hipEvent_t e0, e1;
    PDAL_HIP(hipEventCreate(&e0));
    PDAL_HIP(hipEventCreate(&e1));
    auto keep_if_faster = [&](DevSliced &S, const DevCsr &M, int cols) {
        if (!S.ready) return;
        DBuf<double> xa(cols), ya(M.rows);
        PDAL_HIP(hipMemsetAsync(xa.get(), 0, (size_t)cols * sizeof(double), stream_));
        auto time3 = [&](auto &&launch) {
            launch();
            PDAL_HIP(hipEventRecord(e0, stream_));
            for (int r = 0; r < 3; ++r) launch();
            PDAL_HIP(hipEventRecord(e1, stream_));
            PDAL_LAUNCHED();
            PDAL_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            PDAL_HIP(hipEventElapsedTime(&ms, e0, e1));
            return ms;
        };
        const float t_csr = time3([&] {
            dispatch_group(M.G, [&](auto g) {
                constexpr int G = decltype(g)::value;
                spmv_kernel<G><<<blocks_for((long)M.rows * G), kBlock, 0, stream_>>>(M.view(), xa.get(), ya.get());
            });
        });
        const float t_sl = time3([&] {
            dispatch_unroll(slice_unroll_, [&](auto u) {
                constexpr int U = decltype(u)::value;
                sliced_spmv_kernel<U><<<S.groups, kSliceBlock, S.lds_bytes(), stream_>>>(S.view(), xa.get(), ya.get());
            });
        });
        if (!(t_sl < t_csr)) S = DevSliced{};
    };
    keep_if_faster(sAt_, At_, m_);
    keep_if_faster(sA_, A_, n_);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

void Solver::init_sigma()
{   // sigma mode on CP_AL_SIGMA or CP_AL_SIGMA_MODE (solver.cu:3563-3578)
    const bool sigma_env = env_set("CP_AL_SIGMA");
    sigma_mode_ = sigma_env || env_int("CP_AL_SIGMA_MODE", 0) != 0;
    al_sigma_ = sigma_base_ = sigma_min_ = sigma_max_ = sigma_best_ = 0.0;
    sigma_err_sum_ = sigma_last_err_ = lambda_param_ = 0.0;
    sigma_adaptive_ = false;
    if (sigma_mode_) set_sigma_parameters();
}

void Solver::set_sigma_parameters()
{   // initialize_sigma_parameters, solver.cu:3141-3164 (also the sigma trigger's switch, :1511)
    const bool sigma_env = env_set("CP_AL_SIGMA");
    lambda_param_ = env_double("CP_AL_LAMBDA_PARAM",
                               env_double("CP_AL_LAMBDA_FACTOR", 2.01) * std::fmax(lambda_est_, 1e-12));
    lambda_param_ = std::fmax(lambda_param_, (2.0 + 1e-12) * std::fmax(lambda_est_, 1e-12));
    sigma_base_ = 1.0 / std::sqrt(std::fmax(lambda_param_, 1e-16));
    sigma_min_ = env_double("CP_AL_SIGMA_MIN_FACTOR", 1e-4) * sigma_base_;
    sigma_max_ = env_double("CP_AL_SIGMA_MAX_FACTOR", 100.0) * sigma_base_;
    al_sigma_ = env_double("CP_AL_SIGMA", env_double("CP_AL_SIGMA_INIT_FACTOR", 6.0) * sigma_base_);
    al_sigma_ = std::fmin(std::fmax(al_sigma_, sigma_min_), sigma_max_);
    sigma_best_ = al_sigma_;
    sigma_err_sum_ = 0.0;
    sigma_last_err_ = 0.0;
    sigma_adaptive_ = env_int("CP_AL_SIGMA_ADAPTIVE", sigma_env ? 0 : 1) != 0;
    sigma_mode_ = true;
}

double Solver::controller_clock() const
{   // The reference's cumulative_time_sec (utils.cu:301) as the adaptive-alpha
    // epochs and the dynamic evaluation's time gate read it.
    // PDAL_CONTROLLER_SECONDS_PER_ITERATION > 0 replaces it by the iteration
    // count times that value, so that a test sees the same decisions in every run.
    const double per_iter = env_double("PDAL_CONTROLLER_SECONDS_PER_ITERATION", 0.0);
    return per_iter > 0.0 ? per_iter * (double)total_ : solve_time_;
}

void Solver::init_alpha()
{   // the alpha part of initialize_step_size_and_primal_weight, solver.cu:3579-3584
    al_lambda_ = env_double("CP_AL_LAMBDA", 0.05);
    al_lambda_ = tiers_.init(al_lambda_, controller_clock());
    if (sigma_mode_) tiers_.adaptive = false;
    refresh_step_schedule();
}

void Solver::refresh_step_schedule()
{   // solver.cu:3103-3126
    const double margin = env_double("CP_AL_STEP_MARGIN", 0.998 * 0.998);
    if (sigma_mode_) {  // tau = 1/(sigma lambda_param), dual step = penalty = sigma, no alpha
        const double sigma = std::fmax(al_sigma_, 0.0);
        if (sigma <= 0.0) {
            al_penalty_ = 0.0;
            lambda_param_ = 0.0;
            step_size_ = std::sqrt(std::fmax(margin / std::fmax(lambda_est_, 1e-12), 1e-24));
            al_lambda_ = 0.0;
            return;
        }
        if (!std::isfinite(lambda_param_) || lambda_param_ <= 0.0)
            lambda_param_ = env_double("CP_AL_LAMBDA_PARAM",
                                       env_double("CP_AL_LAMBDA_FACTOR", 2.01) * std::fmax(lambda_est_, 1e-12));
        lambda_param_ = std::fmax(lambda_param_, (2.0 + 1e-12) * std::fmax(lambda_est_, 1e-12));
        const double tau = 1.0 / (sigma * lambda_param_);
        const double rho = sigma;
        al_penalty_ = sigma;
        step_size_ = std::sqrt(std::fmax(tau * rho, 1e-24));
        pw_ = std::sqrt(std::fmax(rho / tau, 1e-24));
        al_lambda_ = 0.0;
        return;
    }
    al_lambda_ = std::fmin(std::fmax(al_lambda_, 0.0), 0.95);
    const double eta_sq = margin * (1.0 - al_lambda_) / std::fmax(lambda_est_, 1e-12);
    step_size_ = std::sqrt(std::fmax(eta_sq, 1e-24));
    const double dual_step = step_size_ * std::fmax(pw_, 1e-12);
    al_penalty_ = dual_step * al_lambda_ / std::fmax(1.0 - al_lambda_, 1e-12);
}

void Solver::init_steps()
{   // initialize_step_size_and_primal_weight, solver.cu:3538-3609
    lambda_est_ = 1.0;
    if (nnz_ != 0) {
        const double sv = power_iteration();
        lambda_est_ = sv * sv;
    }
    pw_ = P_.bound_objective_rescaling ? 1.0 : (c_norm_ + 1.0) / (b_norm_ + 1.0);
    init_sigma();
    init_alpha();
    best_pw_ = pw_;
    warn_unsupported_controllers(false);
    if (P_.verbose) {  // solver.cu:3587-3608
        fprintf(stderr,
                "[box-cpal] pw_init=%.3e al_lambda=%.3e sigma_mode=%d sigma_adapt=%d al_sigma=%.3e sigma_base=%.3e "
                "lambda_param=%.3e eta=%.3e refl=%.3e al_penalty=%.3e lambdaA=%.3e adaptive=%d lambda_grid=",
                pw_, al_lambda_, (int)sigma_mode_, (int)sigma_adaptive_, al_sigma_, sigma_base_, lambda_param_,
                step_size_, P_.reflection_coefficient, al_penalty_, lambda_est_, (int)tiers_.adaptive);
        for (int i = 0; i < tiers_.num; ++i) fprintf(stderr, "%s%.3e", i == 0 ? "" : "/", tiers_.values[i]);
        fprintf(stderr, "\n");
    }
}

void Solver::setup()
{   // initialize_solver_state, solver.cu:1795-2273
    const auto t0 = std::chrono::steady_clock::now();
    upload();
    rescale();
    autotune_lanes();
    setup_sliced();
    x_start_.alloc(n_);
    y_start_.alloc(m_);
    // anchor = current = T-output = Halpern point = start (solver.cu:1954-1965)
    auto copy = [&](DBuf<double> &dst, const DBuf<double> &src, int len) {
        PDAL_HIP(hipMemcpyAsync(dst.get(), src.get(), (size_t)len * sizeof(double), hipMemcpyDeviceToDevice,
                                stream_));
    };
    copy(x_start_, x0_, n_);
    copy(y_start_, y0_, m_);
    // norms of the unscaled working problem (solver.cu:1999-2067)
    const lp_problem_t *w = W_;
    const bool linf = P_.optimality_norm == NORM_TYPE_L_INF;
    double sq = 0.0, mx = 0.0;
    for (int j = 0; j < n_; ++j) {
        const double cj = w->objective_vector[j];
        if (linf) mx = std::max(mx, std::fabs(cj));
        else sq += cj * cj;
    }
    c_norm_ = linf ? mx : std::sqrt(sq);
    sq = 0.0;
    mx = 0.0;
    for (int i = 0; i < m_; ++i) {
        const double lo = w->constraint_lower_bound[i], hi = w->constraint_upper_bound[i];
        if (linf) {
            if (std::isfinite(lo) && lo != hi) mx = std::max(mx, std::fabs(lo));
            if (std::isfinite(hi)) mx = std::max(mx, std::fabs(hi));
        } else {
            if (std::isfinite(lo) && lo != hi) sq += lo * lo;
            if (std::isfinite(hi)) sq += hi * hi;
        }
    }
    b_norm_ = linf ? mx : std::sqrt(sq);
    PDAL_HIP(hipStreamSynchronize(stream_));
    iter_view_ = eval_view_ = View{c_.get(), lb_.get(), ub_.get(), l_.get(), u_.get()};
    eval_x_ = xT_.get();
    init_steps();
    pw_init_ = pw_;
    setup_time_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void Solver::reset_iterates()
{   // anchor = current = T-output = Halpern point = start (solver.cu:1954-1965)
    auto copy = [&](DBuf<double> &dst, const DBuf<double> &src, int len) {
        PDAL_HIP(hipMemcpyAsync(dst.get(), src.get(), (size_t)len * sizeof(double), hipMemcpyDeviceToDevice,
                                stream_));
    };
    for (auto *b : {&x0_, &x_, &xT_, &xH_}) copy(*b, x_start_, n_);
    for (auto *b : {&y0_, &y_, &yT_, &yH_}) copy(*b, y_start_, m_);
    for (auto *b : {&xR_, &dslack_}) PDAL_HIP(hipMemsetAsync(b->get(), 0, (size_t)n_ * sizeof(double), stream_));
    PDAL_HIP(hipMemsetAsync(yR_.get(), 0, (size_t)m_ * sizeof(double), stream_));
    pw_ = pw_init_;
    total_ = 0;
    solve_time_ = 0.0;
    init_sigma();
    init_alpha();
    best_pw_ = pw_;
    pw_err_sum_ = 0.0;
    pw_last_err_ = 0.0;
    best_gap_ = INFINITY;
    total_ = 0;
    inner_ = 0;
    fp_ = fp0_ = 0.0;
    last_trial_fp_ = INFINITY;
    abs_p_ = rel_p_ = abs_d_ = rel_d_ = p_obj_ = d_obj_ = gap_ = rel_gap_ = 0.0;
    reason_ = TERMINATION_REASON_UNSPECIFIED;
    solve_time_ = 0.0;
    last_eval_ = Eval{};
    restart_target_ = 0;
    score_current_ = 0.0;
    score_average_ = score_halpern_ = INFINITY;
    fp_stale_ = false;
    sync_scalars();
}

void Solver::sync_scalars()
{   // sync_step_sizes_to_gpu + sync_inner_count_to_gpu, solver.cu:1738-1755
    IterScalars &s = *h_sc_.get();
    s.tau = step_size_ / pw_;
    s.sigma = step_size_ * pw_;
    s.beta = al_penalty_;
    s.inner = inner_;
    s.pad = 0;
    PDAL_HIP(hipMemcpyAsync(sc_.get(), h_sc_.get(), sizeof(IterScalars), hipMemcpyHostToDevice, stream_));
}

void Solver::launch_iteration(int k_offset, bool major)
{
    const double rho = P_.reflection_coefficient;
    const bool pred = al_penalty_ > 0.0;
    const View &v = iter_view_;
    if (pred) {
        PredictorArgs a{A_.view(), x_.get(), y_.get(), v.l, v.u, yR_.get(), sc_.get()};
        launch_predictor(a);
    }
    PrimalArgs pa{At_.view(), pred ? yR_.get() : y_.get(), v.c, v.lb, v.ub, x0_.get(), x_.get(),
                  xR_.get(), xT_.get(), xH_.get(), dslack_.get(), sc_.get(), rho};
    launch_primal(pa, k_offset, major);
    DualArgs da{A_.view(), xR_.get(), v.l, v.u, y0_.get(), y_.get(), yT_.get(), yH_.get(), yR_.get(),
                sc_.get(), rho};
    launch_dual(da, k_offset, major);
    PDAL_LAUNCHED();
}

void Solver::launch_primal(const PrimalArgs &pa, int k_offset, bool major)
{   // At.y + primal update: the sliced pass when At has a sliced copy
    if (sAt_.ready) {
        const size_t lds = sAt_.lds_bytes();
        dispatch_unroll(slice_unroll_, [&](auto u) {
            constexpr int U = decltype(u)::value;
            if (major) primal_sliced_kernel<U, true><<<sAt_.groups, kSliceBlock, lds, stream_>>>(pa, sAt_.view(), k_offset);
            else primal_sliced_kernel<U, false><<<sAt_.groups, kSliceBlock, lds, stream_>>>(pa, sAt_.view(), k_offset);
        });
        return;
    }
    dispatch_group(At_.G, [&](auto g) {
        constexpr int G = decltype(g)::value;
        if (major) primal_kernel<G, true><<<blocks_for((long)n_ * G), kBlock, 0, stream_>>>(pa, k_offset);
        else primal_kernel<G, false><<<blocks_for((long)n_ * G), kBlock, 0, stream_>>>(pa, k_offset);
    });
}

void Solver::launch_dual(const DualArgs &da, int k_offset, bool major)
{   // A.x_bar + dual update: the sliced pass when A has a sliced copy
    if (sA_.ready) {
        const size_t lds = sA_.lds_bytes();
        dispatch_unroll(slice_unroll_, [&](auto u) {
            constexpr int U = decltype(u)::value;
            if (major) dual_sliced_kernel<U, true><<<sA_.groups, kSliceBlock, lds, stream_>>>(da, sA_.view(), k_offset);
            else dual_sliced_kernel<U, false><<<sA_.groups, kSliceBlock, lds, stream_>>>(da, sA_.view(), k_offset);
        });
        return;
    }
    dispatch_group(A_.G, [&](auto g) {
        constexpr int G = decltype(g)::value;
        if (major) dual_kernel<G, true><<<blocks_for((long)m_ * G), kBlock, 0, stream_>>>(da, k_offset);
        else dual_kernel<G, false><<<blocks_for((long)m_ * G), kBlock, 0, stream_>>>(da, k_offset);
    });
}

void Solver::launch_predictor(const PredictorArgs &a)
{
    if (sA_.ready) {
        dispatch_unroll(slice_unroll_, [&](auto u) {
            constexpr int U = decltype(u)::value;
            predictor_sliced_kernel<U><<<sA_.groups, kSliceBlock, sA_.lds_bytes(), stream_>>>(a, sA_.view());
        });
        return;
    }
    dispatch_group(A_.G, [&](auto g) {
        constexpr int G = decltype(g)::value;
        predictor_kernel<G><<<blocks_for((long)m_ * G), kBlock, 0, stream_>>>(a);
    });
}

void Solver::launch_eval()
{
    const View &v = eval_view_;
    EvalAtArgs ea{At_.view(), yR_.get(), yT_.get(), xR_.get(), eval_x_, x0_.get(), v.c, dslack_.get(),
                  D_.get(), parts_at_.get()};
    dispatch_group(At_.G, [&](auto g) {
        constexpr int G = decltype(g)::value;
        eval_at_kernel<G><<<eval_grid_at_, kBlock, 0, stream_>>>(ea);
    });
    EvalAArgs eb{A_.view(), xR_.get(), xT_.get(), yR_.get(), yT_.get(), y0_.get(), v.l, v.u, E_.get(),
                 parts_a_.get()};
    dispatch_group(A_.G, [&](auto g) {
        constexpr int G = decltype(g)::value;
        eval_a_kernel<G><<<eval_grid_a_, kBlock, 0, stream_>>>(eb);
    });
    reduce_partials_kernel<<<1, kBlock, 0, stream_>>>(parts_at_.get(), eval_grid_at_, kEvalStride, kEvalAtSums,
                                                      QAT_N, evalq_.get());
    reduce_partials_kernel<<<1, kBlock, 0, stream_>>>(parts_a_.get(), eval_grid_a_, kEvalStride, kEvalASums, QA_N,
                                                      evalq_.get() + 8);
    PDAL_LAUNCHED();
}

void Solver::record_graph(int K)
{
    if (graph_exec_) {
        PDAL_HIP(hipGraphExecDestroy(graph_exec_));
        graph_exec_ = nullptr;
    }
    hipGraph_t graph;
    PDAL_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    for (int i = 1; i <= K; ++i) launch_iteration(i, i == 1 || i == K);
    launch_eval();
    PDAL_HIP(hipStreamEndCapture(stream_, &graph));
    PDAL_HIP(hipGraphInstantiate(&graph_exec_, graph, nullptr, nullptr, 0));
    PDAL_HIP(hipGraphDestroy(graph));
    graph_K_ = K;
    graph_pred_ = al_penalty_ > 0.0;
}

void Solver::launch_block(int K)
{   // solver.cu:1301-1341: iteration 1 and K use the major kernels
    if (!use_graph_) {
        for (int i = 1; i <= K; ++i) launch_iteration(i, i == 1 || i == K);
        launch_eval();
        return;
    }
    if (!graph_exec_ || graph_K_ != K || graph_pred_ != (al_penalty_ > 0.0)) record_graph(K);
    PDAL_HIP(hipGraphLaunch(graph_exec_, stream_));
}

void Solver::launch_block_traced(int K, int trace_freq)
{   // the dense trace of solver.cu:1286-1300, 321-358: the block launched iteration by iteration, and at
    // every trace iteration that is not an evaluation the fixed-point metric of x_R - x_T, y_R - y_T as
    // the buffers hold them then (x_R of this iteration; x_T, y_T, y_R of the block's last major step)
    for (int i = 1; i <= K; ++i) {
        launch_iteration(i, i == 1 || i == K);
        const long long it = (long long)total_ + i;
        if (it % trace_freq == 0 && it % K != 0) {
            const double f = delta_score(xR_.get(), xT_.get(), yR_.get(), yT_.get());
            if (std::isfinite(f) && f > 0.0) printf("trace iter %lld fp %.12e\n", it, f);
        }
    }
    launch_eval();
}

Eval Solver::read_eval()
{
    PDAL_HIP(hipMemcpyAsync(h_eval_.get(), evalq_.get(), 16 * sizeof(double), hipMemcpyDeviceToHost, stream_));
    PDAL_HIP(hipStreamSynchronize(stream_));
    const double *q = h_eval_.get();
    Eval e;
    e.cross = q[QAT_CROSS];
    e.pn2 = q[QAT_PN2];
    e.dres2 = q[QAT_DRES2];
    e.cx = q[QAT_CX];
    e.lx = q[QAT_LX];
    e.rx2 = q[QAT_RX2];
    e.dresmax = q[QAT_DRESMAX];
    e.ax2 = q[8 + QA_AX2];
    e.dn2 = q[8 + QA_DN2];
    e.pres2 = q[8 + QA_PRES2];
    e.contrib = q[8 + QA_CONTRIB];
    e.ry2 = q[8 + QA_RY2];
    e.presmax = q[8 + QA_PRESMAX];
    return e;
}

void Solver::fixed_point_error(const Eval &e)
{   // solver.cu:3658-3686
    const double dn = std::sqrt(e.dn2), pn = std::sqrt(e.pn2), an = std::sqrt(e.ax2);
    if (sigma_mode_ && al_sigma_ > 0.0 && lambda_param_ > 0.0) {  // :3664-3671
        const double val = al_sigma_ * (lambda_param_ * pn * pn - an * an) + dn * dn / std::fmax(al_sigma_, 1e-16);
        fp_ = std::sqrt(std::fmax(val, 0.0));
        return;
    }
    const double movement = pn * pn * pw_ + dn * dn / pw_;
    const double interaction = 2 * step_size_ * e.cross;
    const double aug = step_size_ * al_penalty_ * an * an;
    fp_ = std::sqrt(std::fmax(movement + interaction - aug, 0.0));
}

void Solver::residual(const Eval &e)
{   // compute_residual, utils.cu:812-917
    if (P_.optimality_norm == NORM_TYPE_L_INF) {
        abs_p_ = e.presmax;
        abs_d_ = e.dresmax;
    } else {
        abs_p_ = std::sqrt(e.pres2) / b_scale_;
        abs_d_ = std::sqrt(e.dres2);
    }
    abs_d_ /= c_scale_;
    const double bc = b_scale_ * c_scale_;
    p_obj_ = e.cx / bc + obj_const_;
    d_obj_ = (e.lx + e.contrib) / bc + obj_const_;
    rel_p_ = abs_p_ / (1.0 + b_norm_);
    rel_d_ = abs_d_ / (1.0 + c_norm_);
    gap_ = std::fabs(p_obj_ - d_obj_);
    rel_gap_ = gap_ / (1.0 + std::fabs(p_obj_) + std::fabs(d_obj_));
}

bool Solver::should_restart(int K)
{   // should_do_adaptive_restart, utils.cu:319-359
    const restart_parameters_t &rp = P_.restart_params;
    const double suff = env_double("CP_AL_RESTART_SUFFICIENT", rp.sufficient_reduction_for_restart);
    const double nec = env_double("CP_AL_RESTART_NECESSARY", rp.necessary_reduction_for_restart);
    const double art = env_double("CP_AL_RESTART_ARTIFICIAL", rp.artificial_restart_threshold);
    bool r = false;
    if (total_ == K) {
        r = true;
    } else if (total_ > K) {
        if (fp_ <= suff * fp0_) r = true;
        if (fp_ <= nec * fp0_ && fp_ > last_trial_fp_) r = true;
        if (inner_ >= art * total_) r = true;
    }
    last_trial_fp_ = fp_;
    return r;
}

double Solver::delta_normal_motion(const double *rx)
{   // ||A (restart point - x0)|| before the restart copies (solver.cu:2867-2884)
    const int grid = eval_grid_a_;
    dispatch_group(A_.G, [&](auto g) {
        constexpr int G = decltype(g)::value;
        delta_spmv_sumsq_kernel<G><<<grid, kBlock, 0, stream_>>>(A_.view(), rx, x0_.get(), parts_a_.get());
    });
    reduce_partials_kernel<<<1, kBlock, 0, stream_>>>(parts_a_.get(), grid, 1, 1, 1, evalq_.get());
    PDAL_LAUNCHED();
    double ss = 0.0;
    evalq_.download(&ss, 1, stream_);
    PDAL_HIP(hipStreamSynchronize(stream_));
    return std::sqrt(ss);
}

void Solver::perform_restart(const Eval &e)
{   // solver.cu:2821-3101: restart to the selected candidate; PID on the primal weight, or on sigma in sigma mode
    const int target = env_int("CP_AL_RESTART_SELECT", 0) != 0 ? restart_target_ : 0;  // :2823-2834
    const double *rx = target == 1 ? x_.get() : target == 2 ? xH_.get() : xT_.get();
    const double *ry = target == 1 ? y_.get() : target == 2 ? yH_.get() : yT_.get();
    double pd = std::sqrt(e.rx2), dd = std::sqrt(e.ry2);  // the evaluation's ||x_T - x0||, ||y_T - y0||
    if (target != 0) {
        const int grid = eval_grid_at_;
        dist2_kernel<<<grid, kBlock, 0, stream_>>>(rx, x0_.get(), n_, ry, y0_.get(), m_, parts_at_.get());
        reduce_partials_kernel<<<1, kBlock, 0, stream_>>>(parts_at_.get(), grid, 2, 2, 2, scoreq_.get());
        PDAL_LAUNCHED();
        PDAL_HIP(hipMemcpyAsync(h_score_.get(), scoreq_.get(), 2 * sizeof(double), hipMemcpyDeviceToHost, stream_));
        PDAL_HIP(hipStreamSynchronize(stream_));
        pd = std::sqrt(h_score_[0]);
        dd = std::sqrt(h_score_[1]);
    }
    const double ratio = rel_d_ / std::fmax(rel_p_, 1e-300);
    const restart_parameters_t &rp = P_.restart_params;
    if (sigma_mode_) {  // :2857-2966
        if (sigma_adaptive_) {
            const double ax = delta_normal_motion(rx);
            const double dx_metric_sq = lambda_param_ * pd * pd - ax * ax;
            if (dd > 1e-16 && dd < 1e12 && dx_metric_sq > 1e-24 && dx_metric_sq < 1e24 && std::isfinite(al_sigma_) &&
                al_sigma_ > 1e-16 && ratio > 1e-8 && ratio < 1e8) {
                const double gain = env_double("CP_AL_SIGMA_BALANCE_GAIN", 0.15);
                const double smooth = env_double("CP_AL_SIGMA_I_SMOOTH", rp.i_smooth);
                const double kp = env_double("CP_AL_SIGMA_KP", rp.k_p);
                const double ki = env_double("CP_AL_SIGMA_KI", rp.k_i);
                const double kd = env_double("CP_AL_SIGMA_KD", rp.k_d);
                const double rmin = env_double("CP_AL_SIGMA_UPDATE_MIN", 0.60);
                const double rmax = env_double("CP_AL_SIGMA_UPDATE_MAX", 1.70);
                double err = std::log(std::fmax(dd, 1e-300)) - 0.5 * std::log(std::fmax(dx_metric_sq, 1e-300)) -
                             std::log(std::fmax(al_sigma_, 1e-300));
                const double bal = std::fmax(-2.0, std::fmin(0.5 * std::log(std::fmax(ratio, 1e-300)), 2.0));
                err += gain * bal;
                sigma_err_sum_ = smooth * sigma_err_sum_ + err;
                const double de = err - sigma_last_err_;
                const double r = std::fmin(std::fmax(std::exp(kp * err + ki * sigma_err_sum_ + kd * de), rmin), rmax);
                al_sigma_ *= r;
                sigma_last_err_ = err;
            } else {
                al_sigma_ = sigma_best_;
                sigma_err_sum_ = 0.0;
                sigma_last_err_ = 0.0;
            }
        }
        if (!std::isfinite(al_sigma_) || al_sigma_ <= 1e-16) {
            al_sigma_ = sigma_best_;
            sigma_err_sum_ = 0.0;
            sigma_last_err_ = 0.0;
        }
        al_sigma_ = std::fmin(std::fmax(al_sigma_, sigma_min_), sigma_max_);
        const double g = std::fabs(std::log10(std::fmax(rel_d_, 1e-300) / std::fmax(rel_p_, 1e-300)));
        if (std::isfinite(g) && g < best_gap_) {
            best_gap_ = g;
            sigma_best_ = al_sigma_;
        }
        pw_err_sum_ = 0.0;
        pw_last_err_ = 0.0;
    } else if (pd > 1e-16 && dd > 1e-16 && pd < 1e12 && dd < 1e12 && ratio > 1e-8 && ratio < 1e8) {
        const double err = std::log(dd) - std::log(pd) - std::log(pw_);
        pw_err_sum_ *= rp.i_smooth;
        pw_err_sum_ += err;
        const double de = err - pw_last_err_;
        pw_ *= std::exp(rp.k_p * err + rp.k_i * pw_err_sum_ + rp.k_d * de);
        pw_last_err_ = err;
    } else {
        pw_ = best_pw_;
        pw_err_sum_ = 0.0;
        pw_last_err_ = 0.0;
    }
    if (!sigma_mode_) {
        const double g = std::fabs(std::log10(std::fmax(rel_d_, 1e-300) / std::fmax(rel_p_, 1e-300)));
        if (g < best_gap_) {
            best_gap_ = g;
            best_pw_ = pw_;
        }
        al_lambda_ = tiers_.on_restart(al_lambda_, rel_p_, rel_d_, rel_gap_, fp_, total_, controller_clock(),
                                       P_.verbose);  // :2996
    }
    refresh_step_schedule();
    const int select_mode = env_int("CP_AL_RESTART_SELECT", 0);
    if (select_mode != 0 && (P_.verbose || env_int("CP_AL_RESTART_SELECT_TRACE", 0) != 0))  // :3000-3015
        fprintf(stderr,
                "[restart-select] iter=%d target=%s score_current=%.12e score_reflected_halpern=%.12e "
                "score_halpern=%.12e candidate=%.12e margin=%.6e mode=%d\n",
                total_, target == 1 ? "reflected_halpern" : target == 2 ? "halpern" : "current", score_current_,
                score_average_, score_halpern_, fp_, env_double("CP_AL_RESTART_SELECT_MARGIN", 1.0), select_mode);
    auto copy = [&](double *dst, const double *src, int len) {
        if (len > 0)
            PDAL_HIP(hipMemcpyAsync(dst, src, (size_t)len * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    };
    if (target == 1) {  // :3019-3043: reflected = old T output; anchor = T output = the current iterate
        copy(xR_.get(), xT_.get(), n_);
        copy(yR_.get(), yT_.get(), m_);
        copy(x0_.get(), x_.get(), n_);
        copy(xT_.get(), x_.get(), n_);
        copy(y0_.get(), y_.get(), m_);
        copy(yT_.get(), y_.get(), m_);
        fp_stale_ = true;
    } else if (target == 2) {  // :3044-3076: reflected = old T output; anchor = current = T output = Halpern point
        copy(xR_.get(), xT_.get(), n_);
        copy(yR_.get(), yT_.get(), m_);
        for (double *d : {x0_.get(), x_.get(), xT_.get()}) copy(d, xH_.get(), n_);
        for (double *d : {y0_.get(), y_.get(), yT_.get()}) copy(d, yH_.get(), m_);
        fp_stale_ = true;
    } else {
        restart_copy_kernel<<<blocks_for((long)n_ + m_), kBlock, 0, stream_>>>(xT_.get(), x0_.get(), x_.get(), n_,
                                                                              yT_.get(), y0_.get(), y_.get(), m_);
        PDAL_LAUNCHED();
    }
    tiers_.last_restart_iter = total_;  // :3098
    inner_ = 0;
    last_trial_fp_ = INFINITY;
}

double Solver::delta_score(const double *cx, const double *rx, const double *cy, const double *ry)
{   // compute_restart_delta_score (solver.cu:2663-2747), the fixed-point metric of (candidate - reference)
    dispatch_group(At_.G, [&](auto g) {
        constexpr int G = decltype(g)::value;
        score_at_kernel<G><<<eval_grid_at_, kBlock, 0, stream_>>>(At_.view(), cy, ry, cx, rx, parts_at_.get());
    });
    dispatch_group(A_.G, [&](auto g) {
        constexpr int G = decltype(g)::value;
        score_a_kernel<G><<<eval_grid_a_, kBlock, 0, stream_>>>(A_.view(), cx, rx, cy, ry, parts_a_.get());
    });
    reduce_partials_kernel<<<1, kBlock, 0, stream_>>>(parts_at_.get(), eval_grid_at_, 2, 2, 2, scoreq_.get());
    reduce_partials_kernel<<<1, kBlock, 0, stream_>>>(parts_a_.get(), eval_grid_a_, 2, 2, 2, scoreq_.get() + 2);
    PDAL_LAUNCHED();
    PDAL_HIP(hipMemcpyAsync(h_score_.get(), scoreq_.get(), 4 * sizeof(double), hipMemcpyDeviceToHost, stream_));
    PDAL_HIP(hipStreamSynchronize(stream_));
    const double pn2 = h_score_[0], cross = h_score_[1], an2 = h_score_[2], dn2 = h_score_[3];
    if (sigma_mode_ && al_sigma_ > 0.0 && lambda_param_ > 0.0) {
        const double val = al_sigma_ * (lambda_param_ * pn2 - an2) + dn2 / std::fmax(al_sigma_, 1e-16);
        return std::sqrt(std::fmax(val, 0.0));
    }
    const double movement = pn2 * pw_ + dn2 / pw_;
    return std::sqrt(std::fmax(movement + 2 * step_size_ * cross - step_size_ * al_penalty_ * an2, 0.0));
}

void Solver::refresh_fp()
{   // compute_fixed_point_error (solver.cu:3611-3686) at the current x_R - x_T, y_R - y_T: from the last
    // evaluation's terms unless a restart has moved those vectors since
    if (fp_stale_) fp_ = delta_score(xR_.get(), xT_.get(), yR_.get(), yT_.get());
    else fixed_point_error(last_eval_);
}

void Solver::choose_restart_candidate()
{   // solver.cu:2749-2819
    const int mode = env_int("CP_AL_RESTART_SELECT", 0);
    restart_target_ = 0;
    score_current_ = fp_;
    score_average_ = INFINITY;
    score_halpern_ = INFINITY;
    if (mode == 0) return;
    if (total_ < env_int("CP_AL_RESTART_SELECT_MIN_ITERS", 0) ||
        solve_time_ < env_double("CP_AL_RESTART_SELECT_MIN_TIME", 0.0))
        return;
    score_average_ = delta_score(xT_.get(), x_.get(), yT_.get(), y_.get());
    const bool three = mode == 3 || env_int("CP_AL_RESTART_SELECT_THREEPOINT", 0) != 0;
    if (three) score_halpern_ = delta_score(xT_.get(), xH_.get(), yT_.get(), yH_.get());
    const double margin = env_double("CP_AL_RESTART_SELECT_MARGIN", 1.0);
    if (mode == 2) {
        restart_target_ = 1;
        fp_ = score_average_;
        return;
    }
    if (!std::isfinite(score_current_)) {
        fp_ = score_current_;
        return;
    }
    int best = 0;
    double score = score_current_;
    if (std::isfinite(score_average_) && score_average_ < score) {
        best = 1;
        score = score_average_;
    }
    if (three && std::isfinite(score_halpern_) && score_halpern_ < score) {
        best = 2;
        score = score_halpern_;
    }
    if (best != 0 && score <= margin * score_current_) {
        restart_target_ = best;
        fp_ = score;
    } else {
        fp_ = score_current_;
    }
}

// ---------------------------------------------------------------------------
// Positive-alpha controllers (solver.cu:187-1057; used at :1202-1222,
// :1372-1494, :1631-1695).  Host bookkeeping at the evaluation points; the
// trial's snapshot holds the scalars and the ten iterate vectors.

double ctl_metric(double v) { return std::fmax(std::fabs(v), 1e-300); }  // alpha_trial_metric, :409-412

double ctl_log_rate(double start, double end, int iters)  // alpha_trial_log_rate, :414-422
{
    if (iters <= 0 || !std::isfinite(start) || !std::isfinite(end)) return NAN;
    const double r = std::log(ctl_metric(start) / ctl_metric(end)) / (double)iters;
    return std::isfinite(r) ? r : NAN;
}

double ctl_predict(double rate, int iters)  // alpha_trial_predict_ratio, :424-428
{
    return std::exp(std::fmax(-700.0, std::fmin(700.0, -rate * (double)iters)));
}

Solver::Scalars Solver::save_scalars() const
{
    Scalars v;
    v.step_size = step_size_;
    v.pw = pw_;
    v.al_lambda = al_lambda_;
    v.al_penalty = al_penalty_;
    v.lambda_param = lambda_param_;
    v.al_sigma = al_sigma_;
    v.sigma_best = sigma_best_;
    v.sigma_err_sum = sigma_err_sum_;
    v.sigma_last_err = sigma_last_err_;
    v.pw_err_sum = pw_err_sum_;
    v.pw_last_err = pw_last_err_;
    v.best_pw = best_pw_;
    v.best_gap = best_gap_;
    v.fp = fp_;
    v.fp0 = fp0_;
    v.last_trial_fp = last_trial_fp_;
    v.abs_p = abs_p_;
    v.rel_p = rel_p_;
    v.abs_d = abs_d_;
    v.rel_d = rel_d_;
    v.p_obj = p_obj_;
    v.d_obj = d_obj_;
    v.gap = gap_;
    v.rel_gap = rel_gap_;
    v.score_current = score_current_;
    v.score_average = score_average_;
    v.score_halpern = score_halpern_;
    v.sigma_mode = sigma_mode_;
    v.fp_stale = fp_stale_;
    v.tiers = tiers_;
    v.total = total_;
    v.inner = inner_;
    v.restart_target = restart_target_;
    v.reason = reason_;
    v.last_eval = last_eval_;
    return v;
}

void Solver::load_scalars(const Scalars &v)
{
    step_size_ = v.step_size;
    pw_ = v.pw;
    al_lambda_ = v.al_lambda;
    al_penalty_ = v.al_penalty;
    lambda_param_ = v.lambda_param;
    al_sigma_ = v.al_sigma;
    sigma_best_ = v.sigma_best;
    sigma_err_sum_ = v.sigma_err_sum;
    sigma_last_err_ = v.sigma_last_err;
    pw_err_sum_ = v.pw_err_sum;
    pw_last_err_ = v.pw_last_err;
    best_pw_ = v.best_pw;
    best_gap_ = v.best_gap;
    fp_ = v.fp;
    fp0_ = v.fp0;
    last_trial_fp_ = v.last_trial_fp;
    abs_p_ = v.abs_p;
    rel_p_ = v.rel_p;
    abs_d_ = v.abs_d;
    rel_d_ = v.rel_d;
    p_obj_ = v.p_obj;
    d_obj_ = v.d_obj;
    gap_ = v.gap;
    rel_gap_ = v.rel_gap;
    score_current_ = v.score_current;
    score_average_ = v.score_average;
    score_halpern_ = v.score_halpern;
    sigma_mode_ = v.sigma_mode;
    fp_stale_ = v.fp_stale;
    tiers_ = v.tiers;
    total_ = v.total;
    inner_ = v.inner;
    restart_target_ = v.restart_target;
    reason_ = v.reason;
    last_eval_ = v.last_eval;
}

void Solver::trial_snapshot()
{   // alpha_trial_snapshot_save, :442-475
    DBuf<double> *xs[5] = {&x0_, &x_, &xT_, &xH_, &xR_}, *ys[5] = {&y0_, &y_, &yT_, &yH_, &yR_};
    for (int k = 0; k < 5; ++k) {
        if (!snap_x_[k].get()) snap_x_[k].alloc(n_);
        if (!snap_y_[k].get()) snap_y_[k].alloc(m_);
        if (n_ > 0)
            PDAL_HIP(hipMemcpyAsync(snap_x_[k].get(), xs[k]->get(), (size_t)n_ * sizeof(double),
                                    hipMemcpyDeviceToDevice, stream_));
        if (m_ > 0)
            PDAL_HIP(hipMemcpyAsync(snap_y_[k].get(), ys[k]->get(), (size_t)m_ * sizeof(double),
                                    hipMemcpyDeviceToDevice, stream_));
    }
    trial_.snap = save_scalars();
    trial_.snap_valid = true;
}

double Solver::trial_restore()
{   // alpha_trial_snapshot_restore, :502-585: the snapshot's scalars and vectors, the elapsed time kept;
    // returns the largest 2-norm change of the ten vectors
    if (!trial_.snap_valid) return INFINITY;
    DBuf<double> *xs[5] = {&x0_, &x_, &xT_, &xH_, &xR_}, *ys[5] = {&y0_, &y_, &yT_, &yH_, &yR_};
    double diff = 0.0;
    for (int k = 0; k < 5; ++k) {
        dist2_kernel<<<eval_grid_at_, kBlock, 0, stream_>>>(xs[k]->get(), snap_x_[k].get(), n_, ys[k]->get(),
                                                             snap_y_[k].get(), m_, parts_at_.get());
        reduce_partials_kernel<<<1, kBlock, 0, stream_>>>(parts_at_.get(), eval_grid_at_, 2, 2, 2, scoreq_.get());
        PDAL_LAUNCHED();
        PDAL_HIP(hipMemcpyAsync(h_score_.get(), scoreq_.get(), 2 * sizeof(double), hipMemcpyDeviceToHost, stream_));
        PDAL_HIP(hipStreamSynchronize(stream_));
        diff = std::fmax(diff, std::fmax(std::sqrt(h_score_[0]), std::sqrt(h_score_[1])));
        if (n_ > 0)
            PDAL_HIP(hipMemcpyAsync(xs[k]->get(), snap_x_[k].get(), (size_t)n_ * sizeof(double),
                                    hipMemcpyDeviceToDevice, stream_));
        if (m_ > 0)
            PDAL_HIP(hipMemcpyAsync(ys[k]->get(), snap_y_[k].get(), (size_t)m_ * sizeof(double),
                                    hipMemcpyDeviceToDevice, stream_));
    }
    const double elapsed = solve_time_;
    load_scalars(trial_.snap);
    solve_time_ = elapsed;
    return diff;
}

void Solver::set_epoch_start()
{   // alpha_trial_set_epoch_start, :605-614
    trial_.epoch_valid = true;
    trial_.ep_total = total_;
    trial_.ep_time = solve_time_;
    trial_.ep_p = ctl_metric(rel_p_);
    trial_.ep_d = ctl_metric(rel_d_);
    trial_.ep_g = ctl_metric(rel_gap_);
}

void Solver::record_zero_epoch()
{   // alpha_trial_record_zero_epoch, :616-678: rates of the epoch that ends at this restart
    Trial &t = trial_;
    if (!t.epoch_valid || t.active || al_lambda_ > 0.0) return;
    ZeroEpoch r;
    r.iterations = total_ - t.ep_total;
    r.seconds = solve_time_ - t.ep_time;
    const double k0 = std::fmax(t.ep_p, std::fmax(t.ep_d, t.ep_g));
    const double k1 = std::fmax(ctl_metric(rel_p_), std::fmax(ctl_metric(rel_d_), ctl_metric(rel_gap_)));
    r.fp_rate = ctl_log_rate(fp0_, fp_, r.iterations);
    r.p_rate = ctl_log_rate(t.ep_p, rel_p_, r.iterations);
    r.d_rate = ctl_log_rate(t.ep_d, rel_d_, r.iterations);
    r.g_rate = ctl_log_rate(t.ep_g, rel_gap_, r.iterations);
    r.kkt_rate = ctl_log_rate(k0, k1, r.iterations);
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
                total_, r.iterations, r.seconds, r.fp_rate, r.p_rate, r.d_rate, r.g_rate, r.kkt_rate, t.count);
}

bool Solver::trial_eligible(int charged) const
{   // alpha_trial_is_eligible, :680-703.  PDAL_ALPHA_TRIAL_MIN_SEC lowers the
    // reference's 2 s floor on the elapsed time, so a test reaches the trial at
    // a deterministic point.
    const Trial &t = trial_;
    const double min_sec = env_double("PDAL_ALPHA_TRIAL_MIN_SEC", 2.0);
    if (!t.enabled || t.started != 0 || t.count < 2 || charged < 10000 || solve_time_ < min_sec ||
        reason_ != TERMINATION_REASON_UNSPECIFIED)
        return false;
    const int hmax = std::max(t.hist[0].iterations, t.hist[1].iterations);
    const double hsec = std::fmax(2.0, std::fmax(t.hist[0].seconds, t.hist[1].seconds));
    const int rem_it = P_.termination_criteria.iteration_limit - charged;
    const double rem_sec = P_.termination_criteria.time_sec_limit - solve_time_;
    return hmax > 0 && rem_it >= 3 * hmax && rem_sec >= 3.0 * hsec;
}

void Solver::trial_start(int charged)
{   // alpha_trial_start, :704-761
    Trial &t = trial_;
    trial_snapshot();
    t.started = 1;
    t.active = true;
    t.t_total = total_;
    t.t_charged = charged;
    t.t_time = solve_time_;
    t.t_p = ctl_metric(rel_p_);
    t.t_d = ctl_metric(rel_d_);
    t.t_g = ctl_metric(rel_gap_);
    t.max_iters = 2 * std::max(t.hist[0].iterations, t.hist[1].iterations);
    t.max_sec = 2.0 * std::fmax(2.0, std::fmax(t.hist[0].seconds, t.hist[1].seconds));
    sigma_mode_ = false;
    tiers_.adaptive = false;
    al_lambda_ = 0.002;
    refresh_step_schedule();
    refresh_fp();
    t.t_fp = ctl_metric(fp_);
    if (!std::isfinite(t.t_fp) || t.t_fp <= 0.0) {
        t.active = false;
        t.rejected = 1;
        const double d = trial_restore();
        fprintf(stderr,
                "[alpha-trial-restore] reason=invalid-start charged_iter=%d trajectory_iter=%d max_vector_l2=%.12e\n",
                charged, total_, d);
        t.snap_valid = false;
        return;
    }
    fprintf(stderr,
            "[alpha-trial-start] charged_iter=%d trajectory_iter=%d solve_sec=%.12e alpha=2.000000000000e-03 "
            "fp=%.12e primal=%.12e dual=%.12e gap=%.12e max_K=%d max_sec=%.12e\n",
            charged, total_, solve_time_, t.t_fp, t.t_p, t.t_d, t.t_g, t.max_iters, t.max_sec);
}

bool Solver::trial_accepts(int charged, const char **why) const
{   // alpha_trial_accepts, :777-861: the epoch at alpha > 0 against the rates of the last two zero-alpha epochs
    const Trial &t = trial_;
    const int k = total_ - t.t_total;
    if (k <= 0) {
        *why = "empty-epoch";
        return false;
    }
    const ZeroEpoch &h0 = t.hist[0], &h1 = t.hist[1];
    const double fp_pred = ctl_predict(0.5 * (h0.fp_rate + h1.fp_rate), k);
    const double p_pred = t.t_p * ctl_predict(0.5 * (h0.p_rate + h1.p_rate), k);
    const double d_pred = t.t_d * ctl_predict(0.5 * (h0.d_rate + h1.d_rate), k);
    const double g_pred = t.t_g * ctl_predict(0.5 * (h0.g_rate + h1.g_rate), k);
    const double k_pred = std::fmax(t.t_p, std::fmax(t.t_d, t.t_g)) * ctl_predict(0.5 * (h0.kkt_rate + h1.kkt_rate), k);
    const double fp_obs = ctl_metric(fp_) / t.t_fp;
    const double p_obs = ctl_metric(rel_p_), d_obs = ctl_metric(rel_d_), g_obs = ctl_metric(rel_gap_);
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
            *why, ok ? 1 : 0, charged, total_, k, fp_obs, fp_pred, k_obs, k_pred, p_obs, p_pred, d_obs, d_pred, g_obs,
            g_pred);
    return ok;
}

bool Solver::pulse_eligible(int charged, int K)
{   // alpha_direct_pulse_is_eligible, :906-958: the gate passes only when forced
    Pulse &d = pulse_;
    const Trial &h = trial_;
    if (!d.enabled || d.started != 0 || d.gate_eval != 0 || d.phase != kPulseIdle || h.count < 2 ||
        charged < 10000 || reason_ != TERMINATION_REASON_UNSPECIFIED || K != 200)
        return false;
    for (const ZeroEpoch &e : h.hist)
        if (!e.valid || e.iterations <= 0 || !std::isfinite(e.fp_rate)) return false;
    if (P_.termination_criteria.iteration_limit - charged < 400) return false;
    d.gate_eval = 1;
    d.gate_anchor = charged;
    d.gate_prev_k = h.hist[0].iterations;
    d.gate_recent_k = h.hist[1].iterations;
    d.gate_dfp = h.hist[1].fp_rate - h.hist[0].fp_rate;
    d.gate_pass = d.force ? 1 : 0;
    d.gate_rej = d.force ? 0 : 1;
    fprintf(stderr,
            "[alpha-deterministic-tail-anchor] charged_iter=%d trajectory_iter=%d previous_K=%d recent_K=%d "
            "previous_fp_rate=%.12e recent_fp_rate=%.12e delta_fp_rate=%.12e pass=%d force_pulse=%d\n",
            charged, total_, d.gate_prev_k, d.gate_recent_k, h.hist[0].fp_rate, h.hist[1].fp_rate, d.gate_dfp,
            d.gate_pass, d.force ? 1 : 0);
    return d.force;
}

void Solver::pulse_start(int charged, int K)
{   // alpha_direct_pulse_start, :960-982: one evaluation block at alpha = 0.002
    Pulse &d = pulse_;
    d.started = 1;
    d.phase = kPulsePositive;
    d.pulse_iters = K;
    sigma_mode_ = false;
    tiers_.adaptive = false;
    al_lambda_ = 0.002;
    refresh_step_schedule();
    fprintf(stderr,
            "[alpha-direct-pulse-start] charged_iter=%d trajectory_iter=%d alpha=2.000000000000e-03 "
            "pulse_iterations=%d\n",
            charged, total_, d.pulse_iters);
}

void Solver::pulse_finish(int charged)
{   // alpha_direct_pulse_finish_positive, :984-1025
    Pulse &d = pulse_;
    if (!d.enabled || d.phase != kPulsePositive) return;
    d.pos_p = ctl_metric(rel_p_);
    d.pos_d = ctl_metric(rel_d_);
    d.pos_g = ctl_metric(rel_gap_);
    d.pos_reason = reason_;
    reason_ = TERMINATION_REASON_UNSPECIFIED;
    sigma_mode_ = false;
    tiers_.adaptive = false;
    al_lambda_ = 0.0;
    refresh_step_schedule();
    refresh_fp();
    d.pos_fp = ctl_metric(fp_);
    choose_restart_candidate();
    reason_ = d.pos_reason;
    d.completed = 1;
    d.aborted = d.pos_reason != TERMINATION_REASON_UNSPECIFIED && d.pos_reason != TERMINATION_REASON_OPTIMAL;
    d.phase = kPulseDone;
    fprintf(stderr,
            "[alpha-direct-pulse-complete] charged_iter=%d trajectory_iter=%d positive_iters=%d fp0map=%.12e "
            "primal=%.12e dual=%.12e gap=%.12e term=%d\n",
            charged, total_, d.pulse_iters, d.pos_fp, d.pos_p, d.pos_d, d.pos_g, (int)d.pos_reason);
}

int print_frequency(int iter)
{   // get_print_frequency, utils.cu:637-647
    int step = 10;
    long long threshold = 1000;
    while (iter >= threshold) {
        step *= 10;
        threshold *= 10;
    }
    return step;
}

void Solver::display() const
{   // display_iteration_stats, utils.cu:613-635
    if (!P_.verbose) return;
    if (total_ % print_frequency(total_) == 0)
        printf("%6d %.1e | %8.1e  %8.1e | %.1e %.1e %.1e | %.1e %.1e %.1e | %.1e \n", total_, solve_time_, p_obj_,
               d_obj_, abs_p_, abs_d_, gap_, rel_p_, rel_d_, rel_gap_, fp_);
}

cp_al_result_t *Solver::make_result(const lp_problem_t *original, int charged)
{   // create_result_from_state, solver.cu:3828-3909
    DBuf<double> rc(n_), xo(n_), yo(m_);
    dispatch_group(At_.G, [&](auto g) {
        constexpr int G = decltype(g)::value;
        reduced_cost_kernel<G><<<blocks_for((long)n_ * G), kBlock, 0, stream_>>>(
            At_.view(), yT_.get(), c_.get(), D_.get(), lb_.get(), ub_.get(), c_scale_, rc.get());
    });
    unscale_kernel<<<blocks_for((long)n_ + m_), kBlock, 0, stream_>>>(xT_.get(), D_.get(), b_scale_, xo.get(), n_,
                                                                     yT_.get(), E_.get(), c_scale_, yo.get(), m_);
    PDAL_LAUNCHED();
    cp_al_result_t *r = (cp_al_result_t *)calloc(1, sizeof(cp_al_result_t));
    r->primal_solution = (double *)malloc((size_t)std::max(n_, 1) * sizeof(double));
    r->dual_solution = (double *)malloc((size_t)std::max(m_, 1) * sizeof(double));
    r->reduced_cost = (double *)malloc((size_t)std::max(n_, 1) * sizeof(double));
    xo.download(r->primal_solution, n_, stream_);
    yo.download(r->dual_solution, m_, stream_);
    rc.download(r->reduced_cost, n_, stream_);
    PDAL_HIP(hipStreamSynchronize(stream_));
    r->num_variables = original->num_variables;
    r->num_constraints = original->num_constraints;
    r->num_nonzeros = original->constraint_matrix_num_nonzeros;
    r->total_count = charged;
    r->rescaling_time_sec = rescale_time_;
    r->cumulative_time_sec = solve_time_;
    r->absolute_primal_residual = abs_p_;
    r->relative_primal_residual = rel_p_;
    r->absolute_dual_residual = abs_d_;
    r->relative_dual_residual = rel_d_;
    r->primal_objective_value = p_obj_;
    r->dual_objective_value = d_obj_;
    r->objective_gap = gap_;
    r->relative_objective_gap = rel_gap_;
    r->termination_reason = reason_;
    r->feasibility_polishing_time = polish_time_;
    r->feasibility_iteration = polish_iters_;
    return r;
}

cp_al_result_t *Solver::run(const lp_problem_t *original)
{
    setup();
    return solve_once(original);
}

cp_al_result_t *Solver::solve_once(const lp_problem_t *original)
{   // optimize main loop, solver.cu:1197-1736
    reset_iterates();
    display();
    trial_ = Trial{};
    trial_.enabled = env_int("CP_AL_ALPHA_ROLLBACK_TRIAL", 0) != 0;  // :587-603
    trial_.trace = env_int("CP_AL_ALPHA_ROLLBACK_TRACE", 0) != 0;
    trial_.force_reject = env_int("CP_AL_ALPHA_ROLLBACK_FORCE_REJECT", 0) != 0;
    pulse_ = Pulse{};
    pulse_.enabled = env_int("CP_AL_ALPHA_DIRECT_PULSE", 0) != 0;  // :884-904
    pulse_.trace = env_int("CP_AL_ALPHA_DIRECT_PULSE_TRACE", 0) != 0;
    pulse_.force = env_int("CP_AL_DETERMINISTIC_PULSE_FORCE", 0) != 0;
    if (trial_.enabled && pulse_.enabled)  // :1216-1222 (an error instead of exit)
        throw HipError("rollback trial and direct pulse cannot be enabled together");
    if (trial_.enabled)
        fprintf(stderr,
                "[alpha-trial-config] enabled=1 alpha=2.000000000000e-03 history=2 fp_gain=0.10 kkt_gain=0.05 "
                "component_slack=0.05 force_reject=%d\n",
                trial_.force_reject ? 1 : 0);
    if (pulse_.enabled)
        fprintf(stderr,
                "[alpha-direct-pulse-config] enabled=1 alpha=2.000000000000e-03 pulse_iterations=200 "
                "history_epochs=2 deterministic_anchor_iter_ge=10000 wall_time_features=0 tail_only=1 "
                "force_pulse=%d\n",
                pulse_.force ? 1 : 0);
    const int trial_test_limit = env_int("CP_AL_ALPHA_ROLLBACK_TEST_TRAJECTORY_LIMIT", 0);
    const int pulse_test_limit = env_int("CP_AL_ALPHA_DIRECT_PULSE_TEST_TRAJECTORY_LIMIT", 0);
    const auto t_start = std::chrono::steady_clock::now();
    int charged = total_;
    bool do_restart = false;
    const termination_criteria_t &tc = P_.termination_criteria;
    bool have_eval = false;
    // dynamic evaluation frequency (solver.cu:1228-1260, 1533-1584)
    const int base_k = P_.termination_evaluation_frequency;
    int eval_k = base_k;
    const int fast_k = std::min(std::max(env_int("CP_AL_DYNAMIC_EVAL_FAST", 50), 3), base_k);
    const bool dyn_residual = env_int("CP_AL_DYNAMIC_EVAL", 0) != 0;
    const double dyn_after_sec = env_double("CP_AL_DYNAMIC_EVAL_AFTER_SEC", -1.0);
    const int dyn_after_iters = env_int("CP_AL_DYNAMIC_EVAL_AFTER_ITERS", -1);
    const bool dyn_time = dyn_after_sec > 0.0, dyn_iters = dyn_after_iters > 0;
    const bool dyn = (dyn_residual || dyn_time || dyn_iters) && fast_k < base_k;
    const int dyn_min_iters = std::max(env_int("CP_AL_DYNAMIC_EVAL_MIN_ITERS", base_k), base_k);
    const double dyn_small = env_double("CP_AL_DYNAMIC_EVAL_SMALL_RES", 1e-6);
    const double dyn_large = env_double("CP_AL_DYNAMIC_EVAL_LARGE_RES", 1e-4);
    const double dyn_ratio = env_double("CP_AL_DYNAMIC_EVAL_RATIO", 1e4);
    const int dyn_bootstrap = std::max(env_int("CP_AL_DYNAMIC_EVAL_BOOTSTRAP_ITERS", 0), 0);
    bool dyn_latched = false;
    if (dyn && dyn_bootstrap > 0) eval_k = fast_k;
    // sigma trigger (solver.cu:1496-1528): switch to sigma mode once the primal
    // residual lags the dual one by the given ratio
    const bool trigger_on = env_int("CP_AL_SIGMA_TRIGGER", 0) != 0;
    const int trigger_min_iters = env_int("CP_AL_SIGMA_TRIGGER_MIN_ITERS", 200);
    const double trigger_primal = env_double("CP_AL_SIGMA_TRIGGER_PRIMAL", 1e-2);
    const double trigger_dual = env_double("CP_AL_SIGMA_TRIGGER_DUAL", 1e-6);
    const double trigger_ratio = env_double("CP_AL_SIGMA_TRIGGER_RATIO", 1e6);
    // verbose diagnostics: the dense trace and the progress lines (solver.cu:321-375)
    auto log_freq = [](const char *name) {
        const char *raw = getenv(name);
        const int v = (raw && raw[0]) ? atoi(raw) : 0;
        return v > 0 ? v : 0;
    };
    const int trace_freq = P_.verbose ? log_freq("CP_AL_LOG_FREQ") : 0;
    const int progress_freq = P_.verbose ? log_freq("CP_AL_PROGRESS_LOG_FREQ") : 0;
    auto positive_charge = [&]() {
        const int pos = charged - trial_.t_charged;
        trial_.positive_charged += pos;
        trial_.positive_epochs += pos > 0 ? 1 : 0;
        return pos;
    };
    while (charged < tc.iteration_limit) {
        const int K = eval_k;
        if (do_restart) {  // :1279-1284
            refresh_fp();
            fp0_ = fp_;
            do_restart = false;
        }
        sync_scalars();
        if (trace_freq > 0) launch_block_traced(K, trace_freq);
        else launch_block(K);
        const Eval e = read_eval();
        last_eval_ = e;
        fp_stale_ = false;
        have_eval = true;
        fixed_point_error(e);
        choose_restart_candidate();  // :1351
        residual(e);
        inner_ += K;
        total_ += K;
        charged += K;
        if (P_.verbose) {  // :1356-1363, 329-375
            if (std::isfinite(fp_) && fp_ > 0.0) printf("checkpoint iter %lld fp %.12e\n", (long long)total_, fp_);
            if (progress_freq > 0 && total_ % progress_freq == 0) {
                const double bal = std::fabs(std::log10(std::fmax(rel_d_, 1e-300) / std::fmax(rel_p_, 1e-300)));
                printf("progress iter %lld time %.6e rpr %.12e rdr %.12e rgap %.12e fp %.12e pw %.12e pw_best %.12e "
                       "bal %.12e\n",
                       (long long)total_, solve_time_, rel_p_, rel_d_, rel_gap_,
                       (std::isfinite(fp_) && fp_ > 0.0) ? fp_ : NAN, pw_, best_pw_, bal);
            }
        }
        if (total_ % print_frequency(total_) == 0) display();
        solve_time_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        if (rel_d_ < tc.eps_feasible_relative && rel_p_ < tc.eps_feasible_relative &&
            rel_gap_ < tc.eps_optimal_relative)
            reason_ = TERMINATION_REASON_OPTIMAL;
        else if (total_ >= tc.iteration_limit)
            reason_ = TERMINATION_REASON_ITERATION_LIMIT;
        else if (solve_time_ >= tc.time_sec_limit)
            reason_ = TERMINATION_REASON_TIME_LIMIT;
        if (reason_ == TERMINATION_REASON_UNSPECIFIED && charged >= tc.iteration_limit)
            reason_ = TERMINATION_REASON_ITERATION_LIMIT;
        // the controllers' test limits on the trajectory count (:1378-1391)
        if (reason_ == TERMINATION_REASON_UNSPECIFIED && trial_.enabled && trial_test_limit > 0 && !trial_.active &&
            trial_.rejected > 0 && total_ >= trial_test_limit)
            reason_ = TERMINATION_REASON_ITERATION_LIMIT;
        if (reason_ == TERMINATION_REASON_UNSPECIFIED && pulse_.enabled && pulse_test_limit > 0 &&
            pulse_.phase == kPulseDone && total_ >= pulse_test_limit)
            reason_ = TERMINATION_REASON_ITERATION_LIMIT;
        if (pulse_.phase == kPulsePositive) {  // :1416-1425
            pulse_finish(charged);
            if (reason_ == TERMINATION_REASON_UNSPECIFIED && pulse_test_limit > 0 && total_ >= pulse_test_limit)
                reason_ = TERMINATION_REASON_ITERATION_LIMIT;
        }
        if (reason_ != TERMINATION_REASON_UNSPECIFIED) {  // :1426-1474
            if (trial_.active) {
                positive_charge();
                if (reason_ == TERMINATION_REASON_OPTIMAL) {
                    trial_.accepted += 1;
                    trial_.active = false;
                    al_lambda_ = 0.0;
                    refresh_step_schedule();
                    fprintf(stderr,
                            "[alpha-trial-decision] reason=terminal-optimal accept=1 charged_iter=%d "
                            "trajectory_iter=%d K=%d\n",
                            charged, total_, total_ - trial_.t_total);
                } else {
                    const termination_reason_t why = reason_;
                    const double d = trial_restore();
                    reason_ = why;
                    trial_.rejected += 1;
                    trial_.active = false;
                    fprintf(stderr,
                            "[alpha-trial-restore] reason=charged-termination charged_iter=%d trajectory_iter=%d "
                            "max_vector_l2=%.12e\n",
                            charged, total_, d);
                }
                trial_.snap_valid = false;
            }
            break;
        }
        if (trial_.active && (charged - trial_.t_charged >= trial_.max_iters ||
                              solve_time_ - trial_.t_time >= trial_.max_sec)) {  // safety cap, :1476-1494
            const int pos = positive_charge();
            const double d = trial_restore();
            trial_.rejected += 1;
            trial_.active = false;
            set_epoch_start();
            fprintf(stderr,
                    "[alpha-trial-restore] reason=safety-cap charged_iter=%d trajectory_iter=%d positive_iters=%d "
                    "max_vector_l2=%.12e\n",
                    charged, total_, pos, d);
            trial_.snap_valid = false;
            do_restart = true;
            continue;
        }
        bool trigger_restart = false;
        if (trigger_on && !sigma_mode_ && total_ >= trigger_min_iters) {  // :1496-1528
            const double rp = std::fmax(rel_p_, 1e-300), rd = std::fmax(rel_d_, 1e-300);
            if (rp >= trigger_primal && rd <= trigger_dual && rp / rd >= trigger_ratio) {
                set_sigma_parameters();
                tiers_.adaptive = false;
                trigger_restart = true;
                if (P_.verbose)
                    printf("sigma-trigger iter %d rpr %.12e rdr %.12e ratio %.12e sigma %.12e base %.12e\n", total_,
                           rel_p_, rel_d_, rp / rd, al_sigma_, sigma_base_);
            }
        }
        if (dyn && !sigma_mode_ && (total_ >= dyn_min_iters || (dyn_bootstrap > 0 && total_ >= dyn_bootstrap))) {
            // :1533-1584
            const double rp = std::fmax(rel_p_, 1e-300), rd = std::fmax(rel_d_, 1e-300);
            const double lo = std::fmin(rp, rd), hi = std::fmax(rp, rd), ratio = hi / lo;
            const bool one_sided = lo <= dyn_small && hi >= dyn_large;
            const bool lag = ratio >= dyn_ratio && hi >= dyn_large;
            const double now = controller_clock();
            const bool time_gate = dyn_time && !cr_early_stopped_ && now >= dyn_after_sec;
            const bool iter_gate = dyn_iters && !cr_early_stopped_ && total_ >= dyn_after_iters;
            if (time_gate || iter_gate || (dyn_residual && (one_sided || lag))) {
                dyn_latched = true;
                if (eval_k != fast_k) {
                    eval_k = fast_k;
                    if (P_.verbose || dyn_time || dyn_iters)
                        printf("[dynamic-eval] iter=%d solve_sec=%.12e eval_frequency=%d->%d relative_primal=%.12e "
                               "relative_dual=%.12e residual_ratio=%.12e reason=%s\n",
                               total_, now, K, eval_k, rel_p_, rel_d_, ratio,
                               time_gate ? "time" : iter_gate ? "iteration" : "residual");
                }
            } else if (!dyn_latched && dyn_bootstrap > 0 && eval_k == fast_k && total_ >= dyn_bootstrap) {
                eval_k = base_k;
                if (P_.verbose)
                    printf("dynamic-eval iter %d eval_freq %d->%d rpr %.12e rdr %.12e ratio %.12e\n", total_, K,
                           eval_k, rel_p_, rel_d_, ratio);
            }
        }
        do_restart = trigger_restart || should_restart(K);
        if (!do_restart && tiers_.force_restart(total_)) {  // :1590-1600
            do_restart = true;
            if (P_.verbose) printf("lambda-race artificial restart iter %d fp %.12e\n", total_, fp_);
        }
        if (do_restart) {  // :1631-1695
            if (P_.verbose) printf("restart iter %d fp %.12e\n", total_, fp_);
            if (trial_.active) {
                const int pos = positive_charge();
                const char *why = nullptr;
                if (!trial_accepts(charged, &why)) {
                    const double d = trial_restore();
                    trial_.rejected += 1;
                    trial_.active = false;
                    set_epoch_start();
                    fprintf(stderr,
                            "[alpha-trial-restore] reason=%s charged_iter=%d trajectory_iter=%d positive_iters=%d "
                            "max_vector_l2=%.12e\n",
                            why, charged, total_, pos, d);
                    trial_.snap_valid = false;
                    do_restart = true;
                    continue;
                }
                trial_.accepted += 1;
                trial_.active = false;
                al_lambda_ = 0.0;
                refresh_step_schedule();
                trial_.snap_valid = false;
            } else {
                record_zero_epoch();
            }
            perform_restart(last_eval_);
            set_epoch_start();
            if (trial_eligible(charged)) trial_start(charged);
            if (pulse_eligible(charged, K)) pulse_start(charged, K);
        }
    }
    if (reason_ == TERMINATION_REASON_UNSPECIFIED) {  // :1704-1709
        reason_ = TERMINATION_REASON_ITERATION_LIMIT;
        if (!have_eval) {
            launch_eval();
            last_eval_ = read_eval();
        }
        residual(last_eval_);
        display();
    }
    if (trial_.enabled)  // alpha_trial_log_final, :863-882
        fprintf(stderr,
                "[alpha-trial-final] charged_iter=%d trajectory_iter=%d started=%d accepted=%d rejected=%d "
                "positive_epochs=%d positive_charged_iters=%d final_alpha=%.12e final_penalty=%.12e\n",
                charged, total_, trial_.started, trial_.accepted, trial_.rejected, trial_.positive_epochs,
                trial_.positive_charged, al_lambda_, al_penalty_);
    if (pulse_.enabled)  // alpha_direct_pulse_log_final, :1027-1057
        fprintf(stderr,
                "[alpha-direct-pulse-final] charged_iter=%d trajectory_iter=%d started=%d completed=%d aborted=%d "
                "gate_evaluated=%d gate_passed=%d gate_rejected=%d gate_anchor_charged_iter=%d gate_previous_K=%d "
                "gate_recent_K=%d gate_delta_fp_rate=%.12e positive_charged_iters=%d final_alpha=%.12e "
                "final_penalty=%.12e\n",
                charged, total_, pulse_.started, pulse_.completed, pulse_.aborted, pulse_.gate_eval,
                pulse_.gate_pass, pulse_.gate_rej, pulse_.gate_anchor, pulse_.gate_prev_k, pulse_.gate_recent_k,
                pulse_.gate_dfp, pulse_.completed ? pulse_.pulse_iters : 0, al_lambda_, al_penalty_);
    polish_time_ = 0.0;
    polish_iters_ = 0;
    if (P_.feasibility_polishing && reason_ != TERMINATION_REASON_DUAL_INFEASIBLE &&
        reason_ != TERMINATION_REASON_PRIMAL_INFEASIBLE)  // :1715-1719
        feasibility_polish(t_start);
    return make_result(original, charged);
}

// Launch `iters` plain iterations one kernel at a time with hipEvents around
// each SpMV-fused kernel; out = {avg ms primal, avg ms dual, avg ms predictor
// (0 when beta == 0), avg ms per whole iteration, avg ms eval pass}.
void Solver::time_kernels(int iters, double *out)
{
    hipEvent_t ev[8];
    for (auto &e : ev) PDAL_HIP(hipEventCreate(&e));
    double tp = 0, td = 0, tpr = 0;
    const double rho = P_.reflection_coefficient;
    const bool pred = al_penalty_ > 0.0;
    sync_scalars();
    for (int it = 0; it < iters; ++it) {
        PDAL_HIP(hipEventRecord(ev[0], stream_));
        if (pred) {
            PredictorArgs a{A_.view(), x_.get(), y_.get(), l_.get(), u_.get(), yR_.get(), sc_.get()};
            launch_predictor(a);
        }
        PDAL_HIP(hipEventRecord(ev[1], stream_));
        PrimalArgs pa{At_.view(), pred ? yR_.get() : y_.get(), c_.get(), lb_.get(), ub_.get(), x0_.get(), x_.get(),
                      xR_.get(), xT_.get(), xH_.get(), dslack_.get(), sc_.get(), rho};
        launch_primal(pa, it + 1, false);
        PDAL_HIP(hipEventRecord(ev[2], stream_));
        DualArgs da{A_.view(), xR_.get(), l_.get(), u_.get(), y0_.get(), y_.get(), yT_.get(), yH_.get(),
                    yR_.get(), sc_.get(), rho};
        launch_dual(da, it + 1, false);
        PDAL_HIP(hipEventRecord(ev[3], stream_));
        PDAL_LAUNCHED();
        PDAL_HIP(hipEventSynchronize(ev[3]));
        float a = 0, b = 0, c = 0;
        PDAL_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
        PDAL_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
        PDAL_HIP(hipEventElapsedTime(&c, ev[2], ev[3]));
        tpr += a;
        tp += b;
        td += c;
    }
    PDAL_HIP(hipEventRecord(ev[4], stream_));
    launch_eval();
    PDAL_HIP(hipEventRecord(ev[5], stream_));
    PDAL_HIP(hipEventSynchronize(ev[5]));
    float te = 0;
    PDAL_HIP(hipEventElapsedTime(&te, ev[4], ev[5]));
    out[0] = tp / iters;
    out[1] = td / iters;
    out[2] = pred ? tpr / iters : 0.0;
    out[3] = (tp + td + (pred ? tpr : 0.0)) / iters;
    out[4] = te;
    out[5] = sAt_.ready ? -sAt_.slices : At_.G;  // a negative lane count: the sliced pass with that many slices
    out[6] = sA_.ready ? -sA_.slices : A_.G;
    for (auto &e : ev) (void)hipEventDestroy(e);
}

}  // namespace

void reset_rng()
{
    g_gen.seed(1);
    g_dist.reset();
}

// ---------------------------------------------------------------------------
// Feasibility polishing (solver.cu:3912-4456, utils.cu:1194-1440).  Two
// phases after the main solve, each a restarted Halpern run of the same
// kernels on a feasibility problem: the primal phase with c = 0 (dual iterates
// from zero), the dual phase with every finite bound set to 0 (primal iterates
// from zero); each keeps the primal weight fixed, restarts only its own side
// and stops when its relative residual reaches eps_feas_polish_relative.

void Solver::launch_polish_block(int K)
{   // iterations 2..K (K major) + the evaluation pass, captured once per phase
    auto body = [&]() {
        for (int i = 2; i <= K; ++i) launch_iteration(i, i == K);
        launch_eval();
    };
    if (!use_graph_) {
        body();
        return;
    }
    if (!polish_graph_) {
        hipGraph_t graph;
        PDAL_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        body();
        PDAL_HIP(hipStreamEndCapture(stream_, &graph));
        PDAL_HIP(hipGraphInstantiate(&polish_graph_, graph, nullptr, nullptr, 0));
        PDAL_HIP(hipGraphDestroy(graph));
    }
    PDAL_HIP(hipGraphLaunch(polish_graph_, stream_));
}

Solver::PolishOut Solver::polish_phase(bool primal, double pw0, std::chrono::steady_clock::time_point t_start)
{   // primal_feasibility_polish / dual_feasibility_polish, solver.cu:3980-4127
    const auto t0 = std::chrono::steady_clock::now();
    const termination_criteria_t &tc = P_.termination_criteria;
    const int K = P_.termination_evaluation_frequency;
    const bool linf = P_.optimality_norm == NORM_TYPE_L_INF;
    pw_ = pw0;
    best_pw_ = pw0;
    total_ = 0;
    inner_ = 0;
    fp_ = fp0_ = last_trial_fp_ = INFINITY;
    if (polish_graph_) {
        PDAL_HIP(hipGraphExecDestroy(polish_graph_));
        polish_graph_ = nullptr;
    }
    if (P_.verbose) {  // print_initial_feas_polish_info, utils.cu:1229-1245
        printf("---------------------------------------------------------------------------------------\n");
        printf("Starting %s Feasibility Polishing Phase with relative tolerance %.2e\n", primal ? "Primal" : "Dual",
               tc.eps_feas_polish_relative);
        printf("---------------------------------------------------------------------------------------\n");
        printf("%s %s |  %s  | %s | %s \n", "  iter", "  time ", primal ? "pr obj" : "du obj",
               primal ? " abs pr res " : " abs du res ", primal ? " rel pr res " : " rel du res ");
        printf("---------------------------------------------------------------------------------------\n");
    }
    PolishOut out{TERMINATION_REASON_UNSPECIFIED, 0, 0.0, 0.0, 0.0, 0.0};
    auto phase_fp = [&](const Eval &e) { return primal ? e.pn2 * pw_ : e.dn2 / pw_; };
    bool do_restart = false;
    while (out.reason == TERMINATION_REASON_UNSPECIFIED) {
        sync_scalars();
        launch_iteration(1, true);
        if (do_restart) {  // fixed-point error after the block's first iteration
            launch_eval();
            fp_ = phase_fp(read_eval());
            fp0_ = fp_;
            do_restart = false;
        }
        launch_polish_block(K);
        const Eval e = read_eval();
        fp_ = phase_fp(e);
        const double bc = b_scale_ * c_scale_;
        if (primal) {  // compute_primal_feas_polish_residual, utils.cu:1341-1392
            out.abs = (linf ? e.presmax : std::sqrt(e.pres2)) / b_scale_;
            out.rel = out.abs / (1.0 + b_norm_);
            out.obj = e.cx / bc;  // the phase's objective constant is 0
        } else {  // compute_dual_feas_polish_residual, utils.cu:1394-1440
            out.abs = (linf ? e.dresmax : std::sqrt(e.dres2)) / c_scale_;
            out.rel = out.abs / (1.0 + c_norm_);
            out.obj = (e.lx + e.contrib) / bc + obj_const_;
        }
        inner_ += K;
        total_ += K;
        out.time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        // check_feas_polishing_termination_criteria, utils.cu:1194-1227
        if (out.rel <= tc.eps_feas_polish_relative)
            out.reason = TERMINATION_REASON_FEAS_POLISH_SUCCESS;
        else if (total_ >= tc.iteration_limit)
            out.reason = TERMINATION_REASON_ITERATION_LIMIT;
        else if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count() >=
                 tc.time_sec_limit)
            out.reason = TERMINATION_REASON_TIME_LIMIT;
        if (P_.verbose && total_ % print_frequency(total_) == 0)
            printf("%6d %.1e | %8.1e |    %.1e   |   %.1e   \n", total_, out.time, out.obj, out.abs, out.rel);
        do_restart = should_restart(K);
        if (do_restart) {  // perform_primal_restart / perform_dual_restart: anchor = current = T, own side only
            const size_t bytes = (size_t)(primal ? n_ : m_) * sizeof(double);
            double *T = primal ? xT_.get() : yT_.get();
            if (bytes) {
                PDAL_HIP(hipMemcpyAsync(primal ? x0_.get() : y0_.get(), T, bytes, hipMemcpyDeviceToDevice, stream_));
                PDAL_HIP(hipMemcpyAsync(primal ? x_.get() : y_.get(), T, bytes, hipMemcpyDeviceToDevice, stream_));
            }
            inner_ = 0;
            last_trial_fp_ = INFINITY;
        }
    }
    out.total = total_;
    if (polish_graph_) {
        PDAL_HIP(hipGraphExecDestroy(polish_graph_));
        polish_graph_ = nullptr;
    }
    return out;
}

void Solver::feasibility_polish(std::chrono::steady_clock::time_point t_start)
{   // feasibility_polish, solver.cu:3912-3978
    const auto t0 = std::chrono::steady_clock::now();
    const termination_criteria_t &tc = P_.termination_criteria;
    if (rel_p_ < tc.eps_feas_polish_relative && rel_d_ < tc.eps_feas_polish_relative) {
        printf("Skipping feasibility polishing as the solution is already sufficiently feasible.\n");
        fflush(stdout);
        return;
    }
    const double pw0 = P_.bound_objective_rescaling ? 1.0 : (c_norm_ + 1.0) / (b_norm_ + 1.0);
    if (graph_exec_) {  // the main block's graph holds the main views
        PDAL_HIP(hipGraphExecDestroy(graph_exec_));
        graph_exec_ = nullptr;
    }
    auto copy = [&](double *dst, const double *src, int len) {
        if (len > 0)
            PDAL_HIP(hipMemcpyAsync(dst, src, (size_t)len * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    };
    auto zero = [&](double *v, int len) {
        if (len > 0) fill_kernel<<<blocks_for(len), kBlock, 0, stream_>>>(v, 0.0, len);
    };
    // the main state's vectors the two phases start from or fall back to
    DBuf<double> sx(n_), sdslack(n_), sy0(m_), sy(m_), syT(m_), syH(m_), syR(m_);
    copy(sx.get(), xT_.get(), n_);
    copy(sdslack.get(), dslack_.get(), n_);
    copy(sy0.get(), y0_.get(), m_);
    copy(sy.get(), y_.get(), m_);
    copy(syT.get(), yT_.get(), m_);
    copy(syH.get(), yH_.get(), m_);
    copy(syR.get(), yR_.get(), m_);
    // the phases' problem data: c = 0; finite bounds = 0
    DBuf<double> zc(n_), zlb(n_), zub(n_), zl(m_), zu(m_);
    zero(zc.get(), n_);
    copy(zlb.get(), lb_.get(), n_);
    copy(zub.get(), ub_.get(), n_);
    copy(zl.get(), l_.get(), m_);
    copy(zu.get(), u_.get(), m_);
    for (auto [p, len] : {std::pair<double *, int>{zlb.get(), n_}, {zub.get(), n_}, {zl.get(), m_}, {zu.get(), m_}})
        if (len > 0) zero_finite_kernel<<<blocks_for(len), kBlock, 0, stream_>>>(p, len);
    PDAL_LAUNCHED();

    // primal phase: c = 0, dual iterates from zero, primal iterates continue
    for (double *v : {y0_.get(), y_.get(), yT_.get(), yH_.get(), yR_.get()}) zero(v, m_);
    zero(dslack_.get(), n_);
    iter_view_ = View{zc.get(), lb_.get(), ub_.get(), l_.get(), u_.get()};
    eval_view_ = View{c_.get(), lb_.get(), ub_.get(), l_.get(), u_.get()};
    eval_x_ = xT_.get();
    const PolishOut pr = polish_phase(true, pw0, t_start);
    if (pr.reason == TERMINATION_REASON_FEAS_POLISH_SUCCESS) {
        copy(sx.get(), xT_.get(), n_);
        abs_p_ = pr.abs;
        rel_p_ = pr.rel;
        p_obj_ = pr.obj;
    }
    polish_iters_ += pr.total - 1;

    // dual phase: finite bounds = 0, dual iterates from the main state's, primal from zero
    copy(y0_.get(), sy0.get(), m_);
    copy(y_.get(), sy.get(), m_);
    copy(yT_.get(), syT.get(), m_);
    copy(yH_.get(), syH.get(), m_);
    copy(yR_.get(), syR.get(), m_);
    copy(dslack_.get(), sdslack.get(), n_);
    for (double *v : {x0_.get(), x_.get(), xT_.get(), xH_.get(), xR_.get()}) zero(v, n_);
    iter_view_ = View{c_.get(), zlb.get(), zub.get(), zl.get(), zu.get()};
    eval_view_ = View{c_.get(), lb_.get(), ub_.get(), l_.get(), u_.get()};
    eval_x_ = sx.get();  // the dual objective pairs the slack with the (polished) primal solution
    const PolishOut du = polish_phase(false, pw0, t_start);
    if (du.reason == TERMINATION_REASON_FEAS_POLISH_SUCCESS) {
        abs_d_ = du.abs;
        rel_d_ = du.rel;
        d_obj_ = du.obj;
    } else {
        copy(yT_.get(), syT.get(), m_);
    }
    polish_iters_ += du.total - 1;
    copy(xT_.get(), sx.get(), n_);  // the result's x
    gap_ = std::fabs(p_obj_ - d_obj_);
    rel_gap_ = gap_ / (1.0 + std::fabs(p_obj_) + std::fabs(d_obj_));

    iter_view_ = eval_view_ = View{c_.get(), lb_.get(), ub_.get(), l_.get(), u_.get()};
    eval_x_ = xT_.get();
    PDAL_HIP(hipStreamSynchronize(stream_));
    // cp_al_feas_polish_final_log, utils.cu:1247-1268
    if (P_.verbose) printf("---------------------------------------------------------------------------------------\n");
    printf("Feasibility Polishing Summary\n");
    printf("  Primal Status        : %s\n", termination_reason_to_string(pr.reason));
    printf("  Primal Iterations    : %d\n", pr.total);
    printf("  Primal Time Usage    : %.3g sec\n", pr.time);
    printf("  Dual Status          : %s\n", termination_reason_to_string(du.reason));
    printf("  Dual Iterations      : %d\n", du.total);
    printf("  Dual Time Usage      : %.3g sec\n", du.time);
    printf("  Primal Residual      : %.3e\n", pr.rel);
    printf("  Dual Residual        : %.3e\n", du.rel);
    printf("  Primal Dual Gap      : %.3e\n",
           std::fabs(pr.obj - du.obj) / (1.0 + std::fabs(pr.obj) + std::fabs(du.obj)));
    fflush(stdout);
    polish_time_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

cp_al_result_t *optimize_gpu(const cp_al_parameters_t *params, lp_problem_t *problem)
{
    Solver s(params, problem);
    return s.run(problem);
}

int time_kernels_gpu(const cp_al_parameters_t *params, const lp_problem_t *problem, int iters, double *out)
{
    Solver s(params, problem);
    s.setup();
    s.reset_iterates();
    s.time_kernels(iters, out);
    return 0;
}

struct Session {
    cp_al_parameters_t params;
    lp_problem_t *problem;  // borrowed: must outlive the session
    Solver solver;
    Session(const cp_al_parameters_t *p, lp_problem_t *pr) : params(*p), problem(pr), solver(&params, pr) {}
};

void *session_create(const cp_al_parameters_t *params, lp_problem_t *problem)
{
    auto *s = new Session(params, problem);
    try {
        s->solver.setup();
    } catch (...) {
        delete s;
        throw;
    }
    return s;
}

cp_al_result_t *session_solve(void *h) { return static_cast<Session *>(h)->solver.solve_once(static_cast<Session *>(h)->problem); }

int session_time_kernels(void *h, int iters, double *out)
{
    auto *s = static_cast<Session *>(h);
    s->solver.time_kernels(iters, out);
    return 0;
}

double session_setup_seconds(void *h) { return static_cast<Session *>(h)->solver.setup_seconds(); }

void session_free(void *h) { delete static_cast<Session *>(h); }

__global__ void __launch_bounds__(kBlock) copy_calibration_kernel(const double *__restrict__ src,
                                                                  double *__restrict__ dst, long n)
{
    for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) dst[i] = src[i];
}

double copy_calibration(long long bytes, int reps)
{
    const long n = (long)(bytes / (long long)sizeof(double));
    if (n <= 0 || reps <= 0) return -1.0;
    DBuf<double> a(n), b(n);
    PDAL_HIP(hipMemset(a.get(), 0, (size_t)n * sizeof(double)));
    hipEvent_t e0, e1;
    PDAL_HIP(hipEventCreate(&e0));
    PDAL_HIP(hipEventCreate(&e1));
    const int grid = 256 * 16;
    copy_calibration_kernel<<<grid, kBlock>>>(a.get(), b.get(), n);
    PDAL_HIP(hipEventRecord(e0, nullptr));
    for (int r = 0; r < reps; ++r) copy_calibration_kernel<<<grid, kBlock>>>(a.get(), b.get(), n);
    PDAL_HIP(hipEventRecord(e1, nullptr));
    PDAL_LAUNCHED();
    PDAL_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    PDAL_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / reps;
}

__global__ void __launch_bounds__(kBlock) gather_index_kernel(int *__restrict__ idx, long count, long len)
{   // uniform indices in [0, len) from a 64-bit mix of the position (splitmix64 finalizer)
    for (long k = (long)blockIdx.x * kBlock + threadIdx.x; k < count; k += (long)gridDim.x * kBlock) {
        unsigned long long z = (unsigned long long)k + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        idx[k] = (int)(z % (unsigned long long)len);
    }
}

// U independent index loads and gathers in flight per thread per round (the
// grid-stride positions k, k + S, ..., k + (U-1) S), their sum kept so no load is
// dead; bounds checked per position.
template <int U>
__global__ void __launch_bounds__(kBlock) gather_calibration_kernel(const int *__restrict__ idx,
                                                                    const double *__restrict__ v, long count,
                                                                    double *__restrict__ out)
{
    const long stride = (long)gridDim.x * kBlock;
    double acc = 0.0;
    for (long k = (long)blockIdx.x * kBlock + threadIdx.x; k < count; k += U * stride) {
        int ix[U];
#pragma unroll
        for (int u = 0; u < U; ++u) ix[u] = k + u * stride < count ? idx[k + u * stride] : -1;
        double g[U];
#pragma unroll
        for (int u = 0; u < U; ++u) g[u] = ix[u] >= 0 ? v[ix[u]] : 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u) acc += g[u];
    }
    out[(long)blockIdx.x * kBlock + threadIdx.x] = acc;
}

double gather_calibration(long long count, long long len, int reps)
{   // the fastest of 1, 4 and 8 gathers in flight per thread over three grid sizes
    if (count <= 0 || len <= 0 || len > INT32_MAX || reps <= 0) return -1.0;
    const int grids[3] = {256 * 4, 256 * 16, 256 * 64};
    DBuf<int> idx(count);
    DBuf<double> v(len), out((size_t)grids[2] * kBlock);
    PDAL_HIP(hipMemset(v.get(), 0, (size_t)len * sizeof(double)));
    gather_index_kernel<<<256 * 16, kBlock>>>(idx.get(), count, len);
    hipEvent_t e0, e1;
    PDAL_HIP(hipEventCreate(&e0));
    PDAL_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    for (int grid : grids) {
        for (int u : {1, 4, 8}) {
            auto launch = [&]() {
                if (u == 1) gather_calibration_kernel<1><<<grid, kBlock>>>(idx.get(), v.get(), count, out.get());
                else if (u == 4) gather_calibration_kernel<4><<<grid, kBlock>>>(idx.get(), v.get(), count, out.get());
                else gather_calibration_kernel<8><<<grid, kBlock>>>(idx.get(), v.get(), count, out.get());
            };
            launch();
            PDAL_HIP(hipEventRecord(e0, nullptr));
            for (int r = 0; r < reps; ++r) launch();
            PDAL_HIP(hipEventRecord(e1, nullptr));
            PDAL_LAUNCHED();
            PDAL_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            PDAL_HIP(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, ms / reps);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

}  // namespace lp
}  // namespace pdal
