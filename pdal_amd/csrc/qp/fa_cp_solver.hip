// FA_CP optimize(): host driver of the MI355X QP path.
//
// Reference: QP/FA_CP/src/solver.cu:74-168 (main loop), preconditioner.c:36-331
// (scaling, here on the device), solver_state.cu:220-877 (state, Q handling,
// inner solver), utils.cu:65-315 (power iterations), :438-537 (termination,
// restart, defaults), :705-1132 (residual, infeasibility), fa_cp_core_op.cu:
// 93-1239 (iteration kernels, BB inner solve, Halpern, restart, sigma rules,
// fixed-point error, result).
//
// MI355X design: one HIP stream, FP64 throughout.  An iteration without an
// inner solve is two launches: the At.y SpMV fused with the LP / diagonal-Q
// prox and the A.x_bar SpMV fused with the dual update, each also applying the
// Halpern step when no fixed-point evaluation needs the pre-Halpern point.
// The projected-BB inner solve runs device-resident: its step size, direction
// norm and stopping test live in a BBState on the device, every inner kernel
// is gated on the state's done flag, and the host launches inner steps in
// chunks and reads one flag per chunk.  Checkpoints (residual, infeasibility,
// restart, fixed-point error) are SpMVs with fused deterministic reductions;
// the host reads a few doubles and runs the scalar logic.
//
// Built with -DPDAL_LIN_CP_AL the same driver is the linearized CP-AL solver
// RHR-Lin-CP-AL (QP/LIN_CP_AL): the shared upload, scaling, power iterations,
// residual, infeasibility and termination; its own step sizes
// (cp_al_core_op.cu:609-691), update (:496-575, one projected gradient step),
// relaxed Halpern step (:577-585), per-iteration fixed-point error (:693-770),
// restart test and restart (utils.cu:451-463, :595-607), main loop
// (solver.cu:34-114); feasibility polishing runs FA_CP's phases with this
// iteration (cp_al_core_op.cu:839-1375).
//
// Built with -DPDAL_PLANTED_FA_CP it is RHR-FA-CP, the planted-KKT FA_CP
// (QP/PLANTED_FA_CP): this iteration with equality AL kernels, the AL branch
// only when al_sigma != 0, its inner-loop rule, the fixed-point error without
// the clamp at zero, and its own AL sigma rule at restarts
// (fa_cp_core_op.cu:610-747) instead of FA_CP's; its feasibility polishing is
// FA_CP's code (fa_cp_core_op.cu:1009-1502).
//
// Built with -DPDAL_PLANTED_CP_AL it is RHR-CP-AL, the planted-KKT CP-AL
// (QP/PLANTED_CP_AL): FA_CP's main loop, restart test and Halpern step over the
// linearized solver's sigma / lambda steps, with its own set-up
// (cp_al_core_op.cu:805-858), update (:531-635: equality rows, LP and diagonal Q
// in closed form, FA_CP's BB inner solve with the effective dual fixed for the
// other Q types, T-outputs in major-kernel iterations only), fixed-point error
// on the reflection (:860-935), and at restarts a sigma controller in the AL
// branch or the primal-weight PID (:655-803).
#include "fa_cp_internal.hpp"

#include "../common/csr_host.hpp"
#include "../common/hip_util.hpp"
#include "fa_cp_kernels.hpp"
#include "lin_cp_al_kernels.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <utility>
#include <vector>

namespace pdal {
namespace qp {

namespace {

std::mt19937 g_gen(1);                              // QP/FA_CP/src/utils.cu:28 (process-global)
std::normal_distribution<double> g_dist(0.0, 1.0);  // utils.cu:29

constexpr int kMaxPartBlocks = 2048;

inline int stride_grid(long len) { return std::max(1, std::min(1024, blocks_for(len))); }
inline int row_grid(long rows, int G) { return std::max(1, std::min(kMaxPartBlocks, blocks_for(rows * G))); }

int print_frequency(int iter)  // utils.cu:690-703
{
    int step = 10;
    long long threshold = 1000;
    while (iter >= threshold) {
        step *= 10;
        threshold *= 10;
    }
    return step;
}

double sane_positive(double v, double fb) { return (std::isfinite(v) && v > 0.0) ? v : fb; }
double score_component(double v) { return std::isfinite(v) ? std::fmax(v, 1e-300) : INFINITY; }
double clampd_h(double v, double lo, double hi) { return std::fmin(std::fmax(v, lo), hi); }

HostCsr host_component(const CsrComponent *c, int rows, int cols, int nnz)
{
    HostCsr M;
    M.rows = rows;
    M.cols = cols;
    M.rp.assign((size_t)rows + 1, 0);
    if (c && c->row_ptr && nnz > 0) {
        M.rp.assign(c->row_ptr, c->row_ptr + rows + 1);
        M.ci.assign(c->col_ind, c->col_ind + nnz);
        M.v.assign(c->val, c->val + nnz);
    }
    return M;
}

// Scalar slots of the device reduction array.
enum Slot {
    S_RA = 0,                 // RA_N = 3
    S_RT = S_RA + RA_N,       // RT_N = 9
    S_FP = S_RT + RT_N,       // cross, pn2
    S_DN2 = S_FP + FP_N,      // ||dy||^2
    S_PROJ = S_DN2 + 1,       // max |dx|, max |dy|
    S_RAY = S_PROJ + 2,       // RY_N = 5
    S_QXMAX = S_RAY + RY_N,   // max |Q dx|
    S_PD2 = S_QXMAX + 1,
    S_DD2 = S_PD2 + 1,
    S_NORMAL = S_DD2 + 1,
    S_PW0 = S_NORMAL + 1,     // power iteration: sumsq, <v,Mv>, (unused), resid
    S_BOUND = S_PW0 + 4,      // bound-objective sums
    S_SEL = S_BOUND + 2,      // restart candidate: ||dx||^2, ||dy||^2, <At dy, dx>
    S_COUNT = S_SEL + 3
};

class Solver {
  public:
    Solver(const params_t &p, const qp_problem_t *w) : P_(p), W_(w) {}
    ~Solver()
    {
        if (bb_graph_) (void)hipGraphExecDestroy(bb_graph_);
        if (stream_) (void)hipStreamDestroy(stream_);
    }
    result_t *run();

  private:
    // ---- set-up ----
    void upload();
    void rescale();
    void scale_all(const double *e, const double *d);
    void init_state();
    double max_singular_value();
    double q_max_eig();
    double q_min_eig(double lmax);

    // ---- device helpers ----
    void reduce(int nparts, int nsum, int nq, int slot)
    {
        reduce_partials_kernel<<<1, kBlock, 0, stream_>>>(parts_.get(), nparts, kEvalStride, nsum, nq,
                                                          dscal_.get() + slot);
        PDAL_LAUNCHED();
    }
    void fetch(int lo, int cnt)
    {
        PDAL_HIP(hipMemcpyAsync(hscal_.get() + lo, dscal_.get() + lo, (size_t)cnt * sizeof(double),
                                hipMemcpyDeviceToHost, stream_));
        PDAL_HIP(hipStreamSynchronize(stream_));
    }
    double hs(int slot) { return hscal_[slot]; }
    void spmv(const DevCsr &M, const double *x, double *y, const BBState *gate = nullptr)
    {
        dispatch_group(M.G, [&](auto g) {
            constexpr int G = decltype(g)::value;
            gated_spmv_kernel<G><<<blocks_for((long)M.rows * G), kBlock, 0, stream_>>>(M.view(), x, y, gate);
        });
        PDAL_LAUNCHED();
    }
    void obj_product(const double *xv);
    void copy(double *dst, const double *src, int len)
    {
        if (len > 0)
            PDAL_HIP(hipMemcpyAsync(dst, src, (size_t)len * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    }

    // ---- iteration ----
    StepScalars scalars() const;
    bool major_kernel() const { return major_ || ((total_ + 2) % print_frequency(total_ + 2)) == 0; }
    void update(bool fuse_halpern);
    void primal_bb(const StepScalars &sc, bool fuse_halpern);
    void bb_inner_step(const StepScalars &sc, bool test_first);
    void bb_loop_test();
    void halpern();
    void fixed_point_error();
    void compute_residual();
    void infeasibility_info();
    void check_termination();
    bool should_restart();
#if !defined(PDAL_CP_AL_TYPES) && !defined(PDAL_PLANTED_FA_CP)
    void choose_restart_candidate();  // fa_cp_core_op.cu:1130-1169
    int restart_target_ = 0;          // 0: the T output, 1: the current iterate
#endif
    void set_sigma_active(double sigma, double zero_thr);
    double clipped_update(double cur, double ratio) const;
    void update_al_sigma(double pd, double dd, double normal);
    void perform_restart();
    // planted-KKT FA_CP (PDAL_PLANTED_FA_CP builds)
    int last_sigma_update_iter_ = 0;
    double inner_tol0_ = 0.0;
    void planted_sigma_on_restart();
    // linearized CP-AL (PDAL_LIN_CP_AL builds)
    double sigma_ = 0.0, sigma_base_ = 1.0, sigma_min_ = 0.0, sigma_max_ = 0.0, lambda_est_ = 1.0;
    double lin_tau() const { return 1.0 / std::fmax(sigma_ * lambda_est_, 1e-16); }
    void lin_init_steps();
    void lin_update();
    void lin_fixed_point_error();
    bool lin_should_restart() const;
    void lin_restart();
    void lin_halpern();
    result_t *run_linearized();
    // planted-KKT CP-AL (PDAL_PLANTED_CP_AL builds)
    double best_sigma_ = 0.0, best_term_metric_ = INFINITY, best_sigma_gap_ = INFINITY;
    double pcal_tau() const
    {   // cpal_effective_primal_step_size, PLANTED_CP_AL cp_al_core_op.cu:255-265
        double t = lin_tau();
        if (nonconvexity_ < 0) {
            t = std::fmax(t, -1.01 * std::fmin(0.0, nonconvexity_));
            t /= 2;
        }
        return t;
    }
    void pcal_init_steps();
    void pcal_update(bool fuse_halpern);
    void pcal_fixed_point_error();
    void pcal_reset_sigma();
    void pcal_sigma_controller(double pd, double dd, double ax);
    void pid_primal_weight(double pd, double dd);
    // the BB inner solve recomputes the AL effective dual at each inner step
    // (FA_CP's full AL subproblem) except in the planted CP-AL, whose BB keeps
    // the product fixed at the outer iterate
#ifdef PDAL_PLANTED_CP_AL
    bool bb_al() const { return false; }
#else
    bool bb_al() const { return use_al_; }
#endif
    result_t *make_result();
    void print_initial_info() const;
    void display_stats() const;
#ifndef PDAL_LIN_CP_AL
    bool trajectory_open() const;             // initialize_trajectory_csv (solver.cu:35-48)
    bool trajectory_row(bool restart) const;  // append_trajectory_row (solver.cu:50-71)
#endif
    // feasibility polishing (fa_cp_core_op.cu:1242-1735)
    struct PolishOut {
        termination_reason_t reason;
        int total;
        double abs, rel, obj, time;
    };
    struct PolishRefs {  // the main problem's data the phases' residuals read
        const double *c, *l, *u, *x;
    };
    void feasibility_polish();
    PolishOut polish_phase(bool primal, double pw0, const PolishRefs &ref);
    void polish_residual(bool primal, const PolishRefs &ref, PolishOut &out);
    double polish_fixed_point_error(bool primal);
    void drop_bb_graph()
    {
        if (bb_graph_) PDAL_HIP(hipGraphExecDestroy(bb_graph_));
        bb_graph_ = nullptr;
        g_tau_ = g_al_ = g_tol_ = NAN;
        g_limit_ = -1;
    }
    double polish_time_ = 0.0;
    int polish_iters_ = 0;

    params_t P_;
    const qp_problem_t *W_;
    hipStream_t stream_ = nullptr;

    int n_ = 0, m_ = 0, r_ = 0;
    double obj_const_ = 0.0;
    HostCsr hQ_;  // structure of Q for the type test
    DevCsr A_, At_, Q_, Qt_, R_, Rt_;
    quad_obj_type_t qtype_ = FA_CP_NON_Q;
    bool is_lp_ = true;
    DBuf<double> c_, lb_, ub_, l_, u_, E_, D_, qdiag_;
    double b_scale_ = 1.0, c_scale_ = 1.0, c_norm_ = 0.0, b_norm_ = 0.0, q_norm_ = 0.0, nonconvexity_ = 0.0;

    DBuf<double> x0_, x_, xT_, xR_, dual_prod_, dslack_, dx_, qx_, grad_, dir_;
    DBuf<double> y0_, y_, yT_, yR_, prim_prod_, al_u_, pslack_, dy_, rx_;
    DBuf<double> parts_, bbparts_, bbparts_g_, dscal_, scratch_e_, scratch_d_;
    PinnedBuf<double> hscal_{S_COUNT};
    DBuf<BBState> dbb_;
    PinnedBuf<BBState> hbb_{1};

    // inner solver
    bool has_inner_ = false;
    int inner_limit_ = 0;
    double inner_tol_ = 0.0;
    int bb_chunk_ = 4;
    // hipGraph of kGraphSteps gated inner steps, re-captured when its baked
    // arguments (tau, AL penalty, tolerance, limit, AL branch) change
    static constexpr int kGraphSteps = 8;
    bool use_graph_ = true;
    hipGraphExec_t bb_graph_ = nullptr;
    double g_tau_ = NAN, g_al_ = NAN, g_tol_ = NAN;
    int g_limit_ = -1;
    bool g_use_al_ = false;
    void launch_inner_steps(const StepScalars &sc, int steps);

    // AL sigma state (fa_cp_core_op.cu:642-906)
    bool al_available_ = false, use_al_ = false;
    double al_sigma_ = 0.0, best_al_sigma_ = 0.0, best_sigma_score_ = INFINITY, last_sigma_score_ = INFINITY;
    double last_sigma_rp_ = INFINITY, last_sigma_rd_ = INFINITY, last_sigma_gap_ = INFINITY;
    int sigma_update_count_ = 0, sigma_dir_ = 1;

    double step_ = 0.0, pw_ = 1.0, best_pw_ = 1.0, pw_err_sum_ = 0.0, pw_last_err_ = 0.0, best_gap_ = INFINITY;
    int total_ = 0, inner_ = 0;
    bool major_ = false;
    double fp_ = INFINITY, fp0_ = INFINITY, last_trial_fp_ = INFINITY;
    double abs_p_ = 0, rel_p_ = 0, abs_d_ = 0, rel_d_ = 0, p_obj_ = 0, d_obj_ = 0, gap_ = 0, rel_gap_ = 0;
    double max_p_ray_ = 0, max_d_ray_ = 0, p_ray_obj_ = 0, d_ray_obj_ = 0;
    termination_reason_t reason_ = TERMINATION_REASON_UNSPECIFIED;
    double rescale_time_ = 0.0, solve_time_ = 0.0;
};

// ---------------------------------------------------------------------------
// Set-up

void Solver::upload()
{
    const qp_problem_t *w = W_;
    n_ = w->num_variables;
    m_ = w->num_constraints;
    r_ = w->num_rank_lowrank_obj;
    obj_const_ = w->objective_constant;
    PDAL_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    use_graph_ = !(getenv("PDAL_NO_GRAPH") && atoi(getenv("PDAL_NO_GRAPH")) != 0);

    HostCsr hA = host_component(w->constraint_matrix, m_, n_, w->constraint_matrix_num_nonzeros);
    hQ_ = host_component(w->objective_sparse_matrix, n_, n_, w->objective_sparse_matrix_num_nonzeros);
    HostCsr hR = host_component(w->objective_lowrank_matrix, r_, n_, w->objective_lowrank_matrix_num_nonzeros);
    HostCsr hAt = transpose_csr(m_, n_, hA.rp.data(), hA.ci.data(), hA.v.data());
    HostCsr hQt = transpose_csr(n_, n_, hQ_.rp.data(), hQ_.ci.data(), hQ_.v.data());
    HostCsr hRt = transpose_csr(r_, n_, hR.rp.data(), hR.ci.data(), hR.v.data());
    auto put = [&](DevCsr &Dm, const HostCsr &H, const char *env) {
        Dm.rows = H.rows;
        Dm.nnz = H.nnz();
        Dm.rp.alloc((size_t)H.rows + 1);
        Dm.ci.alloc(Dm.nnz);
        Dm.v.alloc(Dm.nnz);
        Dm.rp.upload(H.rp.data(), (size_t)H.rows + 1, stream_);
        Dm.ci.upload(H.ci.data(), Dm.nnz, stream_);
        Dm.v.upload(H.v.data(), Dm.nnz, stream_);
        Dm.G = pick_group(env, Dm.nnz, H.rows);
    };
    put(A_, hA, "PDAL_QP_G_A");
    put(At_, hAt, "PDAL_QP_G_AT");
    put(Q_, hQ_, "PDAL_QP_G_Q");
    put(Qt_, hQt, nullptr);
    put(R_, hR, "PDAL_QP_G_R");
    put(Rt_, hRt, "PDAL_QP_G_RT");
    PDAL_HIP(hipStreamSynchronize(stream_));  // host matrices go out of scope

    auto vec = [&](DBuf<double> &b, const double *src, int len) {
        b.alloc(len);
        b.upload(src, len, stream_);
    };
    vec(c_, w->objective_vector, n_);
    vec(lb_, w->variable_lower_bound, n_);
    vec(ub_, w->variable_upper_bound, n_);
    vec(l_, w->constraint_lower_bound, m_);
    vec(u_, w->constraint_upper_bound, m_);
    E_.alloc(m_);
    D_.alloc(n_);
    scratch_e_.alloc(m_);
    scratch_d_.alloc(n_);
    parts_.alloc((size_t)kMaxPartBlocks * kEvalStride);
    bbparts_.alloc(kMaxPartBlocks);
    bbparts_g_.alloc(kMaxPartBlocks);
    dscal_.alloc(S_COUNT);
    dbb_.alloc(1);
    PDAL_HIP(hipMemsetAsync(dscal_.get(), 0, S_COUNT * sizeof(double), stream_));
    PDAL_HIP(hipMemsetAsync(dbb_.get(), 0, sizeof(BBState), stream_));
    for (DBuf<double> *v : {&x0_, &x_, &xT_, &xR_, &dual_prod_, &dslack_, &dx_, &qx_, &grad_, &dir_}) {
        v->alloc(n_);
        PDAL_HIP(hipMemsetAsync(v->get(), 0, (size_t)std::max(n_, 1) * sizeof(double), stream_));
    }
    for (DBuf<double> *v : {&y0_, &y_, &yT_, &yR_, &prim_prod_, &al_u_, &pslack_, &dy_}) {
        v->alloc(m_);
        PDAL_HIP(hipMemsetAsync(v->get(), 0, (size_t)std::max(m_, 1) * sizeof(double), stream_));
    }
    rx_.alloc(r_);
    PDAL_HIP(hipMemsetAsync(rx_.get(), 0, (size_t)std::max(r_, 1) * sizeof(double), stream_));
}

void Solver::scale_all(const double *e, const double *d)
{   // scale_problem, preconditioner.c:92-138, plus the transposed copies
    scale_vectors_kernel<<<blocks_for((long)n_ + m_), kBlock, 0, stream_>>>(c_.get(), lb_.get(), ub_.get(), d,
                                                                            D_.get(), n_, l_.get(), u_.get(), e,
                                                                            E_.get(), m_);
    divide_rows_kernel<0><<<blocks_for(m_), kBlock, 0, stream_>>>(A_.view(), A_.v.get(), e, d);
    divide_rows_kernel<0><<<blocks_for(n_), kBlock, 0, stream_>>>(At_.view(), At_.v.get(), d, e);
    if (Q_.nnz > 0) {
        divide_rows_kernel<0><<<blocks_for(n_), kBlock, 0, stream_>>>(Q_.view(), Q_.v.get(), d, d);
        divide_rows_kernel<0><<<blocks_for(n_), kBlock, 0, stream_>>>(Qt_.view(), Qt_.v.get(), d, d);
    }
    if (R_.nnz > 0) {
        divide_rows_kernel<1><<<blocks_for(r_), kBlock, 0, stream_>>>(R_.view(), R_.v.get(), nullptr, d);
        divide_rows_kernel<2><<<blocks_for(n_), kBlock, 0, stream_>>>(Rt_.view(), Rt_.v.get(), d, nullptr);
    }
    PDAL_LAUNCHED();
}

void Solver::rescale()
{   // rescale_problem, preconditioner.c:291-331
    fill_kernel<<<blocks_for(m_), kBlock, 0, stream_>>>(E_.get(), 1.0, m_);
    fill_kernel<<<blocks_for(n_), kBlock, 0, stream_>>>(D_.get(), 1.0, n_);
    double *e = scratch_e_.get(), *d = scratch_d_.get();
    const int rows = std::max(m_, n_);
    for (int it = 0; it < P_.l_inf_ruiz_iterations; ++it) {  // ruiz_rescaling, :140-203
        ruiz_factors_kernel<<<blocks_for(rows), kBlock, 0, stream_>>>(A_.view(), At_.view(), Qt_.view(), e, d);
        PDAL_LAUNCHED();
        scale_all(e, d);
    }
    if (P_.has_pock_chambolle_alpha) {  // pock_chambolle_rescaling, :205-247
        pc_factors_kernel<<<blocks_for(rows), kBlock, 0, stream_>>>(A_.view(), At_.view(), Qt_.view(),
                                                                    P_.pock_chambolle_alpha, e, d);
        PDAL_LAUNCHED();
        scale_all(e, d);
    }
    if (P_.bound_objective_rescaling) {  // bound_obj_rescaling, :249-289
        const int nb = stride_grid(std::max(m_, n_));
        bound_obj_sums_kernel<<<nb, kBlock, 0, stream_>>>(l_.get(), u_.get(), m_, c_.get(), n_, parts_.get());
        PDAL_LAUNCHED();
        reduce(nb, 2, 2, S_BOUND);
        fetch(S_BOUND, 2);
        b_scale_ = 1.0 / (std::sqrt(hs(S_BOUND)) + 1.0);
        c_scale_ = 1.0 / (std::sqrt(hs(S_BOUND + 1)) + 1.0);
        bound_obj_apply_kernel<<<blocks_for((long)m_ + n_), kBlock, 0, stream_>>>(
            l_.get(), u_.get(), m_, lb_.get(), ub_.get(), c_.get(), n_, b_scale_, c_scale_);
        const double fq = c_scale_ / b_scale_;
        if (Q_.nnz > 0) {
            mul_all_kernel<<<blocks_for(Q_.nnz), kBlock, 0, stream_>>>(Q_.v.get(), Q_.nnz, fq);
            mul_all_kernel<<<blocks_for(Qt_.nnz), kBlock, 0, stream_>>>(Qt_.v.get(), Qt_.nnz, fq);
        }
        if (R_.nnz > 0) {
            const double fr = std::sqrt(c_scale_ / b_scale_);
            mul_all_kernel<<<blocks_for(R_.nnz), kBlock, 0, stream_>>>(R_.v.get(), R_.nnz, fr);
            mul_all_kernel<<<blocks_for(Rt_.nnz), kBlock, 0, stream_>>>(Rt_.v.get(), Rt_.nnz, fr);
        }
        PDAL_LAUNCHED();
    }
    PDAL_HIP(hipStreamSynchronize(stream_));
}

double Solver::max_singular_value()
{   // estimate_maximum_singular_value, utils.cu:65-170
    std::vector<double> v0(m_);
    for (int i = 0; i < m_; ++i) v0[i] = g_dist(g_gen);
    double *v = al_u_.get(), *nv = dy_.get(), *w = dx_.get();  // free before the iteration starts
    PDAL_HIP(hipMemcpyAsync(v, v0.data(), (size_t)m_ * sizeof(double), hipMemcpyHostToDevice, stream_));
    const int nbm = stride_grid(m_);
    double sig = 1.0;
    for (int it = 0; it < P_.sv_max_iter; ++it) {
        dot_sumsq_kernel<<<nbm, kBlock, 0, stream_>>>(v, v, m_, parts_.get());
        reduce(nbm, 2, 2, S_PW0);
        normalize_dev_kernel<<<blocks_for(m_), kBlock, 0, stream_>>>(v, nv, m_, dscal_.get() + S_PW0);
        spmv(At_, nv, w);
        spmv(A_, w, v);
        dot_sumsq_kernel<<<nbm, kBlock, 0, stream_>>>(nv, v, m_, parts_.get());
        reduce(nbm, 2, 2, S_PW0 + 1);
        sv_resid_kernel<<<nbm, kBlock, 0, stream_>>>(nv, v, m_, dscal_.get() + S_PW0 + 1, parts_.get());
        reduce(nbm, 1, 1, S_PW0 + 3);
        PDAL_LAUNCHED();
        fetch(S_PW0, 4);
        sig = hs(S_PW0 + 1);
        if (std::sqrt(hs(S_PW0 + 3)) < P_.sv_tol) break;
    }
    PDAL_HIP(hipMemsetAsync(al_u_.get(), 0, (size_t)std::max(m_, 1) * sizeof(double), stream_));
    PDAL_HIP(hipMemsetAsync(dy_.get(), 0, (size_t)std::max(m_, 1) * sizeof(double), stream_));
    PDAL_HIP(hipMemsetAsync(dx_.get(), 0, (size_t)std::max(n_, 1) * sizeof(double), stream_));
    return std::sqrt(sig);
}

double Solver::q_max_eig()
{   // estimate_maximum_eigenvalue, utils.cu:172-238 (glibc rand() start)
    std::vector<double> v0(n_);
    for (int i = 0; i < n_; ++i) v0[i] = (double)rand() / RAND_MAX;
    double *v = grad_.get(), *av = dir_.get();
    PDAL_HIP(hipMemcpyAsync(v, v0.data(), (size_t)n_ * sizeof(double), hipMemcpyHostToDevice, stream_));
    const int nb = stride_grid(n_);
    double lam = 0.0;
    for (int it = 0; it < P_.sv_max_iter; ++it) {
        dot_sumsq_kernel<<<nb, kBlock, 0, stream_>>>(v, v, n_, parts_.get());
        reduce(nb, 2, 2, S_PW0);
        normalize_dev_kernel<<<blocks_for(n_), kBlock, 0, stream_>>>(v, v, n_, dscal_.get() + S_PW0);
        spmv(Q_, v, av);
        dot_sumsq_kernel<<<nb, kBlock, 0, stream_>>>(v, av, n_, parts_.get());
        reduce(nb, 2, 2, S_PW0 + 1);
        PDAL_LAUNCHED();
        fetch(S_PW0 + 1, 1);
        const double old = lam;
        lam = hs(S_PW0 + 1);
        if (it > 0 && std::fabs(lam - old) < P_.sv_tol) break;
        std::swap(v, av);
    }
    return lam;
}

double Solver::q_min_eig(double lmax)
{   // estimate_minimum_eigenvalue, utils.cu:240-315
    std::vector<double> v0(n_);
    for (int i = 0; i < n_; ++i) v0[i] = (double)rand() / RAND_MAX;
    double *v = grad_.get(), *av = dir_.get(), *sh = dx_.get();
    PDAL_HIP(hipMemcpyAsync(v, v0.data(), (size_t)n_ * sizeof(double), hipMemcpyHostToDevice, stream_));
    const int nb = stride_grid(n_);
    double mu = 0.0;
    for (int it = 0; it < P_.sv_max_iter; ++it) {
        dot_sumsq_kernel<<<nb, kBlock, 0, stream_>>>(v, v, n_, parts_.get());
        reduce(nb, 2, 2, S_PW0);
        normalize_dev_kernel<<<blocks_for(n_), kBlock, 0, stream_>>>(v, v, n_, dscal_.get() + S_PW0);
        spmv(Q_, v, av);
        shift_dot_kernel<<<nb, kBlock, 0, stream_>>>(v, av, lmax, sh, n_, parts_.get());
        reduce(nb, 1, 1, S_PW0 + 1);
        PDAL_LAUNCHED();
        fetch(S_PW0 + 1, 1);
        const double old = mu;
        mu = hs(S_PW0 + 1);
        if (it > 0 && std::fabs(mu - old) < P_.sv_tol) break;
        std::swap(v, sh);
    }
    for (double *b : {grad_.get(), dir_.get(), dx_.get()})
        PDAL_HIP(hipMemsetAsync(b, 0, (size_t)std::max(n_, 1) * sizeof(double), stream_));
    return lmax - mu;
}

void Solver::init_state()
{   // initialize_solver_state, solver_state.cu:494-877
    const qp_problem_t *w = W_;
    al_available_ = P_.use_al_qp;
    use_al_ = P_.use_al_qp;
#if defined(PDAL_PLANTED_FA_CP)
    // planted solver_state.cu:498-504: the AL branch only with a nonzero sigma
    use_al_ = P_.use_al_qp && P_.al_sigma != 0.0;
    al_sigma_ = use_al_ ? P_.al_sigma : 0.0;
#elif !defined(PDAL_CP_AL_TYPES)
    al_sigma_ = P_.al_sigma;
    best_al_sigma_ = P_.al_sigma;
#endif
    if (w->primal_start || w->dual_start) {  // solver_state.cu:630-657
        if (w->primal_start) x_.upload(w->primal_start, n_, stream_);
        if (w->dual_start) y_.upload(w->dual_start, m_, stream_);
        start_scale_kernel<<<blocks_for((long)n_ + m_), kBlock, 0, stream_>>>(
            x_.get(), D_.get(), b_scale_, w->primal_start ? n_ : 0, y_.get(), E_.get(), c_scale_,
            w->dual_start ? m_ : 0);
        PDAL_LAUNCHED();
        copy(x0_.get(), x_.get(), n_);
        copy(xT_.get(), x_.get(), n_);
        copy(y0_.get(), y_.get(), m_);
        copy(yT_.get(), y_.get(), m_);
    }
    const bool linf = P_.optimality_norm == NORM_TYPE_L_INF;
    double sq = 0.0, mx = 0.0;
    for (int i = 0; i < n_; ++i) {
        const double cv = w->objective_vector[i];
        if (linf) mx = std::max(mx, std::fabs(cv));
        else sq += cv * cv;
    }
    c_norm_ = linf ? mx : std::sqrt(sq);
    sq = mx = 0.0;
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

    // detect_q_type, utils.cu:1576-1616
    if (Q_.nnz == 0 && R_.nnz == 0) qtype_ = FA_CP_NON_Q;
    else if (R_.nnz > 0) qtype_ = Q_.nnz > 0 ? FA_CP_LOW_RANK_PLUS_SPARSE_Q : FA_CP_LOW_RANK_Q;
    else {
        qtype_ = FA_CP_DIAG_Q;
        for (int i = 0; i < n_ && qtype_ == FA_CP_DIAG_Q; ++i)
            for (int k = hQ_.rp[i]; k < hQ_.rp[i + 1]; ++k)
                if (hQ_.ci[k] != i) {
                    qtype_ = FA_CP_SPARSE_Q;
                    break;
                }
    }
    hQ_ = HostCsr{};
    if (qtype_ == FA_CP_DIAG_Q) {  // solver_state.cu:462-492
        qdiag_.alloc(n_);
        PDAL_HIP(hipMemsetAsync(qdiag_.get(), 0, (size_t)std::max(n_, 1) * sizeof(double), stream_));
        qdiag_kernel<<<blocks_for(n_), kBlock, 0, stream_>>>(Q_.view(), n_, qdiag_.get());
        PDAL_LAUNCHED();
        std::vector<double> qd(n_);
        qdiag_.download(qd.data(), n_, stream_);
        PDAL_HIP(hipStreamSynchronize(stream_));
        double mxe = 0.0, mne = 0.0;
        for (double dv : qd) {
            mxe = std::fmax(mxe, std::fabs(dv));
            mne = std::fmin(mne, dv);
        }
        q_norm_ = mxe;
        nonconvexity_ = mne;
    } else if (qtype_ == FA_CP_SPARSE_Q) {
        q_norm_ = q_max_eig();
        nonconvexity_ = q_min_eig(q_norm_);
    }
#if defined(PDAL_PLANTED_FA_CP)
    // planted solver_state.cu:305-345: the inner loop for non-diagonal Q, and for diagonal Q in the AL branch
    const bool inner_loop = qtype_ != FA_CP_NON_Q && (qtype_ != FA_CP_DIAG_Q || use_al_);
    has_inner_ = inner_loop;
    if (inner_loop) {
        inner_limit_ = 1000;
        inner_tol_ = inner_tol0_ = 1e-3;
    }
#elif defined(PDAL_CP_AL_TYPES)
    // LIN_CP_AL / PLANTED_CP_AL solver_state.cu:305-340: an inner loop only for the
    // non-diagonal Q types (the linearized update never runs it; the planted one's BB does)
    const bool simple = qtype_ == FA_CP_NON_Q || qtype_ == FA_CP_DIAG_Q;
    has_inner_ = !simple;
    if (!simple) {
        inner_limit_ = 1000;
        inner_tol_ = 1e-3;
    }
#else
    if (al_available_ && P_.al_sigma <= P_.sigma_zero_threshold) {
        use_al_ = false;
        al_sigma_ = 0.0;
    }
    // initialize_inner_solver, solver_state.cu:305-360
    has_inner_ = use_al_ || al_available_;
    const bool simple = qtype_ == FA_CP_NON_Q || qtype_ == FA_CP_DIAG_Q;
    if (!simple) has_inner_ = true;
    if (!simple || al_available_) {
        inner_limit_ = 1000;
        inner_tol_ = 1e-3;
    }
#endif
    is_lp_ = qtype_ == FA_CP_NON_Q;
}

// ---------------------------------------------------------------------------
// Products and iteration

void Solver::obj_product(const double *xv)
{   // update_obj_product, solver_state.cu:371-442
    switch (qtype_) {
    case FA_CP_SPARSE_Q:
        dispatch_group(Q_.G, [&](auto g) {
            constexpr int G = decltype(g)::value;
            qx_kernel<G, 0><<<blocks_for((long)n_ * G), kBlock, 0, stream_>>>(Q_.view(), Rt_.view(), xv, nullptr,
                                                                             qx_.get(), n_, nullptr);
        });
        break;
    case FA_CP_DIAG_Q:
        diag_mul_kernel<<<blocks_for(n_), kBlock, 0, stream_>>>(qdiag_.get(), xv, qx_.get(), n_);
        break;
    case FA_CP_LOW_RANK_Q:
        spmv(R_, xv, rx_.get());
        dispatch_group(Rt_.G, [&](auto g) {
            constexpr int G = decltype(g)::value;
            qx_kernel<G, 1><<<blocks_for((long)n_ * G), kBlock, 0, stream_>>>(Q_.view(), Rt_.view(), xv, rx_.get(),
                                                                             qx_.get(), n_, nullptr);
        });
        break;
    case FA_CP_LOW_RANK_PLUS_SPARSE_Q:
        dispatch_group(Q_.G, [&](auto g) {
            constexpr int G = decltype(g)::value;
            qx_kernel<G, 0><<<blocks_for((long)n_ * G), kBlock, 0, stream_>>>(Q_.view(), Rt_.view(), xv, nullptr,
                                                                             qx_.get(), n_, nullptr);
        });
        spmv(R_, xv, rx_.get());
        dispatch_group(Rt_.G, [&](auto g) {
            constexpr int G = decltype(g)::value;
            qx_kernel<G, 2><<<blocks_for((long)n_ * G), kBlock, 0, stream_>>>(Q_.view(), Rt_.view(), xv, rx_.get(),
                                                                             qx_.get(), n_, nullptr);
        });
        break;
    default: return;
    }
    PDAL_LAUNCHED();
}

StepScalars Solver::scalars() const
{   // fa_cp_update, fa_cp_core_op.cu:531-545
    StepScalars sc{};
    double tau = step_ / pw_;
    if (nonconvexity_ < 0) {
        tau = std::fmax(tau, -1.01 * std::fmin(0.0, nonconvexity_));
        tau /= 2;
    }
    sc.tau = tau;
    sc.sigma = step_ * pw_;
    sc.al_sigma = use_al_ ? al_sigma_ : 0.0;
    sc.weight = (double)(inner_ + 1) / (inner_ + 2);
    sc.rho = P_.reflection_coefficient;
#ifdef PDAL_CP_AL_TYPES
    sc.tau_res = use_al_ ? lin_tau() : step_ / pw_;  // LIN_CP_AL utils.cu:855-858 (PLANTED_CP_AL :871-874)
#else
    sc.tau_res = step_ / pw_;
#endif
    return sc;
}

void Solver::bb_inner_step(const StepScalars &sc, bool test_first)
{   // one pass of the loop body in primal_BB_step_size_update, :452-500.  The
    // loop test of the previous direction update runs at the head of the
    // gradient kernel when test_first (else a separate kernel ran it); the BB
    // step runs at the head of the direction kernel.  The gradient partials
    // go to bbparts_g_, the direction partials to bbparts_.
    BBState *bb = dbb_.get();
    if (bb_al()) {
        dispatch_group(A_.G, [&](auto g) {
            constexpr int G = decltype(g)::value;
            al_effective_dual_kernel<G><<<blocks_for((long)m_ * G), kBlock, 0, stream_>>>(
                A_.view(), xT_.get(), y_.get(), l_.get(), u_.get(), prim_prod_.get(), al_u_.get(), sc, bb);
        });
        spmv(At_, al_u_.get(), dual_prod_.get(), bb);
    }
    const int nbd = stride_grid(n_);
    BBArgs a{c_.get(), lb_.get(), ub_.get(), x_.get(), dual_prod_.get(), qx_.get(), xT_.get(), grad_.get(),
             dir_.get(), bbparts_g_.get(), bbparts_.get(), nbd, bb, sc, n_, inner_tol_, inner_limit_};
    int nb = 0;
    auto grad = [&](auto gtag, auto mtag, const DevCsr &M) {
        constexpr int G = decltype(gtag)::value;
        constexpr int MODE = decltype(mtag)::value;
        nb = std::max(1, std::min(kMaxPartBlocks, blocks_for((long)n_ * G)));
        if (test_first)
            bb_grad_kernel<G, MODE, true><<<nb, kBlock, 0, stream_>>>(a, Q_.view(), M.view(), rx_.get(), qdiag_.get());
        else
            bb_grad_kernel<G, MODE, false><<<nb, kBlock, 0, stream_>>>(a, Q_.view(), M.view(), rx_.get(), qdiag_.get());
    };
    switch (qtype_) {
    case FA_CP_SPARSE_Q:
        dispatch_group(Q_.G, [&](auto g) { grad(g, std::integral_constant<int, 0>{}, Rt_); });
        break;
    case FA_CP_LOW_RANK_Q:
        spmv(R_, xT_.get(), rx_.get(), bb);
        dispatch_group(Rt_.G, [&](auto g) { grad(g, std::integral_constant<int, 1>{}, Rt_); });
        break;
    case FA_CP_LOW_RANK_PLUS_SPARSE_Q:  // the sparse part's product inside the gradient kernel
        spmv(R_, xT_.get(), rx_.get(), bb);
        dispatch_group(Rt_.G, [&](auto g) { grad(g, std::integral_constant<int, 5>{}, Rt_); });
        break;
    case FA_CP_DIAG_Q: grad(std::integral_constant<int, 1>{}, std::integral_constant<int, 4>{}, Rt_); break;
    default: grad(std::integral_constant<int, 1>{}, std::integral_constant<int, 3>{}, Rt_); break;
    }
    PDAL_LAUNCHED();
    BBArgs ad = a;
    ad.parts = bbparts_.get();
    ad.prev = bbparts_g_.get();
    ad.nprev = nb;
    bb_dir_kernel<<<nbd, kBlock, 0, stream_>>>(ad);
    PDAL_LAUNCHED();
}

void Solver::bb_loop_test()
{   // the loop test of the last direction update of a chunk, on its own
    bb_check_kernel<false><<<1, kBlock, 0, stream_>>>(bbparts_.get(), stride_grid(n_), dbb_.get(), inner_tol_,
                                                      inner_limit_);
    PDAL_LAUNCHED();
}

void Solver::launch_inner_steps(const StepScalars &sc, int steps)
{   // every chunk of steps tests the loop after its last direction update, so
    // the host's read of the done flag after a chunk is current
    if (!use_graph_) {
        for (int k = 0; k < steps; ++k) bb_inner_step(sc, k > 0);
        bb_loop_test();
        return;
    }
    if (!bb_graph_ || g_tau_ != sc.tau || g_al_ != sc.al_sigma || g_tol_ != inner_tol_ || g_limit_ != inner_limit_ ||
        g_use_al_ != bb_al()) {
        if (bb_graph_) PDAL_HIP(hipGraphExecDestroy(bb_graph_));
        bb_graph_ = nullptr;
        hipGraph_t graph;
        PDAL_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < kGraphSteps; ++k) bb_inner_step(sc, k > 0);
        bb_loop_test();
        PDAL_HIP(hipStreamEndCapture(stream_, &graph));
        PDAL_HIP(hipGraphInstantiate(&bb_graph_, graph, nullptr, nullptr, 0));
        PDAL_HIP(hipGraphDestroy(graph));
        g_tau_ = sc.tau;
        g_al_ = sc.al_sigma;
        g_tol_ = inner_tol_;
        g_limit_ = inner_limit_;
        g_use_al_ = bb_al();
    }
    for (int k = 0; k < steps; k += kGraphSteps) PDAL_HIP(hipGraphLaunch(bb_graph_, stream_));
}

void Solver::primal_bb(const StepScalars &sc, bool fuse_halpern)
{   // primal_BB_step_size_update, fa_cp_core_op.cu:445-503
    if (bb_al()) {  // compute_al_effective_dual_product at x (:275-303)
        dispatch_group(A_.G, [&](auto g) {
            constexpr int G = decltype(g)::value;
            al_effective_dual_kernel<G><<<blocks_for((long)m_ * G), kBlock, 0, stream_>>>(
                A_.view(), x_.get(), y_.get(), l_.get(), u_.get(), prim_prod_.get(), al_u_.get(), sc, nullptr);
        });
        spmv(At_, al_u_.get(), dual_prod_.get());
    }
    obj_product(x_.get());
    BBArgs a{c_.get(), lb_.get(), ub_.get(), x_.get(), dual_prod_.get(), qx_.get(), xT_.get(), grad_.get(),
             dir_.get(), bbparts_.get(), nullptr, 0, dbb_.get(), sc, n_, inner_tol_, inner_limit_};
    const int nbd = stride_grid(n_);
    bb_init_kernel<<<nbd, kBlock, 0, stream_>>>(a);
    bb_check_kernel<true><<<1, kBlock, 0, stream_>>>(bbparts_.get(), nbd, dbb_.get(), inner_tol_, inner_limit_);
    PDAL_LAUNCHED();
    // Inner steps in chunks, each ending with one read of the done flag.  The
    // first chunk is sized by the previous outer iteration's inner count; steps
    // launched after the loop has stopped return at once (gated kernels).
    int chunk = bb_chunk_;
    while (true) {
        launch_inner_steps(sc, chunk);
        PDAL_HIP(hipMemcpyAsync(hbb_.get(), dbb_.get(), sizeof(BBState), hipMemcpyDeviceToHost, stream_));
        PDAL_HIP(hipStreamSynchronize(stream_));
        if (hbb_[0].done) break;
        chunk = 8;
    }
    bb_chunk_ = std::max(1, std::min(64, hbb_[0].it - 1));
    if (fuse_halpern)
        bb_final_kernel<true><<<blocks_for(n_), kBlock, 0, stream_>>>(x_.get(), xT_.get(), xR_.get(), x0_.get(), n_,
                                                                     dbb_.get(), sc);
    else
        bb_final_kernel<false><<<blocks_for(n_), kBlock, 0, stream_>>>(x_.get(), xT_.get(), xR_.get(), x0_.get(),
                                                                      n_, dbb_.get(), sc);
    PDAL_LAUNCHED();
}

void Solver::update(bool fuse)
{   // fa_cp_update, fa_cp_core_op.cu:531-630 (+ the Halpern step when fused)
#ifdef PDAL_PLANTED_CP_AL
    pcal_update(fuse);
    return;
#endif
    const StepScalars sc = scalars();
    const bool mk = major_kernel();
    const bool simple = !use_al_ && (qtype_ == FA_CP_NON_Q || qtype_ == FA_CP_DIAG_Q);
    if (simple) {
        SimplePrimalArgs a{At_.view(), y_.get(), c_.get(), lb_.get(), ub_.get(), qdiag_.get(), x0_.get(),
                           x_.get(), xR_.get(), xT_.get(), dslack_.get(), dual_prod_.get(), sc};
        const bool diag = qtype_ == FA_CP_DIAG_Q;
        dispatch_group(At_.G, [&](auto g) {
            constexpr int G = decltype(g)::value;
            const int nb = blocks_for((long)n_ * G);
            const int key = (mk ? 4 : 0) | (diag ? 2 : 0) | (fuse ? 1 : 0);
            switch (key) {
            case 0: simple_primal_kernel<G, false, false, false><<<nb, kBlock, 0, stream_>>>(a); break;
            case 1: simple_primal_kernel<G, false, false, true><<<nb, kBlock, 0, stream_>>>(a); break;
            case 2: simple_primal_kernel<G, false, true, false><<<nb, kBlock, 0, stream_>>>(a); break;
            case 3: simple_primal_kernel<G, false, true, true><<<nb, kBlock, 0, stream_>>>(a); break;
            case 4: simple_primal_kernel<G, true, false, false><<<nb, kBlock, 0, stream_>>>(a); break;
            case 5: simple_primal_kernel<G, true, false, true><<<nb, kBlock, 0, stream_>>>(a); break;
            case 6: simple_primal_kernel<G, true, true, false><<<nb, kBlock, 0, stream_>>>(a); break;
            default: simple_primal_kernel<G, true, true, true><<<nb, kBlock, 0, stream_>>>(a); break;
            }
        });
        PDAL_LAUNCHED();
    } else {
        if (!use_al_) spmv(At_, y_.get(), dual_prod_.get());
        primal_bb(sc, fuse);
    }
    DualArgs d{A_.view(), xR_.get(), l_.get(), u_.get(), y0_.get(), y_.get(), yR_.get(), yT_.get(), sc};
    dispatch_group(A_.G, [&](auto g) {
        constexpr int G = decltype(g)::value;
        const int nb = blocks_for((long)m_ * G);
        const int key = (mk ? 4 : 0) | (use_al_ ? 2 : 0) | (fuse ? 1 : 0);
        switch (key) {
        case 0: dual_kernel<G, false, false, false><<<nb, kBlock, 0, stream_>>>(d); break;
        case 1: dual_kernel<G, false, false, true><<<nb, kBlock, 0, stream_>>>(d); break;
        case 2: dual_kernel<G, false, true, false><<<nb, kBlock, 0, stream_>>>(d); break;
        case 3: dual_kernel<G, false, true, true><<<nb, kBlock, 0, stream_>>>(d); break;
        case 4: dual_kernel<G, true, false, false><<<nb, kBlock, 0, stream_>>>(d); break;
        case 5: dual_kernel<G, true, false, true><<<nb, kBlock, 0, stream_>>>(d); break;
        case 6: dual_kernel<G, true, true, false><<<nb, kBlock, 0, stream_>>>(d); break;
        default: dual_kernel<G, true, true, true><<<nb, kBlock, 0, stream_>>>(d); break;
        }
    });
    PDAL_LAUNCHED();
}

void Solver::halpern()
{   // halpern_update, fa_cp_core_op.cu:632-640
    halpern_kernel<<<blocks_for((long)n_ + m_), kBlock, 0, stream_>>>(x_.get(), xR_.get(), x0_.get(), n_, y_.get(),
                                                                      yR_.get(), y0_.get(), m_, scalars());
    PDAL_LAUNCHED();
}

void Solver::fixed_point_error()
{   // compute_fixed_point_error, fa_cp_core_op.cu:1042-1089
#ifdef PDAL_PLANTED_CP_AL
    pcal_fixed_point_error();
    return;
#endif
    const int nbm = stride_grid(m_);
    diff_norm_kernel<<<nbm, kBlock, 0, stream_>>>(yR_.get(), y_.get(), dy_.get(), m_, parts_.get());
    reduce(nbm, 1, 1, S_DN2);
    int nb = 1;
    dispatch_group(At_.G, [&](auto g) {
        constexpr int G = decltype(g)::value;
        nb = row_grid(n_, G);
        fp_at_kernel<G><<<nb, kBlock, 0, stream_>>>(At_.view(), yR_.get(), y_.get(), xR_.get(), x_.get(), dx_.get(),
                                                     dual_prod_.get(), parts_.get());
    });
    PDAL_LAUNCHED();
    reduce(nb, 2, 2, S_FP);
    fetch(S_FP, 3);
    const double dn = std::sqrt(hs(S_DN2)), pn = std::sqrt(hs(S_FP + FP_PN2));
    const double movement = pn * pn * pw_ + dn * dn / pw_;
    const double cross = hs(S_FP + FP_CROSS);
#ifdef PDAL_PLANTED_FA_CP
    fp_ = std::sqrt(movement + 2 * step_ * cross);  // planted: no clamp at zero (:928-930)
#else
    fp_ = std::sqrt(std::fmax(movement + 2 * step_ * cross, 0.0));
#endif
    if (has_inner_) inner_tol_ = std::fmin(inner_tol_, std::fmax(0.0005 * pn / step_ * pw_, 1e-9));
}

void Solver::compute_residual()
{   // compute_residual, utils.cu:889-1026
    int nba = 1, nbt = 1;
    dispatch_group(A_.G, [&](auto g) {
        constexpr int G = decltype(g)::value;
        nba = row_grid(m_, G);
        res_a_kernel<G><<<nba, kBlock, 0, stream_>>>(A_.view(), xT_.get(), yT_.get(), l_.get(), u_.get(), E_.get(),
                                                     parts_.get());
    });
    PDAL_LAUNCHED();
    reduce(nba, 2, RA_N, S_RA);
    obj_product(xT_.get());
    const StepScalars sc = scalars();
    dispatch_group(At_.G, [&](auto g) {
        constexpr int G = decltype(g)::value;
        nbt = row_grid(n_, G);
        if (is_lp_)
            res_at_kernel<G, false><<<nbt, kBlock, 0, stream_>>>(At_.view(), yT_.get(), xT_.get(), qx_.get(),
                                                                 c_.get(), lb_.get(), ub_.get(), D_.get(),
                                                                 dslack_.get(), sc, parts_.get());
        else
            res_at_kernel<G, true><<<nbt, kBlock, 0, stream_>>>(At_.view(), yT_.get(), xT_.get(), qx_.get(), c_.get(),
                                                                lb_.get(), ub_.get(), D_.get(), dslack_.get(), sc,
                                                                parts_.get());
    });
    PDAL_LAUNCHED();
    reduce(nbt, 6, RT_N, S_RT);
    fetch(S_RA, RA_N + RT_N);
    const bool linf = P_.optimality_norm == NORM_TYPE_L_INF;
    const double *ra = hscal_.get() + S_RA, *rt = hscal_.get() + S_RT;
    abs_p_ = (linf ? ra[RA_PRESMAX] : std::sqrt(ra[RA_PRES2])) / b_scale_;
    abs_d_ = (linf ? rt[RT_DRESMAX] : std::sqrt(rt[RT_DRES2])) / c_scale_;
    const double half = qtype_ == FA_CP_NON_Q ? 0.0 : 0.5 * rt[RT_XQX];
    const double bc = b_scale_ * c_scale_;
    p_obj_ = (rt[RT_CX] + half) / bc + obj_const_;
    d_obj_ = (rt[RT_SX] + ra[RA_CONTRIB] - half) / bc + obj_const_;
    rel_p_ = abs_p_ / (1.0 + b_norm_);
    double den;
    if (is_lp_) {
        den = 1.0 + c_norm_;
    } else {  // recover_primal_obj_dual_product, :772-781
        const double qn = linf ? rt[RT_QXDMAX] : std::sqrt(rt[RT_QXD2]);
        const double an = linf ? rt[RT_DPDMAX] : std::sqrt(rt[RT_DPD2]);
        den = 1.0 + std::fmax(c_norm_, std::fmax(qn / c_scale_, an / c_scale_));
    }
    rel_d_ = abs_d_ / den;
    gap_ = std::fabs(p_obj_ - d_obj_);
    rel_gap_ = gap_ / (1.0 + std::fabs(p_obj_) + std::fabs(d_obj_));
}

void Solver::infeasibility_info()
{   // compute_infeasibility_information, utils.cu:1028-1132
    const int nb = stride_grid(std::max(m_, n_));
    ray_project_kernel<<<nb, kBlock, 0, stream_>>>(dx_.get(), lb_.get(), ub_.get(), n_, dy_.get(), l_.get(), u_.get(),
                                                   m_, parts_.get());
    PDAL_LAUNCHED();
    reduce(nb, 0, 2, S_PROJ);
    fetch(S_PROJ, 2);
    const double pinf = hs(S_PROJ), dinf = hs(S_PROJ + 1);
    if (pinf > 0.0) {
        scale_kernel<<<blocks_for(n_), kBlock, 0, stream_>>>(dx_.get(), 1.0 / pinf, n_);
        PDAL_LAUNCHED();
    }
    spmv(A_, dx_.get(), prim_prod_.get());
    spmv(At_, dy_.get(), dual_prod_.get());
    ray_terms_kernel<<<nb, kBlock, 0, stream_>>>(c_.get(), dx_.get(), dual_prod_.get(), lb_.get(), ub_.get(),
                                                 D_.get(), dslack_.get(), n_, dy_.get(), prim_prod_.get(), l_.get(),
                                                 u_.get(), E_.get(), pslack_.get(), m_, parts_.get());
    PDAL_LAUNCHED();
    reduce(nb, 3, RY_N, S_RAY);
    const bool quad = !is_lp_ && qtype_ != FA_CP_NON_Q;
    if (quad) {
        obj_product(dx_.get());
        const int nbn = stride_grid(n_);
        absmax_kernel<<<nbn, kBlock, 0, stream_>>>(qx_.get(), n_, parts_.get());
        PDAL_LAUNCHED();
        reduce(nbn, 0, 1, S_QXMAX);
    }
    fetch(S_RAY, RY_N + 1);
    const double bc = b_scale_ * c_scale_;
    const double *ry = hscal_.get() + S_RAY;
    p_ray_obj_ = ry[RY_CDX] / bc;
    d_ray_obj_ = (ry[RY_PSUM] + ry[RY_DSUM]) / bc;
    max_p_ray_ = ry[RY_PMAX];
    if (quad) max_p_ray_ = std::fmax(max_p_ray_, hs(S_QXMAX) / c_scale_);
    const double dsn = ry[RY_DMAX];
    max_d_ray_ = dsn;
    const double f = std::fmax(dinf, dsn);
    if (f > 0.0) {
        max_d_ray_ /= f;
        d_ray_obj_ /= f;
    } else {
        max_d_ray_ = 0.0;
        d_ray_obj_ = 0.0;
    }
}

void Solver::check_termination()
{   // check_termination_criteria, utils.cu:438-462
    const auto &tc = P_.termination_criteria;
    if (rel_d_ < tc.eps_feasible_relative && rel_p_ < tc.eps_feasible_relative && rel_gap_ < tc.eps_optimal_relative)
        reason_ = TERMINATION_REASON_OPTIMAL;
    else if (d_ray_obj_ > 0.0 && max_d_ray_ / d_ray_obj_ <= tc.eps_infeasible)
        reason_ = TERMINATION_REASON_PRIMAL_INFEASIBLE;
    else if (p_ray_obj_ < 0.0 && max_p_ray_ / (-p_ray_obj_) <= tc.eps_infeasible)
        reason_ = TERMINATION_REASON_DUAL_INFEASIBLE;
    else if (total_ >= tc.iteration_limit)
        reason_ = TERMINATION_REASON_ITERATION_LIMIT;
    else if (solve_time_ >= tc.time_sec_limit)
        reason_ = TERMINATION_REASON_TIME_LIMIT;
}

bool Solver::should_restart()
{   // should_do_adaptive_restart, utils.cu:464-492
    const auto &rp = P_.restart_params;
    const int K = P_.termination_evaluation_frequency;
    bool r = false;
    if (total_ == K) {
        r = true;
    } else if (total_ > K) {
        if (fp_ <= rp.sufficient_reduction_for_restart * fp0_) r = true;
        if (fp_ <= rp.necessary_reduction_for_restart * fp0_ && fp_ > last_trial_fp_) r = true;
        if (inner_ >= rp.artificial_restart_threshold * total_) r = true;
    }
    last_trial_fp_ = fp_;
    return r;
}

#if !defined(PDAL_CP_AL_TYPES) && !defined(PDAL_PLANTED_FA_CP)
void Solver::set_sigma_active(double sigma, double zero_thr)
{   // fa_cp_core_op.cu:667-676
    if (!std::isfinite(sigma) || sigma <= zero_thr) {
        al_sigma_ = 0.0;
        use_al_ = false;
    } else {
        al_sigma_ = sigma;
        use_al_ = al_available_;
    }
}

double Solver::clipped_update(double cur, double ratio) const
{   // fa_cp_core_op.cu:700-713
    const double gain = sane_positive(P_.sigma_update_gain, 0.5);
    double lo = sane_positive(P_.sigma_clip_low, 0.6), hi = sane_positive(P_.sigma_clip_high, 1.7);
    if (lo > hi) std::swap(lo, hi);
    ratio = sane_positive(ratio, 1.0);
    return cur * clampd_h(std::pow(ratio, gain), lo, hi);
}

void Solver::update_al_sigma(double pd, double dd, double normal_motion)
{   // update_al_sigma_at_restart, fa_cp_core_op.cu:722-906
    const auto &P = P_;
    const bool bal = P.sigma_update_rule == FA_CP_SIGMA_UPDATE_BALANCED_GUARDED;
    const bool guarded = P.sigma_update_rule == FA_CP_SIGMA_UPDATE_GUARDED || bal;
    if (!al_available_ || P.sigma_update_rule == FA_CP_SIGMA_UPDATE_FIXED) return;
    if (!guarded && !use_al_) return;
    const double ref = bal ? sane_positive(step_ * pw_, 1.0) : 1.0;
    double smin = sane_positive(P.sigma_min, 1e-8) * ref;
    double smax = sane_positive(P.sigma_max, 10.0) * ref;
    const double zthr = sane_positive(P.sigma_zero_threshold, sane_positive(P.sigma_min, 1e-8)) * ref;
    if (smin > smax) std::swap(smin, smax);
    if (zthr > smin) smin = zthr;
    const double cur = use_al_ ? clampd_h(sane_positive(al_sigma_, smin), smin, smax) : 0.0;
    set_sigma_active(cur, zthr);
    const double rp = sane_positive(rel_p_, 1e-300), rd = sane_positive(rel_d_, 1e-300),
                 rg = sane_positive(rel_gap_, 1e-300);
    const double score =
        std::fmax(score_component(rel_p_), std::fmax(score_component(rel_d_), score_component(rel_gap_)));
    if (!std::isfinite(score)) return;
    if (!std::isfinite(best_sigma_score_) || score < best_sigma_score_) {
        best_sigma_score_ = score;
        best_al_sigma_ = cur;
    }
    if (P.sigma_rollback_tol > 0.0 && (!guarded || cur > zthr) && std::isfinite(best_sigma_score_) &&
        best_sigma_score_ > 0.0 && score > P.sigma_rollback_tol * best_sigma_score_ && std::isfinite(best_al_sigma_) &&
        (guarded ? best_al_sigma_ >= 0.0 : best_al_sigma_ > 0.0)) {
        const double best = guarded && best_al_sigma_ <= zthr ? 0.0 : clampd_h(best_al_sigma_, smin, smax);
        set_sigma_active(best, zthr);
        last_sigma_score_ = score;
        last_sigma_rp_ = rp;
        last_sigma_rd_ = rd;
        last_sigma_gap_ = rg;
        sigma_update_count_++;
        return;
    }
    double prop = cur;
    switch (P.sigma_update_rule) {
    case FA_CP_SIGMA_UPDATE_RESIDUAL_BALANCE: prop = clipped_update(cur, rp / rd); break;
    case FA_CP_SIGMA_UPDATE_KKT_BALANCE: prop = clipped_update(cur, rp / std::fmax(rd, rg)); break;
    case FA_CP_SIGMA_UPDATE_DISPLACEMENT:
        if (std::isfinite(pd) && std::isfinite(dd) && pd > 1e-16 && dd > 1e-16) {
            const double target = clampd_h(dd / pd, smin, smax);
            prop = clipped_update(cur, target / cur);
        }
        break;
    case FA_CP_SIGMA_UPDATE_PROGRESS: {
        const double lo = std::fmax(last_sigma_rd_, last_sigma_gap_);
        const double other = std::fmax(rd, rg);
        if (std::isfinite(last_sigma_rp_) && std::isfinite(lo) && last_sigma_rp_ > 0.0 && lo > 0.0)
            prop = clipped_update(cur, (lo / other) / (last_sigma_rp_ / rp));
        else
            prop = clipped_update(cur, rp / std::fmax(rd, rg));
        break;
    }
    case FA_CP_SIGMA_UPDATE_GREEDY_SCORE:
        if (sigma_update_count_ == 0) sigma_dir_ = (rp >= std::fmax(rd, rg)) ? 1 : -1;
        else if (std::isfinite(score) && std::isfinite(last_sigma_score_) && score > last_sigma_score_)
            sigma_dir_ *= -1;
        prop = clipped_update(cur, sigma_dir_ > 0 ? P.sigma_clip_high : P.sigma_clip_low);
        break;
    case FA_CP_SIGMA_UPDATE_GUARDED:
    case FA_CP_SIGMA_UPDATE_BALANCED_GUARDED: {
        const double nthr = sane_positive(P.sigma_normal_motion_threshold, 1e-10);
        const bool normal_ok = !std::isfinite(normal_motion) || normal_motion > nthr;
        const bool zero_ok = !std::isfinite(normal_motion) || normal_motion > std::fmax(nthr, 2e-1);
        const bool bottleneck =
            rp >= sane_positive(P.sigma_activation_threshold, 2.0) * sane_positive(std::fmax(rd, rg), 1e-300);
        const bool zero_stalled = cur <= zthr && std::isfinite(last_sigma_score_) && score > 0.9 * last_sigma_score_;
        const bool primal_signal = rp >= 0.5 * sane_positive(rd, 1e-300);
        const double act = sane_positive(P.sigma_activation_value, 0.1) * ref;
        const bool scale_ok = !bal || act >= 1e-6;
        const bool activate =
            cur <= zthr ? zero_ok && zero_stalled && primal_signal && scale_ok : normal_ok && bottleneck;
        if (cur <= zthr) prop = activate ? act : 0.0;
        else if (!normal_ok) prop = clipped_update(cur, P.sigma_clip_low);
        else if (activate) prop = clipped_update(cur, rp / std::fmax(rd, rg));
        else prop = clipped_update(cur, P.sigma_clip_low);
        break;
    }
    default: break;
    }
    if (prop > zthr) prop = clampd_h(prop, smin, smax);
    else prop = 0.0;
    set_sigma_active(prop, zthr);
    last_sigma_score_ = score;
    last_sigma_rp_ = rp;
    last_sigma_rd_ = rd;
    last_sigma_gap_ = rg;
    sigma_update_count_++;
}

#endif

#ifdef PDAL_PLANTED_FA_CP
void Solver::planted_sigma_on_restart()
{   // maybe_update_al_sigma_on_restart + apply_new_al_sigma, QP/PLANTED_FA_CP fa_cp_core_op.cu:610-747
    const auto &P = P_;
    if (!P.adaptive_al_sigma || !use_al_) return;
    if (last_sigma_update_iter_ == total_) return;
    const double rp = std::fabs(rel_p_), rd = std::fabs(rel_d_), rg = std::fabs(rel_gap_);
    if (!std::isfinite(rp) || !std::isfinite(rd) || !std::isfinite(rg)) return;
    const double dg = std::fmax(rd, rg);
    const double pref = std::fmax(rp, 1e-12), dgref = std::fmax(dg, 1e-12);
    const double ratio = rp / dgref;
    const bool up = rp > P.al_sigma_primal_trigger * dgref && P.adaptive_al_sigma_mode != ADAPTIVE_AL_SIGMA_MODE_DOWN_ONLY;
    const bool down = dg > P.al_sigma_gap_trigger * pref && P.adaptive_al_sigma_mode != ADAPTIVE_AL_SIGMA_MODE_UP_ONLY;
    double ns = al_sigma_;
    switch (P.adaptive_al_sigma_rule) {
    case ADAPTIVE_AL_SIGMA_RULE_TRIGGER:
        if (up) ns = std::fmin(P.al_sigma_max, al_sigma_ * P.al_sigma_increase_factor);
        else if (down) ns = std::fmax(P.al_sigma_min, al_sigma_ * P.al_sigma_decrease_factor);
        break;
    case ADAPTIVE_AL_SIGMA_RULE_SQRT:
        if (up) ns = std::fmin(P.al_sigma_max, al_sigma_ * std::fmin(P.al_sigma_increase_factor, std::sqrt(ratio)));
        else if (down) ns = std::fmax(P.al_sigma_min, al_sigma_ * std::fmax(P.al_sigma_decrease_factor, std::sqrt(ratio)));
        break;
    case ADAPTIVE_AL_SIGMA_RULE_BANGBANG:
        if (up) ns = P.al_sigma_max;
        else if (down) ns = P.al_sigma_min;
        break;
    default: break;
    }
    if (std::fabs(ns - al_sigma_) <= 1e-12 * std::fmax(1.0, std::fabs(al_sigma_))) return;
    al_sigma_ = ns;
    last_sigma_update_iter_ = total_;
    if (has_inner_) inner_tol_ = inner_tol0_;  // reset_inner_solver_tolerance, :610-614
}
#endif

#ifndef PDAL_LIN_CP_AL
void Solver::pid_primal_weight(double pd, double dd)
{   // the primal-weight PID of perform_restart, fa_cp_core_op.cu:908-1020 (PLANTED_CP_AL :758-786)
    const double ratio = rel_d_ / rel_p_;
    const auto &rp = P_.restart_params;
    if (pd > 1e-16 && dd > 1e-16 && pd < 1e12 && dd < 1e12 && ratio > 1e-8 && ratio < 1e8) {
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
    const double g = std::fabs(std::log10(rel_d_ / rel_p_));
    if (g < best_gap_) {
        best_gap_ = g;
        best_pw_ = pw_;
    }
}

#if !defined(PDAL_CP_AL_TYPES) && !defined(PDAL_PLANTED_FA_CP)
const char *restart_env(const char *name, const char *fallback)  // get_env_value, fa_cp_core_op.cu:35-42
{
    const char *v = getenv(name);
    if ((!v || !v[0]) && fallback) v = getenv(fallback);
    return v;
}

int restart_env_int(const char *name, const char *fallback, int def)  // :44-51
{
    const char *v = restart_env(name, fallback);
    return (!v || !v[0]) ? def : atoi(v);
}

double restart_env_double(const char *name, double def)  // :53-66
{
    const char *v = restart_env(name, name);
    if (!v || !v[0]) return def;
    char *end = nullptr;
    const double x = strtod(v, &end);
    return (end == v || !std::isfinite(x)) ? def : x;
}

int restart_select_mode() { return restart_env_int("FA_CP_RESTART_SELECT", "FA_CP_RESTART_SELECT_LEGACY", 0); }

void Solver::choose_restart_candidate()
{   // :1091-1169: the current iterate as restart point when its score (the fixed-point
    // metric of x_T against it, without the AL term) beats the T output's by the margin
    const int mode = restart_select_mode();
    restart_target_ = 0;
    const double cur = std::isfinite(fp_) ? fp_ : INFINITY;
    if (mode == 0 || total_ == 0 || !std::isfinite(cur)) return;
    if (total_ < restart_env_int("FA_CP_RESTART_SELECT_MIN_ITERS", "FA_CP_RESTART_SELECT_MIN_ITERS", 0) ||
        solve_time_ < restart_env_double("FA_CP_RESTART_SELECT_MIN_TIME", 0.0))
        return;
    const int nbn = stride_grid(n_), nbm = stride_grid(m_);
    diff_norm_kernel<<<nbn, kBlock, 0, stream_>>>(x_.get(), xT_.get(), dx_.get(), n_, parts_.get());
    reduce(nbn, 1, 1, S_SEL);
    diff_norm_kernel<<<nbm, kBlock, 0, stream_>>>(y_.get(), yT_.get(), dy_.get(), m_, parts_.get());
    reduce(nbm, 1, 1, S_SEL + 1);
    spmv(At_, dy_.get(), dual_prod_.get());
    dot_sumsq_kernel<<<nbn, kBlock, 0, stream_>>>(dual_prod_.get(), dx_.get(), n_, parts_.get());
    reduce(nbn, 2, 1, S_SEL + 2);
    PDAL_LAUNCHED();
    fetch(S_SEL, 3);
    const double pn2 = hs(S_SEL), dn2 = hs(S_SEL + 1), cross = hs(S_SEL + 2);
    const double halpern = std::sqrt(std::fmax(pn2 * pw_ + dn2 / pw_ + 2.0 * step_ * cross, 0.0));
    if (!(cur > 0.0)) return;
    const double margin = restart_env_double("FA_CP_RESTART_SELECT_MARGIN", 1.0);
    if (mode == 2 || (std::isfinite(halpern) && halpern < margin * cur)) {
        restart_target_ = 1;
        fp_ = halpern;
    }
}
#endif

void Solver::perform_restart()
{   // perform_restart, fa_cp_core_op.cu:908-1020: to the T output, or to the current iterate
    // when FA_CP's candidate selection chose it
#if !defined(PDAL_CP_AL_TYPES) && !defined(PDAL_PLANTED_FA_CP)
    const bool to_current = restart_select_mode() != 0 && restart_target_ == 1;
#else
    const bool to_current = false;
#endif
    const double *rx = to_current ? x_.get() : xT_.get(), *ry = to_current ? y_.get() : yT_.get();
    const int nbn = stride_grid(n_), nbm = stride_grid(m_);
    diff_norm_kernel<<<nbn, kBlock, 0, stream_>>>(rx, x0_.get(), dx_.get(), n_, parts_.get());
    reduce(nbn, 1, 1, S_PD2);
    diff_norm_kernel<<<nbm, kBlock, 0, stream_>>>(ry, y0_.get(), dy_.get(), m_, parts_.get());
    reduce(nbm, 1, 1, S_DD2);
    PDAL_LAUNCHED();
    fetch(S_PD2, 2);
    const double pd = std::sqrt(hs(S_PD2)), dd = std::sqrt(hs(S_DD2));
#ifdef PDAL_PLANTED_CP_AL
    if (use_al_) {  // ||A (xT - x0)|| for the sigma controller (PLANTED_CP_AL :670-686)
        int nb = 1;
        dispatch_group(A_.G, [&](auto g) {
            constexpr int G = decltype(g)::value;
            nb = row_grid(m_, G);
            spmv_norm_kernel<G><<<nb, kBlock, 0, stream_>>>(A_.view(), dx_.get(), prim_prod_.get(), parts_.get());
        });
        PDAL_LAUNCHED();
        reduce(nb, 1, 1, S_NORMAL);
        fetch(S_NORMAL, 1);
        pcal_sigma_controller(pd, dd, std::sqrt(hs(S_NORMAL)));
    } else {
        pid_primal_weight(pd, dd);
    }
#else
    double normal = INFINITY;
    if (pd > 1e-16 && std::isfinite(pd)) {
        int nb = 1;
        dispatch_group(A_.G, [&](auto g) {
            constexpr int G = decltype(g)::value;
            nb = row_grid(m_, G);
            spmv_norm_kernel<G><<<nb, kBlock, 0, stream_>>>(A_.view(), dx_.get(), prim_prod_.get(), parts_.get());
        });
        PDAL_LAUNCHED();
        reduce(nb, 1, 1, S_NORMAL);
        fetch(S_NORMAL, 1);
        normal = std::sqrt(hs(S_NORMAL)) / pd;
    }
    pid_primal_weight(pd, dd);
#ifdef PDAL_PLANTED_FA_CP
    (void)normal;
    planted_sigma_on_restart();
#else
    update_al_sigma(pd, dd, normal);
#endif
#endif  // PDAL_PLANTED_CP_AL
    if (to_current) {  // :985-1009: anchor and T output = the current iterate
        restart_copy_kernel<<<blocks_for((long)n_ + m_), kBlock, 0, stream_>>>(x_.get(), x0_.get(), xT_.get(), n_,
                                                                               y_.get(), y0_.get(), yT_.get(), m_);
    } else {
        restart_copy_kernel<<<blocks_for((long)n_ + m_), kBlock, 0, stream_>>>(xT_.get(), x0_.get(), x_.get(), n_,
                                                                               yT_.get(), y0_.get(), y_.get(), m_);
    }
    PDAL_LAUNCHED();
    inner_ = 0;
    last_trial_fp_ = INFINITY;
}

#endif  // !PDAL_LIN_CP_AL

result_t *Solver::make_result()
{   // create_result_from_state, fa_cp_core_op.cu:1171-1239
    spmv(At_, yT_.get(), dual_prod_.get());
    // reuse iterate buffers that are no longer needed as outputs
    double *xo = grad_.get(), *rc = dir_.get(), *yo = al_u_.get();
    result_kernel<<<blocks_for((long)n_ + m_), kBlock, 0, stream_>>>(xT_.get(), D_.get(), b_scale_, xo, c_.get(),
                                                                     dual_prod_.get(), c_scale_, rc, n_, yT_.get(),
                                                                     E_.get(), yo, m_);
    PDAL_LAUNCHED();
    PDAL_HIP(hipMemcpyAsync(hbb_.get(), dbb_.get(), sizeof(BBState), hipMemcpyDeviceToHost, stream_));
    result_t *r = (result_t *)calloc(1, sizeof(result_t));
    r->primal_solution = (double *)malloc((size_t)std::max(n_, 1) * sizeof(double));
    r->dual_solution = (double *)malloc((size_t)std::max(m_, 1) * sizeof(double));
    r->reduced_cost = (double *)malloc((size_t)std::max(n_, 1) * sizeof(double));
    PDAL_HIP(hipMemcpyAsync(r->primal_solution, xo, (size_t)n_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    PDAL_HIP(hipMemcpyAsync(r->reduced_cost, rc, (size_t)n_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    PDAL_HIP(hipMemcpyAsync(r->dual_solution, yo, (size_t)m_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    PDAL_HIP(hipStreamSynchronize(stream_));
    r->num_variables = W_->num_variables;
    r->num_constraints = W_->num_constraints;
    r->num_nonzeros = W_->constraint_matrix_num_nonzeros;
    r->total_count = total_;
    r->total_inner_count = (int)hbb_[0].total;
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
    r->max_primal_ray_infeasibility = max_p_ray_;
    r->max_dual_ray_infeasibility = max_d_ray_;
    r->primal_ray_linear_objective = p_ray_obj_;
    r->dual_ray_objective = d_ray_obj_;
#if !defined(PDAL_CP_AL_TYPES) && !defined(PDAL_PLANTED_FA_CP)
    r->final_al_sigma = al_sigma_;
    r->sigma_update_count = sigma_update_count_;
    r->sigma_update_rule = al_available_ ? P_.sigma_update_rule : FA_CP_SIGMA_UPDATE_FIXED;
#endif
    r->termination_reason = reason_;
    r->feasibility_polishing_time = polish_time_;
    r->feasibility_iteration = polish_iters_;
    return r;
}

void Solver::print_initial_info() const
{   // print_initial_info, utils.cu:560-659 (the settings that differ from the defaults are omitted)
    if (P_.verbose < 2) return;
    printf("---------------------------------------------------------------------------------------\n");
    printf("                          %s (MI355X / gfx950 HIP build)\n", kPlantedCpAl ? "RHR-CP-AL" : kLinearized ? "RHR-Lin-CP-AL" : kEqualityAL ? "RHR-FA-CP" : "FA_CP");
    printf("---------------------------------------------------------------------------------------\n");
    printf("problem: %d rows, %d columns, %d nonzeros\n", W_->num_constraints, W_->num_variables,
           W_->constraint_matrix_num_nonzeros);
    printf("settings:\n");
    printf("  iter_limit         : %d\n", P_.termination_criteria.iteration_limit);
    printf("  time_limit         : %.2f sec\n", P_.termination_criteria.time_sec_limit);
    printf("  eps_opt            : %.1e\n", P_.termination_criteria.eps_optimal_relative);
    printf("  eps_feas           : %.1e\n", P_.termination_criteria.eps_feasible_relative);
    printf("  eps_infeas_detect  : %.1e\n", P_.termination_criteria.eps_infeasible);
}

void Solver::display_stats() const
{   // display_iteration_stats, utils.cu:680-688
    if (P_.verbose < 2) return;
    if (total_ % print_frequency(total_) == 0)
        printf("%6d %.1e | %8.1e  %8.1e | %.1e %.1e %.1e | %.1e %.1e %.1e \n", total_, solve_time_, p_obj_, d_obj_,
               abs_p_, abs_d_, gap_, rel_p_, rel_d_, rel_gap_);
}

// ---------------------------------------------------------------------------
// Linearized CP-AL (PDAL_LIN_CP_AL builds; QP/LIN_CP_AL/src/cp_al_core_op.cu)

void Solver::lin_init_steps()
{   // initialize_step_size_and_primal_weight, :609-691
    const double max_sv = A_.nnz > 0 ? max_singular_value() : 0.0;
    const double a_sq = max_sv * max_sv;
    const double half_lf = 0.5 * std::fmax(q_norm_, 0.0);
    pw_ = P_.bound_objective_rescaling ? 1.0 : (c_norm_ + 1.0) / (b_norm_ + 1.0);
    best_pw_ = pw_;
    if (use_al_) {
        if (a_sq <= 1e-16 && half_lf <= 1e-16) {
            lambda_est_ = 1.0;
            sigma_base_ = 1.0;
        } else {  // t = sqrt(lambda) solves t^2 - (L_f / 2) t - 2 ||A||^2 = 0 (with 8.0004 for 8)
            const double t = std::fmax(0.5 * (half_lf + std::sqrt(half_lf * half_lf + 8.0004 * a_sq)), 1e-16);
            lambda_est_ = t * t;
            sigma_base_ = 1.0 / t;
        }
        sigma_min_ = sigma_base_;
        sigma_max_ = 100.0 * sigma_base_;
#ifdef PDAL_CP_AL_TYPES
        const double init = P_.sigma_init;
#else
        const double init = 0.0;
#endif
        sigma_ = init > 0.0 ? init : 6.0 * sigma_base_;
        sigma_ = std::fmin(std::fmax(sigma_, sigma_min_), sigma_max_);
        step_ = lin_tau();
    } else {
        const double wgt = std::fmax(pw_, 1e-16);
        if (a_sq <= 1e-16 && half_lf <= 1e-16) {
            step_ = 1.0;
        } else if (a_sq <= 1e-16) {
            step_ = 0.99 * wgt / std::fmax(half_lf, 1e-16);
        } else {
            const double lc = half_lf / wgt;
            step_ = 0.99 * (-lc + std::sqrt(lc * lc + 4.0 * a_sq)) / (2.0 * a_sq);
        }
    }
}

void Solver::lin_update()
{   // cp_al_update, :496-575: the T-outputs and reflections every iteration
    double tau, dstep, al_sigma = 0.0;
    if (use_al_) {
        tau = lin_tau();
        dstep = std::fmax(sigma_, 1e-16);
        al_sigma = sigma_;
        step_ = tau;
    } else {
        tau = step_ / pw_;
        dstep = step_ * pw_;
    }
    if (nonconvexity_ < 0) {
        tau = std::fmax(tau, -1.01 * std::fmin(0.0, nonconvexity_));
        tau /= 2;
    }
    if (use_al_) {  // compute_al_effective_dual_product, :275-302
        spmv(A_, x_.get(), prim_prod_.get());
        lin_al_dual_kernel<<<blocks_for(m_), kBlock, 0, stream_>>>(prim_prod_.get(), y_.get(), l_.get(), al_sigma,
                                                                   al_u_.get(), m_);
        PDAL_LAUNCHED();
        spmv(At_, al_u_.get(), dual_prod_.get());
    } else {
        spmv(At_, y_.get(), dual_prod_.get());
    }
    if (qtype_ == FA_CP_NON_Q) {  // lp_primal_update, :304-312
        lin_lp_primal_kernel<<<blocks_for(n_), kBlock, 0, stream_>>>(x_.get(), dual_prod_.get(), c_.get(), lb_.get(),
                                                                     ub_.get(), tau, xT_.get(), xR_.get(),
                                                                     dslack_.get(), n_);
    } else {  // primal_gradient_update, :484-494
        obj_product(x_.get());
        lin_grad_primal_kernel<<<blocks_for(n_), kBlock, 0, stream_>>>(x_.get(), qx_.get(), c_.get(),
                                                                       dual_prod_.get(), lb_.get(), ub_.get(), tau,
                                                                       xT_.get(), xR_.get(), n_);
    }
    PDAL_LAUNCHED();
    spmv(A_, xR_.get(), prim_prod_.get());
    lin_dual_kernel<true><<<blocks_for(m_), kBlock, 0, stream_>>>(y_.get(), prim_prod_.get(), l_.get(), dstep, yT_.get(),
                                                            yR_.get(), m_);
    PDAL_LAUNCHED();
}

void Solver::lin_fixed_point_error()
{   // compute_fixed_point_error, :693-770 (delta = T-output - current)
    const int nbn = stride_grid(n_), nbm = stride_grid(m_);
    diff_norm_kernel<<<nbm, kBlock, 0, stream_>>>(yT_.get(), y_.get(), dy_.get(), m_, parts_.get());
    reduce(nbm, 1, 1, S_DN2);
    diff_norm_kernel<<<nbn, kBlock, 0, stream_>>>(xT_.get(), x_.get(), dx_.get(), n_, parts_.get());
    reduce(nbn, 1, 1, S_PD2);
    PDAL_LAUNCHED();
    if (use_al_) {  // <A dx, dy> and ||A dx||^2
        spmv(A_, dx_.get(), prim_prod_.get());
        dot_sumsq_kernel<<<nbm, kBlock, 0, stream_>>>(dy_.get(), prim_prod_.get(), m_, parts_.get());
        reduce(nbm, 2, 2, S_FP);
    } else {  // <At dy, dx>
        spmv(At_, dy_.get(), dual_prod_.get());
        dot_sumsq_kernel<<<nbn, kBlock, 0, stream_>>>(dx_.get(), dual_prod_.get(), n_, parts_.get());
        reduce(nbn, 2, 2, S_FP);
    }
    PDAL_LAUNCHED();
    fetch(S_FP, S_PD2 - S_FP + 1);
    const double pn2 = hs(S_PD2), dn2 = hs(S_DN2), cross = hs(S_FP);
    double fp2;
    if (use_al_) {
        const double an2 = hs(S_FP + 1);
        fp2 = sigma_ * (lambda_est_ * pn2 - an2) + 2.0 * cross + dn2 / sigma_;
    } else {
        const double tau = step_ / pw_, rho = step_ * pw_;
        fp2 = pn2 / std::fmax(tau, 1e-16) + dn2 / std::fmax(rho, 1e-16) + 2.0 * cross;
    }
    fp_ = std::sqrt(std::fmax(0.0, fp2));
}

bool Solver::lin_should_restart() const
{   // should_do_adaptive_restart, LIN_CP_AL utils.cu:451-463
    if (!std::isfinite(fp0_) || !std::isfinite(fp_)) return false;
    if (total_ == inner_) return inner_ >= P_.termination_evaluation_frequency;
    return fp_ <= P_.restart_params.sufficient_reduction_for_restart * fp0_;
}

void Solver::lin_restart()
{   // perform_restart, :595-607: the anchors move to the current iterates
    copy(x0_.get(), x_.get(), n_);
    copy(y0_.get(), y_.get(), m_);
    inner_ = 0;
    last_trial_fp_ = INFINITY;
}

void Solver::lin_halpern()
{   // halpern_update, :577-585
    const double w = (double)(inner_ + 1) / (inner_ + 2);
    lin_halpern_kernel<<<blocks_for((long)n_ + m_), kBlock, 0, stream_>>>(x_.get(), xT_.get(), x0_.get(), n_,
                                                                          y_.get(), yT_.get(), y0_.get(), m_, w,
                                                                          P_.reflection_coefficient);
    PDAL_LAUNCHED();
}

result_t *Solver::run_linearized()
{   // optimize, QP/LIN_CP_AL/src/solver.cu:34-114
    print_initial_info();
    const auto t_setup = std::chrono::steady_clock::now();
    upload();
    rescale();
    rescale_time_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_setup).count();
    init_state();
    if (nonconvexity_ < 0) inner_limit_ = 1;
    lin_init_steps();
    PDAL_HIP(hipStreamSynchronize(stream_));
    const auto t0 = std::chrono::steady_clock::now();
    const int K = P_.termination_evaluation_frequency;
    while (total_ < P_.termination_criteria.iteration_limit) {
        const bool eval = total_ == 0 || total_ % K == 0;
        major_ = eval;
        if (eval || total_ % print_frequency(total_) == 0) {
            compute_residual();
            if (eval && total_ < 3 * K) infeasibility_info();
            solve_time_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            check_termination();
            display_stats();
            if (reason_ != TERMINATION_REASON_UNSPECIFIED) break;
        }
        lin_update();
        lin_fixed_point_error();
        if (!std::isfinite(fp0_)) fp0_ = fp_;
        if (eval && lin_should_restart()) {
            lin_restart();
            fp0_ = fp_;
        }
        lin_halpern();
        inner_++;
        total_++;
    }
    if (reason_ == TERMINATION_REASON_UNSPECIFIED) {
        reason_ = TERMINATION_REASON_ITERATION_LIMIT;
        compute_residual();
        display_stats();
    }
    if (P_.feasibility_polishing && reason_ != TERMINATION_REASON_DUAL_INFEASIBLE &&
        reason_ != TERMINATION_REASON_PRIMAL_INFEASIBLE)  // QP/LIN_CP_AL/src/solver.cu:102-106
        feasibility_polish();
    return make_result();
}

#ifdef PDAL_PLANTED_CP_AL
// ---------------------------------------------------------------------------
// Planted-KKT CP-AL (PDAL_PLANTED_CP_AL builds; QP/PLANTED_CP_AL/src/cp_al_core_op.cu)

void Solver::pcal_init_steps()
{   // initialize_step_size_and_primal_weight, :805-858
    if (use_al_) {
        if (A_.nnz == 0) {
            lambda_est_ = 1.0;
        } else {
            const double sv = max_singular_value();
            lambda_est_ = 2.001 * sv * sv;
        }
        sigma_base_ = 1.0 / std::sqrt(std::fmax(lambda_est_, 1e-16));
        sigma_min_ = 1e-4 * sigma_base_;
        sigma_max_ = 100.0 * sigma_base_;
        sigma_ = P_.sigma_init > 0.0 ? P_.sigma_init : 0.5 * sigma_base_;
        sigma_ = std::fmin(std::fmax(sigma_, sigma_min_), sigma_max_);
        best_sigma_ = sigma_;
        best_sigma_gap_ = INFINITY;
        best_term_metric_ = INFINITY;
        step_ = pcal_tau();
    } else {
        step_ = A_.nnz == 0 ? 1.0 : 0.998 / max_singular_value();
    }
    pw_ = P_.bound_objective_rescaling ? 1.0 : (c_norm_ + 1.0) / (b_norm_ + 1.0);
    best_pw_ = pw_;
}

void Solver::pcal_update(bool fuse)
{   // cp_al_update, :531-635: equality rows; the T-outputs (and the LP dual slack)
    // only in major-kernel iterations
    double tau, dstep;
    if (use_al_) {
        tau = pcal_tau();
        dstep = std::fmax(sigma_, 1e-16);
        step_ = tau;
    } else {
        tau = step_ / pw_;
        dstep = step_ * pw_;
        if (nonconvexity_ < 0) {
            tau = std::fmax(tau, -1.01 * std::fmin(0.0, nonconvexity_));
            tau /= 2;
        }
    }
    const bool mk = major_kernel();
    if (use_al_) {  // compute_al_effective_dual_product, :285-312
        spmv(A_, x_.get(), prim_prod_.get());
        lin_al_dual_kernel<<<blocks_for(m_), kBlock, 0, stream_>>>(prim_prod_.get(), y_.get(), l_.get(), sigma_,
                                                                   al_u_.get(), m_);
        PDAL_LAUNCHED();
        spmv(At_, al_u_.get(), dual_prod_.get());
    } else {
        spmv(At_, y_.get(), dual_prod_.get());
    }
    if (qtype_ == FA_CP_NON_Q || qtype_ == FA_CP_DIAG_Q) {  // lp_primal_update / diag_q_primal_update, :314-356
        const bool diag = qtype_ == FA_CP_DIAG_Q;
        auto launch = [&](auto dtag, auto mtag) {
            pcal_primal_kernel<decltype(dtag)::value, decltype(mtag)::value><<<blocks_for(n_), kBlock, 0, stream_>>>(
                x_.get(), dual_prod_.get(), c_.get(), lb_.get(), ub_.get(), qdiag_.get(), tau, xT_.get(), xR_.get(),
                dslack_.get(), n_);
        };
        using T = std::true_type;
        using F = std::false_type;
        if (diag) {
            if (mk) launch(T{}, T{});
            else launch(T{}, F{});
        } else {
            if (mk) launch(F{}, T{});
            else launch(F{}, F{});
        }
        PDAL_LAUNCHED();
    } else {  // primal_BB_step_size_update, :454-503: the BB with At.u fixed (bb_al() is false)
        StepScalars sc = scalars();
        sc.tau = tau;
        sc.al_sigma = 0.0;
        primal_bb(sc, false);
    }
    spmv(A_, xR_.get(), prim_prod_.get());
    if (mk)
        lin_dual_kernel<true><<<blocks_for(m_), kBlock, 0, stream_>>>(y_.get(), prim_prod_.get(), l_.get(), dstep,
                                                                      yT_.get(), yR_.get(), m_);
    else
        lin_dual_kernel<false><<<blocks_for(m_), kBlock, 0, stream_>>>(y_.get(), prim_prod_.get(), l_.get(), dstep,
                                                                       yT_.get(), yR_.get(), m_);
    PDAL_LAUNCHED();
    if (fuse) halpern();
}

void Solver::pcal_fixed_point_error()
{   // compute_fixed_point_error, :860-935 (delta = reflected - current)
    const int nbn = stride_grid(n_), nbm = stride_grid(m_);
    diff_norm_kernel<<<nbm, kBlock, 0, stream_>>>(yR_.get(), y_.get(), dy_.get(), m_, parts_.get());
    reduce(nbm, 1, 1, S_DN2);
    diff_norm_kernel<<<nbn, kBlock, 0, stream_>>>(xR_.get(), x_.get(), dx_.get(), n_, parts_.get());
    reduce(nbn, 1, 1, S_PD2);
    PDAL_LAUNCHED();
    if (use_al_) {  // <A dx, dy> and ||A dx||^2
        spmv(A_, dx_.get(), prim_prod_.get());
        dot_sumsq_kernel<<<nbm, kBlock, 0, stream_>>>(dy_.get(), prim_prod_.get(), m_, parts_.get());
        reduce(nbm, 2, 2, S_FP);
    } else {  // <At dy, dx>
        spmv(At_, dy_.get(), dual_prod_.get());
        dot_sumsq_kernel<<<nbn, kBlock, 0, stream_>>>(dx_.get(), dual_prod_.get(), n_, parts_.get());
        reduce(nbn, 2, 2, S_FP);
    }
    PDAL_LAUNCHED();
    fetch(S_FP, S_PD2 - S_FP + 1);
    const double pn = std::sqrt(hs(S_PD2)), dn = std::sqrt(hs(S_DN2)), cross = hs(S_FP);
    if (use_al_) {
        const double an = std::sqrt(hs(S_FP + 1));
        const double fp2 = sigma_ * (lambda_est_ * pn * pn - an * an) + 2.0 * cross + dn * dn / sigma_;
        fp_ = std::sqrt(std::fmax(0.0, fp2));
    } else {
        const double movement = pn * pn * pw_ + dn * dn / pw_;
        fp_ = std::sqrt(movement + 2 * step_ * cross);  // no clamp at zero (:921)
    }
    if (has_inner_) {
        const double st = use_al_ ? step_ : step_ / pw_;
        inner_tol_ = std::fmin(inner_tol_, std::fmax(0.0005 * pn / st, 1e-9));
    }
}

void Solver::pcal_reset_sigma()
{   // reset_sigma_controller, :271-283
    double sg = std::fmax(best_sigma_, 1e-16);
    if (sigma_min_ > 0.0) sg = std::fmax(sg, sigma_min_);
    if (sigma_max_ > 0.0) sg = std::fmin(sg, sigma_max_);
    sigma_ = sg;
    step_ = pcal_tau();
}

void Solver::pcal_sigma_controller(double pd, double dd, double ax)
{   // the AL branch of perform_restart, :687-756: a log-damped step toward
    // ||dy|| / sqrt(lambda ||dx||^2 - ||A dx||^2), clipped to [0.6, 1.7] per restart,
    // rolled back to the best sigma when the restart metric worsens by 5 %
    double cur = fp_;
    const double dxm = lambda_est_ * pd * pd - ax * ax;
    const double ratio = rel_d_ / std::fmax(rel_p_, 1e-300);
    const double old = std::fmax(sigma_, 1e-16);
    bool rolled = false;
    if (std::isfinite(cur) && cur < best_term_metric_) {
        best_term_metric_ = cur;
        best_sigma_ = old;
        best_sigma_gap_ = std::fabs(std::log10(std::fmax(rel_d_, 1e-300) / std::fmax(rel_p_, 1e-300)));
    }
    if (!std::isfinite(cur)) {
        pcal_reset_sigma();
        rolled = true;
        cur = INFINITY;
    } else if (std::isfinite(best_term_metric_) && cur > 1.05 * best_term_metric_) {
        pcal_reset_sigma();
        rolled = true;
    } else {
        bool valid = false;
        double est = old;
        if (dd > 1e-16 && dd < 1e12 && dxm > 1e-16 && dxm < 1e24 && std::isfinite(sigma_) && ratio > 1e-8 &&
            ratio < 1e8) {
            est = dd / std::sqrt(dxm);
            valid = std::isfinite(est) && est > 1e-16;
        }
        if (valid) {
            const double target = old * std::exp(0.5 * std::log(est / old));
            if (std::isfinite(target) && target > 1e-16) sigma_ = old * std::fmax(0.60, std::fmin(target / old, 1.70));
        }
    }
    if (!std::isfinite(sigma_) || sigma_ <= 1e-16 || sigma_ >= 1e16) {
        pcal_reset_sigma();
        rolled = true;
    } else {
        sigma_ = std::fmin(std::fmax(sigma_, sigma_min_), sigma_max_);
        step_ = pcal_tau();
    }
    const double g = std::fabs(std::log10(std::fmax(rel_d_, 1e-300) / std::fmax(rel_p_, 1e-300)));
    if (g < best_sigma_gap_) best_sigma_gap_ = g;
    if (rolled) {
        sigma_ = best_sigma_;
        step_ = pcal_tau();
    }
}
#endif  // PDAL_PLANTED_CP_AL

#ifndef PDAL_LIN_CP_AL
bool Solver::trajectory_open() const
{   // the header row; false when the file cannot be written (the reference exits)
    const char *path = P_.trajectory_csv;
    if (!path || !path[0]) return true;
    FILE *f = fopen(path, "w");
    if (!f) {
        perror("Error opening trajectory CSV");
        return false;
    }
#if defined(PDAL_PLANTED_FA_CP) || defined(PDAL_PLANTED_CP_AL)  // the planted variants' solver.cu:35-49
    fprintf(f, "iteration,time_sec,rel_primal,rel_dual,rel_gap,kkt_residual,fixed_point_residual,restart,sigma,"
               "primal_weight,al_sigma,lambda_est,primal_step,dual_step\n");
#else
    fprintf(f, "iteration,time_sec,rel_primal,rel_dual,rel_gap,kkt_residual,fixed_point_residual,restart,al_sigma,"
               "sigma_update_count\n");
#endif
    fclose(f);
    return true;
}

bool Solver::trajectory_row(bool restart) const
{   // one row per checkpoint, written before a planned restart is performed
    const char *path = P_.trajectory_csv;
    if (!path || !path[0]) return true;
    FILE *f = fopen(path, "a");
    if (!f) {
        perror("Error appending trajectory CSV");
        return false;
    }
    const double kkt = std::fmax(rel_p_, std::fmax(rel_d_, rel_gap_));
    const double fpr = total_ == 0 ? NAN : fp_;
#if defined(PDAL_PLANTED_FA_CP) || defined(PDAL_PLANTED_CP_AL)
    double sigma = NAN, al_sigma = NAN, lambda = NAN, pstep = NAN, dstep = NAN;  // compute_trajectory_diagnostics
#ifdef PDAL_PLANTED_CP_AL
    if (use_al_) {
        sigma = sigma_;
        lambda = lambda_est_;
        pstep = step_;
        dstep = sigma_;
    } else if (pw_ > 0.0) {
        pstep = step_ / pw_;
        dstep = step_ * pw_;
    }
#else
    al_sigma = use_al_ ? al_sigma_ : NAN;
    if (pw_ > 0.0) {
        pstep = step_ / pw_;
        dstep = step_ * pw_;
    }
#endif
    fprintf(f, "%d,%.17e,%.17e,%.17e,%.17e,%.17e,%.17e,%d,%.17e,%.17e,%.17e,%.17e,%.17e,%.17e\n", total_,
            solve_time_, rel_p_, rel_d_, rel_gap_, kkt, fpr, restart ? 1 : 0, sigma, pw_, al_sigma, lambda, pstep,
            dstep);
#else
    fprintf(f, "%d,%.17e,%.17e,%.17e,%.17e,%.17e,%.17e,%d,%.17e,%d\n", total_, solve_time_, rel_p_, rel_d_,
            rel_gap_, kkt, fpr, restart ? 1 : 0, al_sigma_, sigma_update_count_);
#endif
    fclose(f);
    return true;
}
#endif

result_t *Solver::run()
{
#ifdef PDAL_LIN_CP_AL
    return run_linearized();
#else   // optimize, QP/FA_CP/src/solver.cu:74-168
    print_initial_info();
    const auto t_setup = std::chrono::steady_clock::now();
    upload();
    rescale();
    rescale_time_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_setup).count();
    init_state();
    if (nonconvexity_ < 0) inner_limit_ = 1;
#ifdef PDAL_PLANTED_CP_AL
    pcal_init_steps();
#else
    // initialize_step_size_and_primal_weight, fa_cp_core_op.cu:1022-1040
    step_ = A_.nnz == 0 ? 1.0 : 0.998 / max_singular_value();
    pw_ = P_.bound_objective_rescaling ? 1.0 : (c_norm_ + 1.0) / (b_norm_ + 1.0);
    best_pw_ = pw_;
#endif
    PDAL_HIP(hipStreamSynchronize(stream_));

    if (!trajectory_open()) return nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    const int K = P_.termination_evaluation_frequency;
    bool do_restart = false;
    while (total_ < P_.termination_criteria.iteration_limit) {
        const bool consider = major_ || total_ == 0;
        bool planned = false;
        if (consider || total_ % print_frequency(total_) == 0) {
            compute_residual();
            if (major_ && total_ < 3 * K) infeasibility_info();
            solve_time_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            check_termination();
            if (consider && reason_ == TERMINATION_REASON_UNSPECIFIED) {
#if !defined(PDAL_CP_AL_TYPES) && !defined(PDAL_PLANTED_FA_CP)
                choose_restart_candidate();  // solver.cu:112
#endif
                planned = should_restart();
            }
            if (!trajectory_row(planned)) return nullptr;
            display_stats();
            if (reason_ != TERMINATION_REASON_UNSPECIFIED) break;
        }
        if (consider) {
            do_restart = planned;
            if (do_restart) perform_restart();
        }
        major_ = ((total_ + 1) % K) == 0;
        const bool need_fp = major_ || do_restart;
        update(!need_fp);
        if (need_fp) {
            fixed_point_error();
            if (do_restart) {
                fp0_ = fp_;
                do_restart = false;
            }
            halpern();
        }
        inner_++;
        total_++;
    }
    if (reason_ == TERMINATION_REASON_UNSPECIFIED) {
        reason_ = TERMINATION_REASON_ITERATION_LIMIT;
        compute_residual();
        if (!trajectory_row(false)) return nullptr;
        display_stats();
    }
    if (P_.feasibility_polishing && reason_ != TERMINATION_REASON_DUAL_INFEASIBLE &&
        reason_ != TERMINATION_REASON_PRIMAL_INFEASIBLE)  // solver.cu:156-160 (planted :187-190)
        feasibility_polish();
    return make_result();
#endif
}

// ---------------------------------------------------------------------------
// Feasibility polishing (fa_cp_core_op.cu:1242-1735, utils.cu:1360-1575; the
// planted solver's QP/PLANTED_FA_CP/src/fa_cp_core_op.cu:1009-1502 is the same
// code, and the linearized solver's QP/LIN_CP_AL/src/cp_al_core_op.cu:839-1375
// differs only in the iteration it runs, see polish_phase): a
// primal phase with c = 0 (Q kept) and a dual phase with every finite bound 0,
// each the main loop's iteration with the primal weight fixed, its own
// fixed-point error (squared distance of the reflected from the current point,
// one side) and residual, restarts of its own side only, until its relative
// residual reaches eps_feas_polish_relative.  The phases run in place: the
// problem vectors are swapped for the phase's copies and the main iterates
// saved around them; the inner solver's state and counter carry on, as the
// reference's phase states share them.

void Solver::polish_residual(bool primal, const PolishRefs &ref, PolishOut &out)
{   // compute_primal_feas_polish_residual / compute_dual_feas_polish_residual, utils.cu:1491-1575
    int nba = 1, nbt = 1;
    const double *xa = primal ? xT_.get() : ref.x;
    dispatch_group(A_.G, [&](auto g) {
        constexpr int G = decltype(g)::value;
        nba = row_grid(m_, G);  // primal residual with the phase's (= original) bounds; dual contributions with the original ones
        res_a_kernel<G><<<nba, kBlock, 0, stream_>>>(A_.view(), xa, yT_.get(), primal ? l_.get() : ref.l,
                                                     primal ? u_.get() : ref.u, E_.get(), parts_.get());
    });
    PDAL_LAUNCHED();
    reduce(nba, 2, RA_N, S_RA);
    const StepScalars sc = scalars();
    dispatch_group(At_.G, [&](auto g) {
        constexpr int G = decltype(g)::value;
        nbt = row_grid(n_, G);  // (c - At.y - slack) D, c.x and slack.x against the original c / polished x
        res_at_kernel<G, false><<<nbt, kBlock, 0, stream_>>>(At_.view(), yT_.get(), primal ? xT_.get() : ref.x,
                                                             qx_.get(), primal ? ref.c : c_.get(), lb_.get(),
                                                             ub_.get(), D_.get(), dslack_.get(), sc, parts_.get());
    });
    PDAL_LAUNCHED();
    reduce(nbt, 6, RT_N, S_RT);
    fetch(S_RA, RA_N + RT_N);
    const bool linf = P_.optimality_norm == NORM_TYPE_L_INF;
    const double *ra = hscal_.get() + S_RA, *rt = hscal_.get() + S_RT;
    const double bc = b_scale_ * c_scale_;
    if (primal) {
        out.abs = (linf ? ra[RA_PRESMAX] : std::sqrt(ra[RA_PRES2])) / b_scale_;
        out.rel = out.abs / (1.0 + b_norm_);
        out.obj = rt[RT_CX] / bc;  // the phase's objective constant is 0
    } else {
        out.abs = (linf ? rt[RT_DRESMAX] : std::sqrt(rt[RT_DRES2])) / c_scale_;
        out.rel = out.abs / (1.0 + c_norm_);
        out.obj = (rt[RT_SX] + ra[RA_CONTRIB]) / bc + obj_const_;
    }
    abs_p_ = primal ? out.abs : 0.0;
    rel_p_ = primal ? out.rel : 0.0;
    abs_d_ = primal ? 0.0 : out.abs;
    rel_d_ = primal ? 0.0 : out.rel;
}

double Solver::polish_fixed_point_error(bool primal)
{   // compute_primal/dual_fixed_point_error, fa_cp_core_op.cu:1714-1735
    const int len = primal ? n_ : m_;
    const int nb = stride_grid(len);
    diff_norm_kernel<<<nb, kBlock, 0, stream_>>>(primal ? xR_.get() : yR_.get(), primal ? x_.get() : y_.get(),
                                                 primal ? dx_.get() : dy_.get(), len, parts_.get());
    reduce(nb, 1, 1, S_DN2);
    fetch(S_DN2, 1);
    const double ss = hs(S_DN2);
#ifdef PDAL_PLANTED_CP_AL
    // compute_primal/dual_fixed_point_error, PLANTED_CP_AL cp_al_core_op.cu:1495-1539
    if (!primal) return ss / (use_al_ ? sigma_ : pw_);
    if (!use_al_) return ss * pw_;
    int nba = 1;
    dispatch_group(A_.G, [&](auto g) {
        constexpr int G = decltype(g)::value;
        nba = row_grid(m_, G);
        spmv_norm_kernel<G><<<nba, kBlock, 0, stream_>>>(A_.view(), dx_.get(), prim_prod_.get(), parts_.get());
    });
    PDAL_LAUNCHED();
    reduce(nba, 1, 1, S_NORMAL);
    fetch(S_NORMAL, 1);
    const double pn = std::sqrt(ss), an = std::sqrt(hs(S_NORMAL));
    return std::fmax(0.0, sigma_ * (lambda_est_ * pn * pn - an * an));
#else
    return primal ? ss * pw_ : ss / pw_;
#endif
}

Solver::PolishOut Solver::polish_phase(bool primal, double pw0, const PolishRefs &ref)
{   // primal_feasibility_polish / dual_feasibility_polish, fa_cp_core_op.cu:1312-1404
    const termination_criteria_t &tc = P_.termination_criteria;
    const int K = P_.termination_evaluation_frequency;
    // scalar resets of initialize_primal/dual_feas_polish_state (:1474-1494, 1626-1646)
    pw_err_sum_ = pw_last_err_ = 0.0;
    pw_ = best_pw_ = pw0;
    fp_ = fp0_ = last_trial_fp_ = INFINITY;
    major_ = false;
    total_ = inner_ = 0;
    best_gap_ = INFINITY;
    drop_bb_graph();
    if (P_.verbose) {  // print_initial_feas_polish_info, utils.cu:1386-1407
        printf("---------------------------------------------------------------------------------------\n");
        printf("Starting %s Feasibility Polishing Phase with relative tolerance %.2e\n", primal ? "Primal" : "Dual",
               tc.eps_feas_polish_relative);
        printf("---------------------------------------------------------------------------------------\n");
    }
    const auto t0 = std::chrono::steady_clock::now();
    PolishOut out{TERMINATION_REASON_UNSPECIFIED, 0, 0.0, 0.0, 0.0, 0.0};
    bool do_restart = false;
    while (out.reason == TERMINATION_REASON_UNSPECIFIED) {
        const bool consider = major_ || total_ == 0;
        if (consider || total_ % print_frequency(total_) == 0) {
            polish_residual(primal, ref, out);
            out.time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            // check_feas_polishing_termination_criteria, utils.cu:1360-1384 (the phase's own clock)
            if (out.rel <= tc.eps_feas_polish_relative) out.reason = TERMINATION_REASON_FEAS_POLISH_SUCCESS;
            else if (total_ >= tc.iteration_limit) out.reason = TERMINATION_REASON_ITERATION_LIMIT;
            else if (out.time >= tc.time_sec_limit) out.reason = TERMINATION_REASON_TIME_LIMIT;
            if (P_.verbose >= 2 && total_ % print_frequency(total_) == 0)
                printf("%6d %.1e | %8.1e |    %.1e   |   %.1e   \n", total_, out.time, out.obj, out.abs, out.rel);
        }
        if (consider) {  // the restart test runs after a stop as well, as in the reference loop
#ifdef PDAL_LIN_CP_AL
            do_restart = lin_should_restart();
#else
            do_restart = should_restart();
#endif
            if (do_restart) {  // perform_primal_restart / perform_dual_restart, :1668-1690
                copy(primal ? x0_.get() : y0_.get(), primal ? xT_.get() : yT_.get(), primal ? n_ : m_);
                copy(primal ? x_.get() : y_.get(), primal ? xT_.get() : yT_.get(), primal ? n_ : m_);
                inner_ = 0;
                last_trial_fp_ = INFINITY;
            }
        }
        major_ = ((total_ + 1) % K) == 0;
#ifdef PDAL_LIN_CP_AL
        // cp_al_update, then halpern_update with the reflection coefficient.  The
        // phase's restart test (should_do_adaptive_restart, LIN_CP_AL utils.cu:451-463)
        // needs a finite reference error, which a phase starts without
        // (initialize_*_feas_polish_state set it to infinity) and only a restart sets:
        // it never fires, so the phase's fixed-point error (:1331-1375), read by
        // nothing else, is not computed.
        lin_update();
        lin_halpern();
#else
        const bool need_fp = major_ || do_restart;
        update(!need_fp);
        if (need_fp) {
            fp_ = polish_fixed_point_error(primal);
            if (do_restart) {
                fp0_ = fp_;
                do_restart = false;
            }
            halpern();
        }
#endif
        inner_++;
        total_++;
    }
    out.total = total_;
    drop_bb_graph();
    return out;
}

void Solver::feasibility_polish()
{   // feasibility_polish, fa_cp_core_op.cu:1242-1310
    const auto t_start = std::chrono::steady_clock::now();
    const termination_criteria_t &tc = P_.termination_criteria;
    if (rel_p_ < tc.eps_feas_polish_relative && rel_d_ < tc.eps_feas_polish_relative) {
        printf("Skipping feasibility polishing as the solution is already sufficiently feasible.\n");
        fflush(stdout);
        return;
    }
    const double pw0 = P_.bound_objective_rescaling ? 1.0 : (c_norm_ + 1.0) / (b_norm_ + 1.0);
    // the main state's scalars and vectors the phases overwrite
    const int m_total = total_, m_inner = inner_;
    const bool m_major = major_;
    const double m_fp = fp_, m_fp0 = fp0_, m_last = last_trial_fp_, m_pw = pw_, m_best_pw = best_pw_;
    const double m_err = pw_err_sum_, m_last_err = pw_last_err_, m_best_gap = best_gap_;
    double abs_p = abs_p_, rel_p = rel_p_, p_obj = p_obj_, abs_d = abs_d_, rel_d = rel_d_, d_obj = d_obj_;
    DBuf<double> sx(n_), sy0(m_), sy(m_), syT(m_), syR(m_), sdp(n_), sds(n_);
    copy(sx.get(), xT_.get(), n_);
    copy(sy0.get(), y0_.get(), m_);
    copy(sy.get(), y_.get(), m_);
    copy(syT.get(), yT_.get(), m_);
    copy(syR.get(), yR_.get(), m_);
    copy(sdp.get(), dual_prod_.get(), n_);
    copy(sds.get(), dslack_.get(), n_);
    auto zero = [&](double *v, int len) {
        if (len > 0) fill_kernel<<<blocks_for(len), kBlock, 0, stream_>>>(v, 0.0, len);
    };

    // primal phase (initialize_primal_feas_polish_state, :1412-1494): c = 0, constant 0,
    // dual iterates, products and slacks zero, primal iterates from the main state
    DBuf<double> zc(n_);
    zero(zc.get(), n_);
    std::swap(c_, zc);  // zc: the original c
    const double oc = obj_const_;
    obj_const_ = 0.0;
    for (double *v : {y0_.get(), y_.get(), yT_.get(), yR_.get(), al_u_.get(), pslack_.get(), dy_.get()}) zero(v, m_);
    for (double *v : {dual_prod_.get(), dslack_.get(), dx_.get()}) zero(v, n_);
    PDAL_LAUNCHED();
    const PolishOut pr = polish_phase(true, pw0, PolishRefs{zc.get(), l_.get(), u_.get(), nullptr});
    std::swap(c_, zc);
    obj_const_ = oc;
    if (pr.reason == TERMINATION_REASON_FEAS_POLISH_SUCCESS) {
        copy(sx.get(), xT_.get(), n_);
        abs_p = pr.abs;
        rel_p = pr.rel;
        p_obj = pr.obj;
    }
    polish_iters_ += pr.total - 1;

    // dual phase (initialize_dual_feas_polish_state, :1539-1646): finite bounds 0, primal
    // iterates and products zero, dual iterates, product and slack from the main state
    DBuf<double> zlb(n_), zub(n_), zl(m_), zu(m_);
    copy(zlb.get(), lb_.get(), n_);
    copy(zub.get(), ub_.get(), n_);
    copy(zl.get(), l_.get(), m_);
    copy(zu.get(), u_.get(), m_);
    for (auto [v, len] : {std::pair<double *, int>{zlb.get(), n_}, {zub.get(), n_}, {zl.get(), m_}, {zu.get(), m_}})
        if (len > 0) zero_finite_kernel<<<blocks_for(len), kBlock, 0, stream_>>>(v, len);
    std::swap(lb_, zlb);  // z*: the original bounds
    std::swap(ub_, zub);
    std::swap(l_, zl);
    std::swap(u_, zu);
    copy(y0_.get(), sy0.get(), m_);
    copy(y_.get(), sy.get(), m_);
    copy(yT_.get(), syT.get(), m_);
    copy(yR_.get(), syR.get(), m_);
    copy(dual_prod_.get(), sdp.get(), n_);
    copy(dslack_.get(), sds.get(), n_);
    for (double *v : {x0_.get(), x_.get(), xT_.get(), xR_.get(), dx_.get()}) zero(v, n_);
    for (double *v : {prim_prod_.get(), al_u_.get(), pslack_.get(), dy_.get()}) zero(v, m_);
    PDAL_LAUNCHED();
    const PolishOut du = polish_phase(false, pw0, PolishRefs{c_.get(), zl.get(), zu.get(), sx.get()});
    std::swap(lb_, zlb);
    std::swap(ub_, zub);
    std::swap(l_, zl);
    std::swap(u_, zu);
    if (du.reason == TERMINATION_REASON_FEAS_POLISH_SUCCESS) {
        abs_d = du.abs;
        rel_d = du.rel;
        d_obj = du.obj;
    } else {
        copy(yT_.get(), syT.get(), m_);
    }
    polish_iters_ += du.total - 1;
    copy(xT_.get(), sx.get(), n_);  // the result's x
    PDAL_HIP(hipStreamSynchronize(stream_));

    total_ = m_total;
    inner_ = m_inner;
    major_ = m_major;
    fp_ = m_fp;
    fp0_ = m_fp0;
    last_trial_fp_ = m_last;
    pw_ = m_pw;
    best_pw_ = m_best_pw;
    pw_err_sum_ = m_err;
    pw_last_err_ = m_last_err;
    best_gap_ = m_best_gap;
    abs_p_ = abs_p;
    rel_p_ = rel_p;
    p_obj_ = p_obj;
    abs_d_ = abs_d;
    rel_d_ = rel_d;
    d_obj_ = d_obj;
    gap_ = std::fabs(p_obj_ - d_obj_);
    rel_gap_ = gap_ / (1.0 + std::fabs(p_obj_) + std::fabs(d_obj_));
    // fa_cp_feas_polish_final_log, utils.cu:1409-1440
    if (P_.verbose) printf("---------------------------------------------------------------------------------------\n");
    printf("Feasibility Polishing Summary\n");
    printf("  Primal Status        : %s\n", termination_reason_name(pr.reason));
    printf("  Primal Iterations    : %d\n", pr.total);
    printf("  Primal Time Usage    : %.3g sec\n", pr.time);
    printf("  Dual Status          : %s\n", termination_reason_name(du.reason));
    printf("  Dual Iterations      : %d\n", du.total);
    printf("  Dual Time Usage      : %.3g sec\n", du.time);
    printf("  Primal Residual      : %.3e\n", pr.rel);
    printf("  Dual Residual        : %.3e\n", du.rel);
    printf("  Primal Dual Gap      : %.3e\n", std::fabs(pr.obj - du.obj) / (1.0 + std::fabs(pr.obj) + std::fabs(du.obj)));
    fflush(stdout);
    polish_time_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
}

}  // namespace

result_t *optimize_gpu(const params_t *params, const qp_problem_t *prob)
{
    Solver s(*params, prob);
    return s.run();
}

void reset_rng()
{
    g_gen.seed(1);
    g_dist.reset();
    srand(1);
}

const char *termination_reason_name(termination_reason_t r)
{
    switch (r) {
    case TERMINATION_REASON_OPTIMAL: return "OPTIMAL";
    case TERMINATION_REASON_PRIMAL_INFEASIBLE: return "PRIMAL_INFEASIBLE";
    case TERMINATION_REASON_DUAL_INFEASIBLE: return "DUAL_INFEASIBLE";
    case TERMINATION_REASON_INFEASIBLE_OR_UNBOUNDED: return "INFEASIBLE_OR_UNBOUNDED";
    case TERMINATION_REASON_TIME_LIMIT: return "TIME_LIMIT";
    case TERMINATION_REASON_ITERATION_LIMIT: return "ITERATION_LIMIT";
    case TERMINATION_REASON_FEAS_POLISH_SUCCESS: return "FEAS_POLISH_SUCCESS";
    default: return "UNSPECIFIED";
    }
}

}  // namespace qp
}  // namespace pdal
