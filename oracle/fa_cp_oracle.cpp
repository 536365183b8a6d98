/*
 * CPU FP64 oracle for the FA_CP (QP) restarted reflected-Halpern solver.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP path (libfa_cp.so) is compared
 * against.  Nothing in the product links, loads or calls it; only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.  It exports
 * the same C names as libfa_cp.so plus fa_cp_oracle_reset_rng().
 *
 * Statement-by-statement restatement of QP/FA_CP:
 *   optimize ............................ src/solver.cu:74-168
 *   rescale_problem (host) .............. src/preconditioner.c:36-331
 *   initialize_solver_state ............. src/solver_state.cu:494-877
 *   quadratic term, Q.x, xQx ............ src/solver_state.cu:220-460
 *   detect_q_type ....................... src/utils.cu:1576-1616
 *   power iterations (A, Q max/min) ..... src/utils.cu:65-315 (mt19937(1) and rand())
 *   termination / restart / defaults .... src/utils.cu:412-537
 *   residual, infeasibility ............. src/utils.cu:705-1132
 *   fa_cp_update, BB inner solve, dual,
 *   Halpern, restart, sigma rules,
 *   fixed-point error, result ........... src/fa_cp_core_op.cu:93-1239
 *   problem construction, conversions ... src/fa_cp.c:28-376, src/utils.cu:1135-1358
 *   feasibility polishing ............... src/fa_cp_core_op.cu:1242-1735, src/utils.cu:1360-1575
 *
 * Deliberate differences shared with the HIP path: wall-clock timing (also
 * in the trajectory CSV's time column, solver.cu:35-71), errors return NULL
 * instead of exit().
 *
 * Built a second time with -DFA_CP_ORACLE_REVERSE_SUMS, every SpMV row, dot
 * product and norm sums its terms in reverse order.  That variant is the same
 * algorithm with a different (equally valid) rounding; tests use the distance
 * between the two builds as the summation-order noise floor of a trajectory.
 *
 * Built with -DORACLE_LIN_CP_AL it is the oracle of the linearized CP-AL
 * solver instead (QP/LIN_CP_AL, types of lin_cp_al_types.h): the same
 * scaling, power iterations, residual, infeasibility and termination code
 * (LIN_CP_AL's utils.cu and solver_state.cu differ from FA_CP's only where
 * marked), with LIN_CP_AL's own step sizes (cp_al_core_op.cu:609-691), update
 * (:496-575: AL effective dual, one projected gradient step, equality dual),
 * relaxed Halpern step (:187-206, 577-585), fixed-point error every iteration
 * (:693-770), restart test (utils.cu:451-463) and restart (:595-607), main loop
 * (solver.cu:34-114); feasibility polishing as FA_CP's with the linearized
 * iteration and fixed-point errors (cp_al_core_op.cu:839-1375).
 *
 * Built with -DORACLE_PLANTED_FA_CP it is the oracle of RHR-FA-CP, the
 * planted-KKT FA_CP solver (QP/PLANTED_FA_CP): FA_CP's iteration and restart
 * test (its utils.cu differs only in defaults and printing), with equality
 * AL kernels (fa_cp_core_op.cu:147-183: right-hand side = lower bound), AL on
 * when al_sigma != 0 (solver_state.cu:498-504), its inner-loop rule
 * (:305-345), a fixed-point error without the clamp at zero (:893-940), and
 * its AL sigma update at restarts (:610-747, off unless adaptive_al_sigma)
 * in place of FA_CP's sigma rules; its feasibility polishing is FA_CP's code.
 *
 * Built with -DORACLE_PLANTED_CP_AL it is the oracle of RHR-CP-AL, the
 * planted-KKT CP-AL solver (QP/PLANTED_CP_AL): FA_CP's main loop and restart
 * test (solver.cu:111-204, utils.cu:451-480) over the linearized solver's
 * sigma / lambda steps with its own set-up (lambda = 2.001 ||A||^2, sigma =
 * sigma_base / 2; cp_al_core_op.cu:805-858), update (equality rows; LP, diagonal
 * Q in closed form, FA_CP's BB inner solve for the other Q types; T-outputs in
 * major-kernel iterations only; :531-635), FA_CP's Halpern step, fixed-point
 * error on the reflection (:860-935), and at restarts a sigma controller (AL
 * branch) or the primal-weight PID (:655-803); its polishing is the linearized
 * solver's code around this iteration.
 */
#if defined(ORACLE_PLANTED_FA_CP)
#include "planted_fa_cp.h"
typedef fa_cp_parameters_t params_t;
typedef fa_cp_result_t result_t;
#elif defined(ORACLE_LIN_CP_AL) || defined(ORACLE_PLANTED_CP_AL)
#define ORACLE_CP_AL_TYPES 1
#if defined(ORACLE_LIN_CP_AL)
#include "lin_cp_al.h"
#else
#include "planted_cp_al.h"
#endif
typedef cp_al_parameters_t params_t;
typedef cp_al_result_t result_t;
// the shared code names the Q types by their FA_CP names
constexpr quad_obj_type_t FA_CP_SPARSE_Q = CP_AL_SPARSE_Q;
constexpr quad_obj_type_t FA_CP_DIAG_Q = CP_AL_DIAG_Q;
constexpr quad_obj_type_t FA_CP_LOW_RANK_PLUS_SPARSE_Q = CP_AL_LOW_RANK_PLUS_SPARSE_Q;
constexpr quad_obj_type_t FA_CP_LOW_RANK_Q = CP_AL_LOW_RANK_Q;
constexpr quad_obj_type_t FA_CP_NON_Q = CP_AL_NON_Q;
#else
#include "fa_cp.h"
typedef fa_cp_parameters_t params_t;
typedef fa_cp_result_t result_t;
#endif

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#define ORACLE_API extern "C" __attribute__((visibility("default")))

namespace {

std::mt19937 g_gen(1);                              // utils.cu:28
std::normal_distribution<double> g_dist(0.0, 1.0);  // utils.cu:29

using Vec = std::vector<double>;

struct Csr {
    int rows = 0, cols = 0;
    std::vector<int> rp, ci;
    Vec v;
    int nnz() const { return rp.empty() ? 0 : rp[rows]; }
};

void spmv(const Csr &A, const double *x, double *y)
{
    for (int i = 0; i < A.rows; ++i) {
        double s = 0.0;
#ifdef FA_CP_ORACLE_REVERSE_SUMS
        for (int k = A.rp[i + 1] - 1; k >= A.rp[i]; --k) s += A.v[k] * x[A.ci[k]];
#else
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) s += A.v[k] * x[A.ci[k]];
#endif
        y[i] = s;
    }
}

double nrm2(const Vec &x)
{
    double s = 0.0;
#ifdef FA_CP_ORACLE_REVERSE_SUMS
    for (size_t i = x.size(); i-- > 0;) s += x[i] * x[i];
#else
    for (double a : x) s += a * a;
#endif
    return std::sqrt(s);
}

double dot(const Vec &a, const Vec &b)
{
    double s = 0.0;
#ifdef FA_CP_ORACLE_REVERSE_SUMS
    for (size_t i = a.size(); i-- > 0;) s += a[i] * b[i];
#else
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
#endif
    return s;
}

double inf_norm(const Vec &x)
{
    double m = 0.0;
    for (double a : x) m = std::max(m, std::fabs(a));
    return m;
}

double sum(const Vec &x)
{
    double s = 0.0;
    for (double a : x) s += a;
    return s;
}

Csr transpose(const Csr &A)
{
    Csr T;
    T.rows = A.cols;
    T.cols = A.rows;
    T.rp.assign(A.cols + 1, 0);
    const int nnz = A.nnz();
    for (int k = 0; k < nnz; ++k) T.rp[A.ci[k] + 1]++;
    for (int j = 0; j < A.cols; ++j) T.rp[j + 1] += T.rp[j];
    T.ci.resize(nnz);
    T.v.resize(nnz);
    std::vector<int> next(T.rp.begin(), T.rp.end() - 1);
    for (int i = 0; i < A.rows; ++i)
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const int p = next[A.ci[k]]++;
            T.ci[p] = i;
            T.v[p] = A.v[k];
        }
    return T;
}

Csr from_component(const CsrComponent *c, int rows, int cols, int nnz)
{
    Csr M;
    M.rows = rows;
    M.cols = cols;
    M.rp.assign(rows + 1, 0);
    if (c && c->row_ptr && nnz > 0) {
        M.rp.assign(c->row_ptr, c->row_ptr + rows + 1);
        M.ci.assign(c->col_ind, c->col_ind + nnz);
        M.v.assign(c->val, c->val + nnz);
    }
    return M;
}

double sane_positive(double v, double fb) { return (std::isfinite(v) && v > 0.0) ? v : fb; }
double score_component(double v) { return std::isfinite(v) ? std::fmax(v, 1e-300) : INFINITY; }
double clampd(double v, double lo, double hi) { return std::fmin(std::fmax(v, lo), hi); }

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

struct S {
    params_t P{};
    int n = 0, m = 0, r = 0;
    double obj_const = 0.0;
    Csr A, At, Q, R, Rt;
    quad_obj_type_t qtype = FA_CP_NON_Q;
    Vec qdiag;
    double q_norm = 0.0, nonconvexity = 0.0;
    bool is_lp = true;
    Vec c, lb, ub, l, u, lfin, ufin, E, D;
    double b_scale = 1.0, c_scale = 1.0, c_norm = 0.0, b_norm = 0.0;

    Vec x0, x, xT, xR, dual_prod, dslack, dres, dx, qx;
    Vec y0, y, yT, yR, prim_prod, al_u, pslack, pres, dy;
    Vec rx, ry;  // R x and scratch

    // inner solver
    bool has_inner = false;
    int inner_limit = 0;
    double inner_tol = 0.0;
    long inner_total = 0;
    Vec grad, dir, gbuf;

    // AL sigma state
    bool al_available = false, use_al = false;
    double al_sigma = 0.0, best_al_sigma = 0.0, best_sigma_score = INFINITY, last_sigma_score = INFINITY;
    double last_sigma_rp = INFINITY, last_sigma_rd = INFINITY, last_sigma_gap = INFINITY;
    int sigma_update_count = 0, sigma_dir = 1;
    // FA_CP restart candidate (choose_restart_candidate, fa_cp_core_op.cu:1130-1169):
    // 0 = the T output, 1 = the current (Halpern) iterate
    int restart_target = 0;

    double step = 0.0, pw = 1.0, best_pw = 1.0, pw_err_sum = 0.0, pw_last_err = 0.0, best_gap = INFINITY;
    // linearized CP-AL: the fixed penalty sigma and its interval, lambda of tau = 1 / (sigma lambda)
    double sigma = 0.0, sigma_base = 1.0, sigma_min = 0.0, sigma_max = 0.0, lambda_est = 1.0;
    // planted CP-AL: the sigma controller's best sigma and its restart metric and residual gap
    double best_sigma = 0.0, best_term_metric = INFINITY, best_sigma_gap = INFINITY;
    // planted FA_CP: iteration of the last AL sigma change, the inner tolerance a change restores
    int last_sigma_update_iter = 0;
    double inner_tol0 = 0.0;
    int total = 0, inner = 0;
    bool major = false;
    double fp = INFINITY, fp0 = INFINITY, last_trial_fp = INFINITY;
    double abs_p = 0, rel_p = 0, abs_d = 0, rel_d = 0, p_obj = 0, d_obj = 0, gap = 0, rel_gap = 0;
    double max_p_ray = 0, max_d_ray = 0, p_ray_obj = 0, d_ray_obj = 0;
    termination_reason_t reason = TERMINATION_REASON_UNSPECIFIED;
    double rescale_time = 0.0, solve_time = 0.0;
};

/* ---------------- preconditioner.c (host) ---------------- */

void scale_problem(S &s, const Vec &e, const Vec &d)
{   // preconditioner.c:92-138
    for (int i = 0; i < s.n; ++i) {
        s.c[i] /= d[i];
        s.ub[i] *= d[i];
        s.lb[i] *= d[i];
    }
    for (int i = 0; i < s.m; ++i) {
        s.l[i] /= e[i];
        s.u[i] /= e[i];
    }
    for (int row = 0; row < s.m; ++row)
        for (int k = s.A.rp[row]; k < s.A.rp[row + 1]; ++k) s.A.v[k] /= (e[row] * d[s.A.ci[k]]);
    for (int row = 0; row < s.n; ++row)
        for (int k = s.Q.rp[row]; k < s.Q.rp[row + 1]; ++k) s.Q.v[k] /= (d[row] * d[s.Q.ci[k]]);
    if (s.R.nnz() > 0)
        for (int row = 0; row < s.r; ++row)
            for (int k = s.R.rp[row]; k < s.R.rp[row + 1]; ++k) s.R.v[k] /= d[s.R.ci[k]];
}

void rescale(S &s)
{   // rescale_problem, preconditioner.c:291-331
    s.E.assign(s.m, 1.0);
    s.D.assign(s.n, 1.0);
    Vec e(s.m), d(s.n);
    for (int it = 0; it < s.P.l_inf_ruiz_iterations; ++it) {  // ruiz_rescaling, :140-203
        std::fill(d.begin(), d.end(), 0.0);
        std::fill(e.begin(), e.end(), 0.0);
        for (int row = 0; row < s.m; ++row)
            for (int k = s.A.rp[row]; k < s.A.rp[row + 1]; ++k) {
                const double v = std::fabs(s.A.v[k]);
                const int col = s.A.ci[k];
                if (v > d[col]) d[col] = v;
                if (v > e[row]) e[row] = v;
            }
        for (int row = 0; row < s.n; ++row)
            for (int k = s.Q.rp[row]; k < s.Q.rp[row + 1]; ++k) {
                const double v = std::fabs(s.Q.v[k]);
                if (v > d[s.Q.ci[k]]) d[s.Q.ci[k]] = v;
            }
        for (int i = 0; i < s.n; ++i) d[i] = (d[i] < 1e-12) ? 1.0 : std::sqrt(d[i]);
        for (int i = 0; i < s.m; ++i) e[i] = (e[i] < 1e-12) ? 1.0 : std::sqrt(e[i]);
        scale_problem(s, e, d);
        for (int i = 0; i < s.n; ++i) s.D[i] *= d[i];
        for (int i = 0; i < s.m; ++i) s.E[i] *= e[i];
    }
    if (s.P.has_pock_chambolle_alpha) {  // pock_chambolle_rescaling, :205-247
        const double a = s.P.pock_chambolle_alpha;
        std::fill(d.begin(), d.end(), 0.0);
        std::fill(e.begin(), e.end(), 0.0);
        for (int row = 0; row < s.m; ++row)
            for (int k = s.A.rp[row]; k < s.A.rp[row + 1]; ++k) {
                const double v = std::fabs(s.A.v[k]);
                d[s.A.ci[k]] += std::pow(v, 2.0 - a);
                e[row] += std::pow(v, a);
            }
        for (int row = 0; row < s.n; ++row)
            for (int k = s.Q.rp[row]; k < s.Q.rp[row + 1]; ++k) d[s.Q.ci[k]] += std::pow(std::fabs(s.Q.v[k]), 2.0 - a);
        for (int i = 0; i < s.n; ++i) d[i] = (d[i] < 1e-12) ? 1.0 : std::sqrt(d[i]);
        for (int i = 0; i < s.m; ++i) e[i] = (e[i] < 1e-12) ? 1.0 : std::sqrt(e[i]);
        scale_problem(s, e, d);
        for (int i = 0; i < s.n; ++i) s.D[i] *= d[i];
        for (int i = 0; i < s.m; ++i) s.E[i] *= e[i];
    }
    if (s.P.bound_objective_rescaling) {  // bound_obj_rescaling, :249-289
        double bsq = 0.0;
        for (int i = 0; i < s.m; ++i) {
            if (std::isfinite(s.l[i]) && s.l[i] != s.u[i]) bsq += s.l[i] * s.l[i];
            if (std::isfinite(s.u[i])) bsq += s.u[i] * s.u[i];
        }
        double csq = 0.0;
        for (int i = 0; i < s.n; ++i) csq += s.c[i] * s.c[i];
        s.b_scale = 1.0 / (std::sqrt(bsq) + 1.0);
        s.c_scale = 1.0 / (std::sqrt(csq) + 1.0);
        for (int i = 0; i < s.m; ++i) {
            s.l[i] *= s.b_scale;
            s.u[i] *= s.b_scale;
        }
        for (int i = 0; i < s.n; ++i) {
            s.lb[i] *= s.b_scale;
            s.ub[i] *= s.b_scale;
            s.c[i] *= s.c_scale;
        }
        for (double &q : s.Q.v) q *= s.c_scale / s.b_scale;
        if (s.R.nnz() > 0) {
            const double f = std::sqrt(s.c_scale / s.b_scale);
            for (double &q : s.R.v) q *= f;
        }
    }
}

/* ---------------- Q helpers (solver_state.cu) ---------------- */

quad_obj_type_t detect(const S &s)
{   // detect_q_type, utils.cu:1576-1616 (called with the scaled problem)
    const int ns = s.Q.nnz(), nr = s.R.nnz();
    if (ns == 0 && nr == 0) return FA_CP_NON_Q;
    if (nr > 0) return ns > 0 ? FA_CP_LOW_RANK_PLUS_SPARSE_Q : FA_CP_LOW_RANK_Q;
    for (int i = 0; i < s.n; ++i)
        for (int k = s.Q.rp[i]; k < s.Q.rp[i + 1]; ++k)
            if (s.Q.ci[k] != i) return FA_CP_SPARSE_Q;
    return FA_CP_DIAG_Q;
}

void obj_product(S &s, const Vec &xv)
{   // update_obj_product, solver_state.cu:371-442
    switch (s.qtype) {
    case FA_CP_NON_Q: return;
    case FA_CP_SPARSE_Q: spmv(s.Q, xv.data(), s.qx.data()); return;
    case FA_CP_DIAG_Q:
        for (int i = 0; i < s.n; ++i) s.qx[i] = s.qdiag[i] * xv[i];
        return;
    case FA_CP_LOW_RANK_Q:
        spmv(s.R, xv.data(), s.rx.data());
        spmv(s.Rt, s.rx.data(), s.qx.data());
        return;
    case FA_CP_LOW_RANK_PLUS_SPARSE_Q: {
        spmv(s.Q, xv.data(), s.qx.data());
        spmv(s.R, xv.data(), s.rx.data());
        for (int i = 0; i < s.n; ++i) {
            double t = 0.0;
            for (int k = s.Rt.rp[i]; k < s.Rt.rp[i + 1]; ++k) t += s.Rt.v[k] * s.rx[s.Rt.ci[k]];
            s.qx[i] = t + s.qx[i];
        }
        return;
    }
    default: return;
    }
}

double max_singular_value(const S &s)
{   // estimate_maximum_singular_value, utils.cu:65-170
    const int m = s.m;
    Vec v(m), nv(m), w(s.n);
    for (int i = 0; i < m; ++i) v[i] = g_dist(g_gen);
    double sig = 1.0;
    for (int it = 0; it < s.P.sv_max_iter; ++it) {
        nv = v;
        const double inv = 1.0 / nrm2(nv);
        for (double &a : nv) a *= inv;
        spmv(s.At, nv.data(), w.data());
        spmv(s.A, w.data(), v.data());
        sig = dot(nv, v);
        const double neg = -sig;
        for (int i = 0; i < m; ++i) nv[i] = neg * nv[i] + v[i];
        if (nrm2(nv) < s.P.sv_tol) break;
    }
    return std::sqrt(sig);
}

double q_max_eig(const S &s)
{   // estimate_maximum_eigenvalue, utils.cu:172-238
    const int n = s.n;
    Vec v(n), av(n);
    for (int i = 0; i < n; ++i) v[i] = (double)rand() / RAND_MAX;
    double lam = 0.0;
    for (int it = 0; it < s.P.sv_max_iter; ++it) {
        const double inv = 1.0 / nrm2(v);
        for (double &a : v) a *= inv;
        spmv(s.Q, v.data(), av.data());
        const double old = lam;
        lam = dot(v, av);
        if (it > 0 && std::fabs(lam - old) < s.P.sv_tol) break;
        v = av;
    }
    return lam;
}

double q_min_eig(const S &s, double lmax)
{   // estimate_minimum_eigenvalue, utils.cu:240-315
    const int n = s.n;
    Vec v(n), av(n), sh(n);
    for (int i = 0; i < n; ++i) v[i] = (double)rand() / RAND_MAX;
    double mu = 0.0;
    for (int it = 0; it < s.P.sv_max_iter; ++it) {
        const double inv = 1.0 / nrm2(v);
        for (double &a : v) a *= inv;
        spmv(s.Q, v.data(), av.data());
        for (int i = 0; i < n; ++i) sh[i] = lmax * v[i] + (-av[i]);
        const double old = mu;
        mu = dot(v, sh);
        if (it > 0 && std::fabs(mu - old) < s.P.sv_tol) break;
        v = sh;
    }
    return lmax - mu;
}

/* ---------------- core op (fa_cp_core_op.cu) ---------------- */

bool major_kernel(const S &s) { return s.major || ((s.total + 2) % print_frequency(s.total + 2)) == 0; }

void al_effective_dual_product(S &s, const Vec &xv, double sigma)
{   // compute_al_effective_dual_product, :275-303 and kernel :181-192
    spmv(s.A, xv.data(), s.prim_prod.data());
    for (int i = 0; i < s.m; ++i) {
#ifdef ORACLE_PLANTED_FA_CP
        s.al_u[i] = s.y[i] - sigma * (s.prim_prod[i] - s.l[i]);  // planted: equality rows (:147-156)
#else
        const double sh = s.prim_prod[i] - s.y[i] / sigma;
        const double pr = std::fmax(s.l[i], std::fmin(sh, s.u[i]));
        s.al_u[i] = s.y[i] - sigma * (s.prim_prod[i] - pr);
#endif
    }
    spmv(s.At, s.al_u.data(), s.dual_prod.data());
}

void primal_bb(S &s, double tau, double al_sigma)
{   // primal_BB_step_size_update, :445-503
    const double inv = 1.0 / tau;
    int it = 1;
    double alpha = 1.0 / inv;
    if (s.use_al) al_effective_dual_product(s, s.x, al_sigma);
    obj_product(s, s.x);
    for (int i = 0; i < s.n; ++i) {  // bb_init kernel :381-399
        const double g = s.qx[i] + s.c[i] - s.dual_prod[i];
        double nx = s.x[i] - alpha * g;
        nx = std::fmax(s.lb[i], std::fmin(nx, s.ub[i]));
        s.xT[i] = nx;
        s.grad[i] = g;
        s.dir[i] = nx - s.x[i];
    }
    while (it < s.inner_limit) {
        const double nd = nrm2(s.dir);
        if (nd <= s.inner_tol) break;
        if (s.use_al) al_effective_dual_product(s, s.xT, al_sigma);
        obj_product(s, s.xT);
        for (int i = 0; i < s.n; ++i) {  // :401-415
            const double last = s.grad[i];
            const double cur = s.qx[i] + s.c[i] - s.dual_prod[i] + inv * (s.xT[i] - s.x[i]);
            s.gbuf[i] = cur - last;
            s.grad[i] = cur;
        }
        const double tmp = dot(s.dir, s.gbuf);
        alpha = nd * nd / tmp;
        for (int i = 0; i < s.n; ++i) {  // :417-430
            const double cs = s.xT[i];
            double nx = cs - alpha * s.grad[i];
            nx = std::fmax(s.lb[i], std::fmin(nx, s.ub[i]));
            s.dir[i] = nx - cs;
            s.xT[i] = nx;
        }
        ++it;
    }
    for (int i = 0; i < s.n; ++i) s.xR[i] = 2.0 * s.xT[i] - s.x[i];
    s.inner_total += it - 1;
}

void fa_cp_update(S &s)
{   // :531-630
    double tau = s.step / s.pw;
    if (s.nonconvexity < 0) {
        tau = std::fmax(tau, -1.01 * std::fmin(0.0, s.nonconvexity));
        tau /= 2;
    }
    const double sigma = s.step * s.pw;
    const double al_sigma = s.use_al ? s.al_sigma : 0.0;
    const bool mk = major_kernel(s);
    if (!s.use_al) spmv(s.At, s.y.data(), s.dual_prod.data());
    if (s.use_al) {
        primal_bb(s, tau, al_sigma);
    } else {
        switch (s.qtype) {
        case FA_CP_NON_Q:  // lp_primal_update, kernels :93-121
            for (int i = 0; i < s.n; ++i) {
                const double xi = s.x[i];
                const double t = xi - tau * (s.c[i] - s.dual_prod[i]);
                const double p = std::fmax(s.lb[i], std::fmin(t, s.ub[i]));
                s.xR[i] = 2.0 * p - xi;
                if (mk) {
                    s.xT[i] = p;
                    s.dslack[i] = (p - t) / tau;
                }
            }
            break;
        case FA_CP_DIAG_Q:  // diag_q_primal_update, kernels :123-154
            for (int i = 0; i < s.n; ++i) {
                const double xi = s.x[i];
                const double t = (xi - tau * (s.c[i] - s.dual_prod[i])) / (1.0 + tau * s.qdiag[i]);
                const double p = std::fmax(s.lb[i], std::fmin(t, s.ub[i]));
                s.xR[i] = 2.0 * p - xi;
                if (mk) s.xT[i] = p;
            }
            break;
        default: primal_bb(s, tau, 0.0); break;
        }
    }
    spmv(s.A, s.xR.data(), s.prim_prod.data());
    if (s.use_al) {  // kernels :194-223
        for (int i = 0; i < s.m; ++i) {
#ifdef ORACLE_PLANTED_FA_CP
            (void)al_sigma;
            const double nd = s.y[i] + sigma * (s.l[i] - s.prim_prod[i]);  // planted: equality rows (:158-183)
#else
            const double den = al_sigma + sigma;
            const double sh = s.prim_prod[i] - s.y[i] / den;
            const double pr = std::fmax(s.l[i], std::fmin(sh, s.u[i]));
            const double nd = s.y[i] + sigma * (pr - s.prim_prod[i]);
#endif
            if (mk) s.yT[i] = nd;
            s.yR[i] = 2.0 * nd - s.y[i];
        }
    } else {  // kernels :156-179
        for (int i = 0; i < s.m; ++i) {
            const double t = s.y[i] / sigma - s.prim_prod[i];
            const double p = std::fmax(-s.u[i], std::fmin(t, -s.l[i]));
            if (mk) {
                s.yT[i] = (t - p) * sigma;
                s.yR[i] = 2.0 * s.yT[i] - s.y[i];
            } else {
                s.yR[i] = 2.0 * (t - p) * sigma - s.y[i];
            }
        }
    }
}

void halpern(S &s)
{   // halpern_update, :632-640 and kernel :225-242
    const double w = (double)(s.inner + 1) / (s.inner + 2);
    const double rho = s.P.reflection_coefficient;
    for (int i = 0; i < s.n; ++i) {
        const double rf = rho * s.xR[i] + (1.0 - rho) * s.x[i];
        s.x[i] = w * rf + (1.0 - w) * s.x0[i];
    }
    for (int i = 0; i < s.m; ++i) {
        const double rf = rho * s.yR[i] + (1.0 - rho) * s.y[i];
        s.y[i] = w * rf + (1.0 - w) * s.y0[i];
    }
}

void fixed_point_error(S &s)
{   // :1042-1089
    for (int i = 0; i < s.n; ++i) s.dx[i] = s.xR[i] - s.x[i];
    for (int i = 0; i < s.m; ++i) s.dy[i] = s.yR[i] - s.y[i];
    spmv(s.At, s.dy.data(), s.dual_prod.data());
    const double dn = nrm2(s.dy), pn = nrm2(s.dx);
    const double movement = pn * pn * s.pw + dn * dn / s.pw;
    const double cross = dot(s.dual_prod, s.dx);
#ifdef ORACLE_PLANTED_FA_CP
    s.fp = std::sqrt(movement + 2 * s.step * cross);  // planted: no clamp at zero (:928-930)
#else
    s.fp = std::sqrt(std::fmax(movement + 2 * s.step * cross, 0.0));
#endif
    if (s.has_inner) s.inner_tol = std::fmin(s.inner_tol, std::fmax(0.0005 * pn / s.step * s.pw, 1e-9));
}

void compute_residual(S &s, norm_type_t norm)
{   // utils.cu:889-1026
    spmv(s.A, s.xT.data(), s.prim_prod.data());
    spmv(s.At, s.yT.data(), s.dual_prod.data());
    obj_product(s, s.xT);
    for (int i = 0; i < s.m; ++i) {
        const double cl = std::fmax(s.l[i], std::fmin(s.prim_prod[i], s.u[i]));
        s.pres[i] = (s.prim_prod[i] - cl) * s.E[i];
        s.pslack[i] = std::fmax(s.yT[i], 0.0) * s.lfin[i] + std::fmin(s.yT[i], 0.0) * s.ufin[i];
    }
    if (s.is_lp) {  // compute_lp_residual_kernel :705-733
        for (int j = 0; j < s.n; ++j) s.dres[j] = (s.c[j] - s.dual_prod[j] - s.dslack[j]) * s.D[j];
    } else {  // compute_qp_residual_kernel :735-770
#ifdef ORACLE_CP_AL_TYPES
        // LIN_CP_AL utils.cu:855-858 (PLANTED_CP_AL :871-874): the AL step 1 / (sigma lambda)
        const double tr = s.use_al ? 1.0 / std::fmax(s.sigma * s.lambda_est, 1e-16) : s.step / s.pw;
#else
        const double tr = s.step / s.pw;
#endif
        for (int j = 0; j < s.n; ++j) {
            const double g = s.qx[j] + s.c[j] - s.dual_prod[j];
            const double t = s.xT[j] - tr * g;
            const double p = std::fmax(s.lb[j], std::fmin(s.ub[j], t));
            const double ds = (p - t) / tr;
            s.dres[j] = (g - ds) * s.D[j];
            s.dslack[j] = ds;
        }
    }
    const bool linf = norm == NORM_TYPE_L_INF;
    s.abs_p = (linf ? inf_norm(s.pres) : nrm2(s.pres)) / s.b_scale;
    s.abs_d = (linf ? inf_norm(s.dres) : nrm2(s.dres)) / s.c_scale;
    const double half = s.qtype == FA_CP_NON_Q ? 0.0 : 0.5 * dot(s.xT, s.qx);
    const double bc = s.b_scale * s.c_scale;
    s.p_obj = (dot(s.c, s.xT) + half) / bc + s.obj_const;
    s.d_obj = (dot(s.dslack, s.xT) + sum(s.pslack) - half) / bc + s.obj_const;
    s.rel_p = s.abs_p / (1.0 + s.b_norm);
    double den;
    if (s.is_lp) {
        den = 1.0 + s.c_norm;
    } else {  // recover_primal_obj_dual_product :772-781 mutates both in place
        for (int j = 0; j < s.n; ++j) {
            s.dual_prod[j] = s.dual_prod[j] * s.D[j];
            s.qx[j] = s.qx[j] * s.D[j];
        }
        const double qn = linf ? inf_norm(s.qx) : nrm2(s.qx);
        const double an = linf ? inf_norm(s.dual_prod) : nrm2(s.dual_prod);
        den = 1.0 + std::fmax(s.c_norm, std::fmax(qn / s.c_scale, an / s.c_scale));
    }
    s.rel_d = s.abs_d / den;
    s.gap = std::fabs(s.p_obj - s.d_obj);
    s.rel_gap = s.gap / (1.0 + std::fabs(s.p_obj) + std::fabs(s.d_obj));
}

void infeasibility_info(S &s)
{   // compute_infeasibility_information, utils.cu:1028-1132
    for (int i = 0; i < s.n; ++i) {
        if (std::isfinite(s.lb[i])) s.dx[i] = std::fmax(s.dx[i], 0.0);
        if (std::isfinite(s.ub[i])) s.dx[i] = std::fmin(s.dx[i], 0.0);
    }
    for (int i = 0; i < s.m; ++i) {
        if (!std::isfinite(s.l[i])) s.dy[i] = std::fmin(s.dy[i], 0.0);
        if (!std::isfinite(s.u[i])) s.dy[i] = std::fmax(s.dy[i], 0.0);
    }
    const double pinf = inf_norm(s.dx);
    if (pinf > 0.0) {
        const double sc = 1.0 / pinf;
        for (double &a : s.dx) a *= sc;
    }
    const double dinf = inf_norm(s.dy);
    spmv(s.A, s.dx.data(), s.prim_prod.data());
    spmv(s.At, s.dy.data(), s.dual_prod.data());
    const double bc = s.b_scale * s.c_scale;
    s.p_ray_obj = dot(s.c, s.dx) / bc;
    for (int i = 0; i < s.m; ++i)
        s.pslack[i] = std::fmax(s.dy[i], 0.0) * s.lfin[i] + std::fmin(s.dy[i], 0.0) * s.ufin[i];
    // the reference passes dual_product as the "dual slack" input and writes dual_slack
    for (int j = 0; j < s.n; ++j) {
        const double lbf = std::isfinite(s.lb[j]) ? s.lb[j] : 0.0;
        const double ubf = std::isfinite(s.ub[j]) ? s.ub[j] : 0.0;
        s.dslack[j] = std::fmax(-s.dual_prod[j], 0.0) * lbf + std::fmin(-s.dual_prod[j], 0.0) * ubf;
    }
    s.d_ray_obj = (sum(s.pslack) + sum(s.dslack)) / bc;
    for (int i = 0; i < s.m; ++i) {
        const double pp = s.prim_prod[i];
        s.pslack[i] = (std::fmax(0.0, -pp) * (double)std::isfinite(s.l[i]) +
                       std::fmax(0.0, pp) * (double)std::isfinite(s.u[i])) *
                      s.E[i];
    }
    for (int j = 0; j < s.n; ++j) {
        const double dp = -s.dual_prod[j];
        s.dslack[j] = (std::fmax(0.0, dp) * (double)!std::isfinite(s.lb[j]) -
                       std::fmin(0.0, dp) * (double)!std::isfinite(s.ub[j])) *
                      s.D[j];
    }
    s.max_p_ray = inf_norm(s.pslack);
    if (!s.is_lp && s.qtype != FA_CP_NON_Q) {
        obj_product(s, s.dx);
        s.max_p_ray = std::fmax(s.max_p_ray, inf_norm(s.qx) / s.c_scale);
    }
    const double dsn = inf_norm(s.dslack);
    s.max_d_ray = dsn;
    const double f = std::fmax(dinf, dsn);
    if (f > 0.0) {
        s.max_d_ray /= f;
        s.d_ray_obj /= f;
    } else {
        s.max_d_ray = 0.0;
        s.d_ray_obj = 0.0;
    }
}

void check_termination(S &s)
{   // utils.cu:438-462
    const auto &tc = s.P.termination_criteria;
    if (s.rel_d < tc.eps_feasible_relative && s.rel_p < tc.eps_feasible_relative && s.rel_gap < tc.eps_optimal_relative)
        s.reason = TERMINATION_REASON_OPTIMAL;
    else if (s.d_ray_obj > 0.0 && s.max_d_ray / s.d_ray_obj <= tc.eps_infeasible)
        s.reason = TERMINATION_REASON_PRIMAL_INFEASIBLE;
    else if (s.p_ray_obj < 0.0 && s.max_p_ray / (-s.p_ray_obj) <= tc.eps_infeasible)
        s.reason = TERMINATION_REASON_DUAL_INFEASIBLE;
    else if (s.total >= tc.iteration_limit)
        s.reason = TERMINATION_REASON_ITERATION_LIMIT;
    else if (s.solve_time >= tc.time_sec_limit)
        s.reason = TERMINATION_REASON_TIME_LIMIT;
}

bool should_restart(S &s)
{   // utils.cu:464-492
    const auto &rp = s.P.restart_params;
    const int K = s.P.termination_evaluation_frequency;
    bool r = false;
    if (s.total == K) {
        r = true;
    } else if (s.total > K) {
        if (s.fp <= rp.sufficient_reduction_for_restart * s.fp0) r = true;
        if (s.fp <= rp.necessary_reduction_for_restart * s.fp0 && s.fp > s.last_trial_fp) r = true;
        if (s.inner >= rp.artificial_restart_threshold * s.total) r = true;
    }
    s.last_trial_fp = s.fp;
    return r;
}

#if !defined(ORACLE_CP_AL_TYPES) && !defined(ORACLE_PLANTED_FA_CP)
void set_sigma_active(S &s, double sigma, double zero_thr)
{   // :667-676
    if (!std::isfinite(sigma) || sigma <= zero_thr) {
        s.al_sigma = 0.0;
        s.use_al = false;
    } else {
        s.al_sigma = sigma;
        s.use_al = s.al_available;
    }
}

double clipped_update(const S &s, double cur, double ratio)
{   // :700-713
    const double gain = sane_positive(s.P.sigma_update_gain, 0.5);
    double lo = sane_positive(s.P.sigma_clip_low, 0.6), hi = sane_positive(s.P.sigma_clip_high, 1.7);
    if (lo > hi) std::swap(lo, hi);
    ratio = sane_positive(ratio, 1.0);
    return cur * clampd(std::pow(ratio, gain), lo, hi);
}

void update_al_sigma(S &s, double pd, double dd, double normal_motion)
{   // update_al_sigma_at_restart, :722-906
    const auto &P = s.P;
    const bool bal = P.sigma_update_rule == FA_CP_SIGMA_UPDATE_BALANCED_GUARDED;
    const bool guarded = P.sigma_update_rule == FA_CP_SIGMA_UPDATE_GUARDED || bal;
    if (!s.al_available || P.sigma_update_rule == FA_CP_SIGMA_UPDATE_FIXED) return;
    if (!guarded && !s.use_al) return;
    const double ref = bal ? sane_positive(s.step * s.pw, 1.0) : 1.0;
    double smin = sane_positive(P.sigma_min, 1e-8) * ref;
    double smax = sane_positive(P.sigma_max, 10.0) * ref;
    const double zthr = sane_positive(P.sigma_zero_threshold, sane_positive(P.sigma_min, 1e-8)) * ref;
    if (smin > smax) std::swap(smin, smax);
    if (zthr > smin) smin = zthr;
    const double cur = s.use_al ? clampd(sane_positive(s.al_sigma, smin), smin, smax) : 0.0;
    set_sigma_active(s, cur, zthr);
    const double rp = sane_positive(s.rel_p, 1e-300), rd = sane_positive(s.rel_d, 1e-300),
                 rg = sane_positive(s.rel_gap, 1e-300);
    const double score = std::fmax(score_component(s.rel_p), std::fmax(score_component(s.rel_d), score_component(s.rel_gap)));
    if (!std::isfinite(score)) return;
    if (!std::isfinite(s.best_sigma_score) || score < s.best_sigma_score) {
        s.best_sigma_score = score;
        s.best_al_sigma = cur;
    }
    if (P.sigma_rollback_tol > 0.0 && (!guarded || cur > zthr) && std::isfinite(s.best_sigma_score) &&
        s.best_sigma_score > 0.0 && score > P.sigma_rollback_tol * s.best_sigma_score &&
        std::isfinite(s.best_al_sigma) && (guarded ? s.best_al_sigma >= 0.0 : s.best_al_sigma > 0.0)) {
        const double best = guarded && s.best_al_sigma <= zthr ? 0.0 : clampd(s.best_al_sigma, smin, smax);
        set_sigma_active(s, best, zthr);
        s.last_sigma_score = score;
        s.last_sigma_rp = rp;
        s.last_sigma_rd = rd;
        s.last_sigma_gap = rg;
        s.sigma_update_count++;
        return;
    }
    double prop = cur;
    switch (P.sigma_update_rule) {
    case FA_CP_SIGMA_UPDATE_RESIDUAL_BALANCE: prop = clipped_update(s, cur, rp / rd); break;
    case FA_CP_SIGMA_UPDATE_KKT_BALANCE: prop = clipped_update(s, cur, rp / std::fmax(rd, rg)); break;
    case FA_CP_SIGMA_UPDATE_DISPLACEMENT:
        if (std::isfinite(pd) && std::isfinite(dd) && pd > 1e-16 && dd > 1e-16) {
            const double target = clampd(dd / pd, smin, smax);
            prop = clipped_update(s, cur, target / cur);
        }
        break;
    case FA_CP_SIGMA_UPDATE_PROGRESS: {
        const double lo = std::fmax(s.last_sigma_rd, s.last_sigma_gap);
        const double other = std::fmax(rd, rg);
        if (std::isfinite(s.last_sigma_rp) && std::isfinite(lo) && s.last_sigma_rp > 0.0 && lo > 0.0)
            prop = clipped_update(s, cur, (lo / other) / (s.last_sigma_rp / rp));
        else
            prop = clipped_update(s, cur, rp / std::fmax(rd, rg));
        break;
    }
    case FA_CP_SIGMA_UPDATE_GREEDY_SCORE:
        if (s.sigma_update_count == 0) s.sigma_dir = (rp >= std::fmax(rd, rg)) ? 1 : -1;
        else if (std::isfinite(score) && std::isfinite(s.last_sigma_score) && score > s.last_sigma_score)
            s.sigma_dir *= -1;
        prop = clipped_update(s, cur, s.sigma_dir > 0 ? P.sigma_clip_high : P.sigma_clip_low);
        break;
    case FA_CP_SIGMA_UPDATE_GUARDED:
    case FA_CP_SIGMA_UPDATE_BALANCED_GUARDED: {
        const double nthr = sane_positive(P.sigma_normal_motion_threshold, 1e-10);
        const bool normal_ok = !std::isfinite(normal_motion) || normal_motion > nthr;
        const bool zero_ok = !std::isfinite(normal_motion) || normal_motion > std::fmax(nthr, 2e-1);
        const bool bottleneck = rp >= sane_positive(P.sigma_activation_threshold, 2.0) *
                                          sane_positive(std::fmax(rd, rg), 1e-300);
        const bool zero_stalled =
            cur <= zthr && std::isfinite(s.last_sigma_score) && score > 0.9 * s.last_sigma_score;
        const bool primal_signal = rp >= 0.5 * sane_positive(rd, 1e-300);
        const double act = sane_positive(P.sigma_activation_value, 0.1) * ref;
        const bool scale_ok = !bal || act >= 1e-6;
        const bool activate = cur <= zthr ? zero_ok && zero_stalled && primal_signal && scale_ok
                                          : normal_ok && bottleneck;
        if (cur <= zthr) prop = activate ? act : 0.0;
        else if (!normal_ok) prop = clipped_update(s, cur, P.sigma_clip_low);
        else if (activate) prop = clipped_update(s, cur, rp / std::fmax(rd, rg));
        else prop = clipped_update(s, cur, P.sigma_clip_low);
        break;
    }
    default: break;
    }
    if (prop > zthr) prop = clampd(prop, smin, smax);
    else prop = 0.0;
    set_sigma_active(s, prop, zthr);
    s.last_sigma_score = score;
    s.last_sigma_rp = rp;
    s.last_sigma_rd = rd;
    s.last_sigma_gap = rg;
    s.sigma_update_count++;
}

const char *restart_env(const char *name, const char *fallback)  // get_env_value, :35-42
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

void choose_restart_candidate(S &s)
{   // :1091-1169: the current iterate as restart point when its score beats the T output's
    // fixed-point error by the margin (mode 2: always)
    const int mode = restart_env_int("FA_CP_RESTART_SELECT", "FA_CP_RESTART_SELECT_LEGACY", 0);
    s.restart_target = 0;
    const double cur = std::isfinite(s.fp) ? s.fp : INFINITY;
    if (mode == 0 || s.total == 0 || !std::isfinite(cur)) return;
    if (s.total < restart_env_int("FA_CP_RESTART_SELECT_MIN_ITERS", "FA_CP_RESTART_SELECT_MIN_ITERS", 0) ||
        s.solve_time < restart_env_double("FA_CP_RESTART_SELECT_MIN_TIME", 0.0))
        return;
    // compute_restart_delta_score(x_T vs the current iterate): the fixed-point metric without the AL term
    for (int i = 0; i < s.n; ++i) s.dx[i] = s.x[i] - s.xT[i];
    for (int i = 0; i < s.m; ++i) s.dy[i] = s.y[i] - s.yT[i];
    spmv(s.At, s.dy.data(), s.dual_prod.data());
    const double pn = nrm2(s.dx), dn = nrm2(s.dy), cross = dot(s.dual_prod, s.dx);
    const double halpern = std::sqrt(std::fmax(pn * pn * s.pw + dn * dn / s.pw + 2.0 * s.step * cross, 0.0));
    if (!(cur > 0.0)) return;
    const double margin = restart_env_double("FA_CP_RESTART_SELECT_MARGIN", 1.0);
    if (mode == 2 || (std::isfinite(halpern) && halpern < margin * cur)) {
        s.restart_target = 1;
        s.fp = halpern;
    }
}

void perform_restart(S &s)
{   // :908-1020: to the T output, or to the current iterate when the candidate selection chose it
    const bool to_current =
        restart_env_int("FA_CP_RESTART_SELECT", "FA_CP_RESTART_SELECT_LEGACY", 0) != 0 && s.restart_target == 1;
    const Vec &rx = to_current ? s.x : s.xT;
    const Vec &ry = to_current ? s.y : s.yT;
    for (int i = 0; i < s.n; ++i) s.dx[i] = rx[i] - s.x0[i];
    for (int i = 0; i < s.m; ++i) s.dy[i] = ry[i] - s.y0[i];
    const double pd = nrm2(s.dx), dd = nrm2(s.dy);
    double normal = INFINITY;
    if (pd > 1e-16 && std::isfinite(pd)) {
        spmv(s.A, s.dx.data(), s.prim_prod.data());
        normal = nrm2(s.prim_prod) / pd;
    }
    const double ratio = s.rel_d / s.rel_p;
    const auto &rp = s.P.restart_params;
    if (pd > 1e-16 && dd > 1e-16 && pd < 1e12 && dd < 1e12 && ratio > 1e-8 && ratio < 1e8) {
        const double err = std::log(dd) - std::log(pd) - std::log(s.pw);
        s.pw_err_sum *= rp.i_smooth;
        s.pw_err_sum += err;
        const double de = err - s.pw_last_err;
        s.pw *= std::exp(rp.k_p * err + rp.k_i * s.pw_err_sum + rp.k_d * de);
        s.pw_last_err = err;
    } else {
        s.pw = s.best_pw;
        s.pw_err_sum = 0.0;
        s.pw_last_err = 0.0;
    }
    const double g = std::fabs(std::log10(s.rel_d / s.rel_p));
    if (g < s.best_gap) {
        s.best_gap = g;
        s.best_pw = s.pw;
    }
    update_al_sigma(s, pd, dd, normal);
    if (to_current) {  // :985-1009: anchor and T output = the current iterate
        s.x0 = s.x;
        s.y0 = s.y;
        s.xT = s.x;
        s.yT = s.y;
    } else {
        s.x0 = s.xT;
        s.y0 = s.yT;
        s.x = s.xT;
        s.y = s.yT;
    }
    s.inner = 0;
    s.last_trial_fp = INFINITY;
}
#endif  // FA_CP sigma rules and restart

/* ---------------- feasibility polishing (fa_cp_core_op.cu:1242-1735; the planted
   solver's QP/PLANTED_FA_CP/src/fa_cp_core_op.cu:1009-1502 is the same code; the
   linearized solver's QP/LIN_CP_AL/src/cp_al_core_op.cu:839-1375 runs its own
   iteration, restart test and fixed-point errors in the same loop) ---------------- */

#if defined(ORACLE_LIN_CP_AL)
void lin_update(S &s);
void lin_halpern(S &s);
bool lin_should_restart(const S &s);
#elif defined(ORACLE_PLANTED_CP_AL)
void pcal_update(S &s);
#endif

struct PolishOut {
    termination_reason_t reason;
    int total;
    double time;
};

void polish_residual(S &ps, const S &ori, bool primal)
{   // compute_primal_feas_polish_residual / compute_dual_feas_polish_residual, utils.cu:1491-1575
    const bool linf = ps.P.optimality_norm == NORM_TYPE_L_INF;
    const double bc = ps.b_scale * ps.c_scale;
    if (primal) {
        spmv(ps.A, ps.xT.data(), ps.prim_prod.data());
        for (int i = 0; i < ps.m; ++i) {
            const double cl = std::fmax(ps.l[i], std::fmin(ps.prim_prod[i], ps.u[i]));
            ps.pres[i] = (ps.prim_prod[i] - cl) * ps.E[i];
        }
        ps.abs_p = (linf ? inf_norm(ps.pres) : nrm2(ps.pres)) / ps.b_scale;
        ps.rel_p = ps.abs_p / (1.0 + ps.b_norm);
        ps.p_obj = dot(ori.c, ps.xT) / bc + ps.obj_const;
    } else {  // no Q x term: the reference's dual polish residual is c - A^T y - slack
        spmv(ps.At, ps.yT.data(), ps.dual_prod.data());
        for (int j = 0; j < ps.n; ++j) ps.dres[j] = (ps.c[j] - ps.dual_prod[j] - ps.dslack[j]) * ps.D[j];
        for (int i = 0; i < ps.m; ++i)
            ps.pslack[i] = std::fmax(ps.yT[i], 0.0) * ori.lfin[i] + std::fmin(ps.yT[i], 0.0) * ori.ufin[i];
        ps.abs_d = (linf ? inf_norm(ps.dres) : nrm2(ps.dres)) / ps.c_scale;
        ps.rel_d = ps.abs_d / (1.0 + ps.c_norm);
        ps.d_obj = (dot(ps.dslack, ori.xT) + sum(ps.pslack)) / bc + ps.obj_const;
    }
}

PolishOut polish_loop(S &ps, const S &ori, bool primal)
{   // primal_feasibility_polish / dual_feasibility_polish, fa_cp_core_op.cu:1312-1404
    const auto t0 = std::chrono::steady_clock::now();
    const int K = ps.P.termination_evaluation_frequency;
    const auto &tc = ps.P.termination_criteria;
    PolishOut out{TERMINATION_REASON_UNSPECIFIED, 0, 0.0};
    bool do_restart = false;
    while (ps.reason == TERMINATION_REASON_UNSPECIFIED) {
        const bool consider = ps.major || ps.total == 0;
        if (consider || ps.total % print_frequency(ps.total) == 0) {
            polish_residual(ps, ori, primal);
            out.time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            // check_feas_polishing_termination_criteria, utils.cu:1360-1384 (the phase's own clock)
            if ((primal ? ps.rel_p : ps.rel_d) <= tc.eps_feas_polish_relative)
                ps.reason = TERMINATION_REASON_FEAS_POLISH_SUCCESS;
            else if (ps.total >= tc.iteration_limit)
                ps.reason = TERMINATION_REASON_ITERATION_LIMIT;
            else if (out.time >= tc.time_sec_limit)
                ps.reason = TERMINATION_REASON_TIME_LIMIT;
        }
        if (consider) {  // also after a stop, as the reference loop runs it
#ifdef ORACLE_LIN_CP_AL
            do_restart = lin_should_restart(ps);  // never true here: fp0 starts at infinity
#else
            do_restart = should_restart(ps);
#endif
            if (do_restart) {  // perform_primal_restart / perform_dual_restart, :1668-1690
                if (primal) {
                    ps.x0 = ps.xT;
                    ps.x = ps.xT;
                } else {
                    ps.y0 = ps.yT;
                    ps.y = ps.yT;
                }
                ps.inner = 0;
                ps.last_trial_fp = INFINITY;
            }
        }
        ps.major = ((ps.total + 1) % K) == 0;
#ifdef ORACLE_CP_AL_TYPES
#ifdef ORACLE_LIN_CP_AL
        lin_update(ps);
#else
        pcal_update(ps);
#endif
        // compute_primal/dual_fixed_point_error, LIN_CP_AL :1331-1375 (PLANTED_CP_AL :1495-1539)
        if (ps.major || do_restart) {
            if (primal) {
                for (int i = 0; i < ps.n; ++i) ps.dx[i] = ps.xR[i] - ps.x[i];
                const double pn = nrm2(ps.dx);
                if (ps.use_al) {
                    spmv(ps.A, ps.dx.data(), ps.prim_prod.data());
                    const double an = nrm2(ps.prim_prod);
                    ps.fp = std::fmax(0.0, ps.sigma * (ps.lambda_est * pn * pn - an * an));
                } else {
                    ps.fp = pn * pn * ps.pw;
                }
            } else {
                for (int i = 0; i < ps.m; ++i) ps.dy[i] = ps.yR[i] - ps.y[i];
                const double dn = nrm2(ps.dy);
                ps.fp = dn * dn / (ps.use_al ? ps.sigma : ps.pw);
            }
            if (do_restart) {
                ps.fp0 = ps.fp;
                do_restart = false;
            }
        }
#ifdef ORACLE_LIN_CP_AL
        lin_halpern(ps);
#else
        halpern(ps);
#endif
#else
        fa_cp_update(ps);
        if (ps.major || do_restart) {  // compute_primal/dual_fixed_point_error, :1714-1735
            if (primal) {
                for (int i = 0; i < ps.n; ++i) ps.dx[i] = ps.xR[i] - ps.x[i];
                const double pn = nrm2(ps.dx);
                ps.fp = pn * pn * ps.pw;
            } else {
                for (int i = 0; i < ps.m; ++i) ps.dy[i] = ps.yR[i] - ps.y[i];
                const double dn = nrm2(ps.dy);
                ps.fp = dn * dn / ps.pw;
            }
            if (do_restart) {
                ps.fp0 = ps.fp;
                do_restart = false;
            }
        }
        halpern(ps);
#endif
        ps.inner++;
        ps.total++;
    }
    out.reason = ps.reason;
    out.total = ps.total;
    return out;
}

void reset_phase(S &ps, double pw0)
{   // scalar resets of initialize_primal/dual_feas_polish_state, :1474-1494, 1626-1646
    ps.pw_err_sum = ps.pw_last_err = 0.0;
    ps.pw = ps.best_pw = pw0;
    ps.fp = ps.fp0 = ps.last_trial_fp = INFINITY;
    ps.major = false;
    ps.total = ps.inner = 0;
    ps.reason = TERMINATION_REASON_UNSPECIFIED;
    ps.best_gap = INFINITY;
}

void feasibility_polish(S &s, double &polish_time, int &polish_iters)
{   // feasibility_polish, fa_cp_core_op.cu:1242-1310
    const auto t0 = std::chrono::steady_clock::now();
    const auto &tc = s.P.termination_criteria;
    if (s.rel_p < tc.eps_feas_polish_relative && s.rel_d < tc.eps_feas_polish_relative) {
        printf("Skipping feasibility polishing as the solution is already sufficiently feasible.\n");
        fflush(stdout);
        return;
    }
    const double pw0 = s.P.bound_objective_rescaling ? 1.0 : (s.c_norm + 1.0) / (s.b_norm + 1.0);
    auto zero = [](Vec &v) { std::fill(v.begin(), v.end(), 0.0); };

    // primal phase: c = 0, constant 0, dual side zero, primal side from the main state
    S ps = s;
    zero(ps.c);
    ps.obj_const = 0.0;
    for (Vec *v : {&ps.y0, &ps.y, &ps.yT, &ps.yR, &ps.al_u, &ps.pslack, &ps.pres, &ps.dy, &ps.dual_prod, &ps.dslack,
                   &ps.dres, &ps.dx})
        zero(*v);
    reset_phase(ps, pw0);
    ps.rel_d = ps.abs_d = ps.rel_gap = ps.gap = 0.0;
    const PolishOut pr = polish_loop(ps, s, true);
    s.inner_total = ps.inner_total;  // the inner solver is shared with the phases
    if (pr.reason == TERMINATION_REASON_FEAS_POLISH_SUCCESS) {
        s.xT = ps.xT;
        s.abs_p = ps.abs_p;
        s.rel_p = ps.rel_p;
        s.p_obj = ps.p_obj;
    }
    polish_iters += ps.total - 1;

    // dual phase: finite bounds 0, primal side zero, dual side from the main state
    S ds = s;
    for (Vec *v : {&ds.l, &ds.u, &ds.lb, &ds.ub})
        for (double &b : *v)
            if (std::isfinite(b)) b = 0.0;
    zero(ds.lfin);
    zero(ds.ufin);
    for (Vec *v : {&ds.x0, &ds.x, &ds.xT, &ds.xR, &ds.dx, &ds.dres, &ds.prim_prod, &ds.al_u, &ds.pslack, &ds.pres,
                   &ds.dy})
        zero(*v);
    reset_phase(ds, pw0);
    ds.rel_p = ds.abs_p = ds.rel_gap = ds.gap = 0.0;
    const PolishOut du = polish_loop(ds, s, false);
    s.inner_total = ds.inner_total;
    if (du.reason == TERMINATION_REASON_FEAS_POLISH_SUCCESS) {
        s.yT = ds.yT;
        s.abs_d = ds.abs_d;
        s.rel_d = ds.rel_d;
        s.d_obj = ds.d_obj;
    }
    polish_iters += ds.total - 1;
    s.gap = std::fabs(s.p_obj - s.d_obj);
    s.rel_gap = s.gap / (1.0 + std::fabs(s.p_obj) + std::fabs(s.d_obj));
    printf("Feasibility Polishing Summary\n");
    printf("  Primal Status        : %d\n  Primal Iterations    : %d\n", (int)pr.reason, pr.total);
    printf("  Dual Status          : %d\n  Dual Iterations      : %d\n", (int)du.reason, du.total);
    fflush(stdout);
    polish_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

#if defined(ORACLE_PLANTED_FA_CP)

/* ---------------- planted FA_CP restart (QP/PLANTED_FA_CP/src/fa_cp_core_op.cu) ---------------- */

bool planted_apply_sigma(S &s, double nsig)
{   // apply_new_al_sigma, :626-658 (the trial bookkeeping it clears is never read)
    if (std::fabs(nsig - s.al_sigma) <= 1e-12 * std::fmax(1.0, std::fabs(s.al_sigma))) return false;
    s.al_sigma = nsig;
    s.last_sigma_update_iter = s.total;
    if (s.has_inner) s.inner_tol = s.inner_tol0;  // reset_inner_solver_tolerance, :610-614
    return true;
}

void planted_sigma_on_restart(S &s)
{   // maybe_update_al_sigma_on_restart, :670-747
    const auto &P = s.P;
    if (!P.adaptive_al_sigma || !s.use_al) return;
    if (s.last_sigma_update_iter == s.total) return;
    const double rp = std::fabs(s.rel_p), rd = std::fabs(s.rel_d), rg = std::fabs(s.rel_gap);
    if (!std::isfinite(rp) || !std::isfinite(rd) || !std::isfinite(rg)) return;
    const double dg = std::fmax(rd, rg);
    const double pref = std::fmax(rp, 1e-12), dgref = std::fmax(dg, 1e-12);
    const double ratio = rp / dgref;
    const bool up = rp > P.al_sigma_primal_trigger * dgref && P.adaptive_al_sigma_mode != ADAPTIVE_AL_SIGMA_MODE_DOWN_ONLY;
    const bool down = dg > P.al_sigma_gap_trigger * pref && P.adaptive_al_sigma_mode != ADAPTIVE_AL_SIGMA_MODE_UP_ONLY;
    double ns = s.al_sigma;
    switch (P.adaptive_al_sigma_rule) {
    case ADAPTIVE_AL_SIGMA_RULE_TRIGGER:
        if (up) ns = std::fmin(P.al_sigma_max, s.al_sigma * P.al_sigma_increase_factor);
        else if (down) ns = std::fmax(P.al_sigma_min, s.al_sigma * P.al_sigma_decrease_factor);
        break;
    case ADAPTIVE_AL_SIGMA_RULE_SQRT:
        if (up) ns = std::fmin(P.al_sigma_max, s.al_sigma * std::fmin(P.al_sigma_increase_factor, std::sqrt(ratio)));
        else if (down) ns = std::fmax(P.al_sigma_min, s.al_sigma * std::fmax(P.al_sigma_decrease_factor, std::sqrt(ratio)));
        break;
    case ADAPTIVE_AL_SIGMA_RULE_BANGBANG:
        if (up) ns = P.al_sigma_max;
        else if (down) ns = P.al_sigma_min;
        break;
    default: break;
    }
    planted_apply_sigma(s, ns);
}

void perform_restart(S &s)
{   // perform_restart, :811-871: PID primal weight as FA_CP, then the AL sigma rule
    for (int i = 0; i < s.n; ++i) s.dx[i] = s.xT[i] - s.x0[i];
    for (int i = 0; i < s.m; ++i) s.dy[i] = s.yT[i] - s.y0[i];
    const double pd = nrm2(s.dx), dd = nrm2(s.dy);
    const double ratio = s.rel_d / s.rel_p;
    const auto &rp = s.P.restart_params;
    if (pd > 1e-16 && dd > 1e-16 && pd < 1e12 && dd < 1e12 && ratio > 1e-8 && ratio < 1e8) {
        const double err = std::log(dd) - std::log(pd) - std::log(s.pw);
        s.pw_err_sum *= rp.i_smooth;
        s.pw_err_sum += err;
        const double de = err - s.pw_last_err;
        s.pw *= std::exp(rp.k_p * err + rp.k_i * s.pw_err_sum + rp.k_d * de);
        s.pw_last_err = err;
    } else {
        s.pw = s.best_pw;
        s.pw_err_sum = 0.0;
        s.pw_last_err = 0.0;
    }
    const double g = std::fabs(std::log10(s.rel_d / s.rel_p));
    if (g < s.best_gap) {
        s.best_gap = g;
        s.best_pw = s.pw;
    }
    planted_sigma_on_restart(s);
    s.x0 = s.xT;
    s.y0 = s.yT;
    s.x = s.xT;
    s.y = s.yT;
    s.inner = 0;
    s.last_trial_fp = INFINITY;
}

#elif defined(ORACLE_LIN_CP_AL)

/* ---------------- linearized CP-AL (QP/LIN_CP_AL/src/cp_al_core_op.cu) ---------------- */

double lin_primal_step(const S &s) { return 1.0 / std::fmax(s.sigma * s.lambda_est, 1e-16); }  // :253-255
double lin_dual_step(const S &s) { return std::fmax(s.sigma, 1e-16); }                          // :257-259

void lin_init_steps(S &s)
{   // initialize_step_size_and_primal_weight, :609-691
    const double max_sv = s.A.nnz() > 0 ? max_singular_value(s) : 0.0;
    const double a_sq = max_sv * max_sv;
    const double half_lf = 0.5 * std::fmax(s.q_norm, 0.0);
    s.pw = s.P.bound_objective_rescaling ? 1.0 : (s.c_norm + 1.0) / (s.b_norm + 1.0);
    s.best_pw = s.pw;
    if (s.use_al) {
        if (a_sq <= 1e-16 && half_lf <= 1e-16) {
            s.lambda_est = 1.0;
            s.sigma_base = 1.0;
        } else {  // t = sqrt(lambda) solves t^2 - (L_f / 2) t - 2 ||A||^2 = 0 (with 8.0004 for 8)
            const double t = std::fmax(0.5 * (half_lf + std::sqrt(half_lf * half_lf + 8.0004 * a_sq)), 1e-16);
            s.lambda_est = t * t;
            s.sigma_base = 1.0 / t;
        }
        s.sigma_min = s.sigma_base;
        s.sigma_max = 100.0 * s.sigma_base;
        s.sigma = s.P.sigma_init > 0.0 ? s.P.sigma_init : 6.0 * s.sigma_base;
        s.sigma = std::fmin(std::fmax(s.sigma, s.sigma_min), s.sigma_max);
        s.step = lin_primal_step(s);
    } else {
        const double wgt = std::fmax(s.pw, 1e-16);
        if (a_sq <= 1e-16 && half_lf <= 1e-16) {
            s.step = 1.0;
        } else if (a_sq <= 1e-16) {
            s.step = 0.99 * wgt / std::fmax(half_lf, 1e-16);
        } else {
            const double lc = half_lf / wgt;
            s.step = 0.99 * (-lc + std::sqrt(lc * lc + 4.0 * a_sq)) / (2.0 * a_sq);
        }
    }
}

void lin_update(S &s)
{   // cp_al_update, :496-575: the T-outputs (xT, yT) and reflections every iteration
    double tau, dstep, al_sigma = 0.0;
    if (s.use_al) {
        tau = lin_primal_step(s);
        dstep = lin_dual_step(s);
        al_sigma = s.sigma;
        s.step = tau;
    } else {
        tau = s.step / s.pw;
        dstep = s.step * s.pw;
    }
    if (s.nonconvexity < 0) {
        tau = std::fmax(tau, -1.01 * std::fmin(0.0, s.nonconvexity));
        tau /= 2;
    }
    if (s.use_al) {  // compute_al_effective_dual_product :275-302, kernel :149-158 (rhs = l)
        spmv(s.A, s.x.data(), s.prim_prod.data());
        for (int i = 0; i < s.m; ++i) s.al_u[i] = s.y[i] - al_sigma * (s.prim_prod[i] - s.l[i]);
        spmv(s.At, s.al_u.data(), s.dual_prod.data());
    } else {
        spmv(s.At, s.y.data(), s.dual_prod.data());
    }
    if (s.qtype == FA_CP_NON_Q) {  // lp_primal_update :304-312, kernel :73-87
        for (int j = 0; j < s.n; ++j) {
            const double xj = s.x[j];
            const double t = xj - tau * (s.c[j] - s.dual_prod[j]);
            const double p = std::fmax(s.lb[j], std::fmin(t, s.ub[j]));
            s.xR[j] = 2.0 * p - xj;
            s.xT[j] = p;
            s.dslack[j] = (p - t) / tau;
        }
    } else {  // primal_gradient_update :484-494, kernel :353-369 (one projected gradient step)
        obj_product(s, s.x);
        for (int j = 0; j < s.n; ++j) {
            const double xj = s.x[j];
            const double g = s.qx[j] + s.c[j] - s.dual_prod[j];
            const double p = std::fmax(s.lb[j], std::fmin(xj - tau * g, s.ub[j]));
            s.xT[j] = p;
            s.xR[j] = 2.0 * p - xj;
        }
    }
    spmv(s.A, s.xR.data(), s.prim_prod.data());
    for (int i = 0; i < s.m; ++i) {  // kernels :135-147 and :173-185: rhs = l
        const double nd = s.y[i] + dstep * (s.l[i] - s.prim_prod[i]);
        s.yT[i] = nd;
        s.yR[i] = 2.0 * nd - s.y[i];
    }
}

void lin_halpern(S &s)
{   // halpern_update :577-585, kernel :187-206: relaxed (1 + rho) T - rho z
    const double w = (double)(s.inner + 1) / (s.inner + 2);
    const double rho = s.P.reflection_coefficient;
    for (int i = 0; i < s.n; ++i) {
        const double rl = (1.0 + rho) * s.xT[i] - rho * s.x[i];
        s.x[i] = w * rl + (1.0 - w) * s.x0[i];
    }
    for (int i = 0; i < s.m; ++i) {
        const double rl = (1.0 + rho) * s.yT[i] - rho * s.y[i];
        s.y[i] = w * rl + (1.0 - w) * s.y0[i];
    }
}

void lin_fixed_point_error(S &s)
{   // compute_fixed_point_error, :693-770 (delta = T-output - current)
    for (int i = 0; i < s.n; ++i) s.dx[i] = s.xT[i] - s.x[i];
    for (int i = 0; i < s.m; ++i) s.dy[i] = s.yT[i] - s.y[i];
    const double dn = nrm2(s.dy), pn = nrm2(s.dx);
    double fp2;
    if (s.use_al) {
        spmv(s.A, s.dx.data(), s.prim_prod.data());
        const double an = nrm2(s.prim_prod);
        const double inter = dot(s.prim_prod, s.dy);
        fp2 = s.sigma * (s.lambda_est * pn * pn - an * an) + 2.0 * inter + dn * dn / s.sigma;
    } else {
        spmv(s.At, s.dy.data(), s.dual_prod.data());
        const double cross = dot(s.dual_prod, s.dx);
        const double tau = s.step / s.pw, rho = s.step * s.pw;
        fp2 = pn * pn / std::fmax(tau, 1e-16) + dn * dn / std::fmax(rho, 1e-16) + 2.0 * cross;
    }
    s.fp = std::sqrt(std::fmax(0.0, fp2));
}

bool lin_should_restart(const S &s)
{   // should_do_adaptive_restart, utils.cu:451-463
    if (!std::isfinite(s.fp0) || !std::isfinite(s.fp)) return false;
    if (s.total == s.inner) return s.inner >= s.P.termination_evaluation_frequency;
    return s.fp <= s.P.restart_params.sufficient_reduction_for_restart * s.fp0;
}

void lin_restart(S &s)
{   // perform_restart, :595-607: the anchors move to the current iterates
    s.x0 = s.x;
    s.y0 = s.y;
    s.inner = 0;
    s.last_trial_fp = INFINITY;
}

#elif defined(ORACLE_PLANTED_CP_AL)

/* ---------------- planted-KKT CP-AL (QP/PLANTED_CP_AL/src/cp_al_core_op.cu) ---------------- */

double pcal_tau(const S &s)
{   // cpal_effective_primal_step_size, :255-265: 1 / (sigma lambda), nonconvexity-adjusted
    double t = 1.0 / std::fmax(s.sigma * s.lambda_est, 1e-16);
    if (s.nonconvexity < 0) {
        t = std::fmax(t, -1.01 * std::fmin(0.0, s.nonconvexity));
        t /= 2;
    }
    return t;
}

void pcal_init_steps(S &s)
{   // initialize_step_size_and_primal_weight, :805-858
    if (s.use_al) {
        if (s.A.nnz() == 0) {
            s.lambda_est = 1.0;
        } else {
            const double sv = max_singular_value(s);
            s.lambda_est = 2.001 * sv * sv;
        }
        s.sigma_base = 1.0 / std::sqrt(std::fmax(s.lambda_est, 1e-16));
        s.sigma_min = 1e-4 * s.sigma_base;
        s.sigma_max = 100.0 * s.sigma_base;
        s.sigma = s.P.sigma_init > 0.0 ? s.P.sigma_init : 0.5 * s.sigma_base;
        s.sigma = std::fmin(std::fmax(s.sigma, s.sigma_min), s.sigma_max);
        s.best_sigma = s.sigma;
        s.best_sigma_gap = INFINITY;
        s.best_term_metric = INFINITY;
        s.step = pcal_tau(s);
    } else {
        s.step = s.A.nnz() == 0 ? 1.0 : 0.998 / max_singular_value(s);
    }
    s.pw = s.P.bound_objective_rescaling ? 1.0 : (s.c_norm + 1.0) / (s.b_norm + 1.0);
    s.best_pw = s.pw;
}

void pcal_update(S &s)
{   // cp_al_update, :531-635: the T-outputs (and the LP dual slack) only in major-kernel
    // iterations; equality rows (right-hand side l) in both branches
    double tau, dstep;
    if (s.use_al) {
        tau = pcal_tau(s);
        dstep = std::fmax(s.sigma, 1e-16);
        s.step = tau;
    } else {
        tau = s.step / s.pw;
        dstep = s.step * s.pw;
        if (s.nonconvexity < 0) {
            tau = std::fmax(tau, -1.01 * std::fmin(0.0, s.nonconvexity));
            tau /= 2;
        }
    }
    const bool mk = major_kernel(s);
    if (s.use_al) {  // compute_al_effective_dual_product, :285-312 (kernel :149-158)
        spmv(s.A, s.x.data(), s.prim_prod.data());
        for (int i = 0; i < s.m; ++i) s.al_u[i] = s.y[i] - s.sigma * (s.prim_prod[i] - s.l[i]);
        spmv(s.At, s.al_u.data(), s.dual_prod.data());
    } else {
        spmv(s.At, s.y.data(), s.dual_prod.data());
    }
    switch (s.qtype) {
    case FA_CP_NON_Q:  // lp_primal_update, :314-333 (kernels :59-87)
        for (int j = 0; j < s.n; ++j) {
            const double xj = s.x[j];
            const double t = xj - tau * (s.c[j] - s.dual_prod[j]);
            const double p = std::fmax(s.lb[j], std::fmin(t, s.ub[j]));
            s.xR[j] = 2.0 * p - xj;
            if (mk) {
                s.xT[j] = p;
                s.dslack[j] = (p - t) / tau;
            }
        }
        break;
    case FA_CP_DIAG_Q:  // diag_q_primal_update, :335-356 (kernels :89-120)
        for (int j = 0; j < s.n; ++j) {
            const double xj = s.x[j];
            const double t = (xj - tau * (s.c[j] - s.dual_prod[j])) / (1.0 + tau * s.qdiag[j]);
            const double p = std::fmax(s.lb[j], std::fmin(t, s.ub[j]));
            s.xR[j] = 2.0 * p - xj;
            if (mk) s.xT[j] = p;
        }
        break;
    default: {  // primal_BB_step_size_update, :454-503: FA_CP's BB with the product fixed
        const bool al = s.use_al;
        s.use_al = false;
        primal_bb(s, tau, 0.0);
        s.use_al = al;
        break;
    }
    }
    spmv(s.A, s.xR.data(), s.prim_prod.data());
    for (int i = 0; i < s.m; ++i) {  // kernels :122-147, :160-185
        const double nd = s.y[i] + dstep * (s.l[i] - s.prim_prod[i]);
        if (mk) s.yT[i] = nd;
        s.yR[i] = 2.0 * nd - s.y[i];
    }
}

void pcal_fixed_point_error(S &s)
{   // compute_fixed_point_error, :860-935 (delta = reflected - current)
    for (int i = 0; i < s.n; ++i) s.dx[i] = s.xR[i] - s.x[i];
    for (int i = 0; i < s.m; ++i) s.dy[i] = s.yR[i] - s.y[i];
    const double dn = nrm2(s.dy), pn = nrm2(s.dx);
    if (s.use_al) {
        spmv(s.A, s.dx.data(), s.prim_prod.data());
        const double an = nrm2(s.prim_prod);
        const double inter = dot(s.prim_prod, s.dy);
        const double fp2 = s.sigma * (s.lambda_est * pn * pn - an * an) + 2.0 * inter + dn * dn / s.sigma;
        s.fp = std::sqrt(std::fmax(0.0, fp2));
    } else {
        spmv(s.At, s.dy.data(), s.dual_prod.data());
        const double movement = pn * pn * s.pw + dn * dn / s.pw;
        const double cross = dot(s.dual_prod, s.dx);
        s.fp = std::sqrt(movement + 2 * s.step * cross);  // no clamp at zero (:921)
    }
    if (s.has_inner) {
        const double st = s.use_al ? s.step : s.step / s.pw;
        s.inner_tol = std::fmin(s.inner_tol, std::fmax(0.0005 * pn / st, 1e-9));
    }
}

void pcal_reset_sigma(S &s)
{   // reset_sigma_controller, :271-283
    double sg = std::fmax(s.best_sigma, 1e-16);
    if (s.sigma_min > 0.0) sg = std::fmax(sg, s.sigma_min);
    if (s.sigma_max > 0.0) sg = std::fmin(sg, s.sigma_max);
    s.sigma = sg;
    s.step = pcal_tau(s);
}

void pcal_restart(S &s)
{   // perform_restart, :655-803: the AL branch's sigma controller (log-damped step toward
    // ||dy|| / sqrt(lambda ||dx||^2 - ||A dx||^2), clipped to [0.6, 1.7] per restart, rolled
    // back to the best sigma when the restart metric worsens), else the primal-weight PID;
    // then the anchors and iterates move to the T-outputs
    for (int i = 0; i < s.n; ++i) s.dx[i] = s.xT[i] - s.x0[i];
    for (int i = 0; i < s.m; ++i) s.dy[i] = s.yT[i] - s.y0[i];
    const double pd = nrm2(s.dx), dd = nrm2(s.dy);
    if (s.use_al) {
        spmv(s.A, s.dx.data(), s.prim_prod.data());
        const double ax = nrm2(s.prim_prod);
        double cur = s.fp;
        const double dxm = s.lambda_est * pd * pd - ax * ax;
        const double ratio = s.rel_d / std::fmax(s.rel_p, 1e-300);
        const double old = std::fmax(s.sigma, 1e-16);
        bool rolled = false;
        if (std::isfinite(cur) && cur < s.best_term_metric) {
            s.best_term_metric = cur;
            s.best_sigma = old;
            s.best_sigma_gap = std::fabs(std::log10(std::fmax(s.rel_d, 1e-300) / std::fmax(s.rel_p, 1e-300)));
        }
        if (!std::isfinite(cur)) {
            pcal_reset_sigma(s);
            rolled = true;
            cur = INFINITY;
        } else if (std::isfinite(s.best_term_metric) && cur > 1.05 * s.best_term_metric) {
            pcal_reset_sigma(s);
            rolled = true;
        } else {
            bool valid = false;
            double est = old;
            if (dd > 1e-16 && dd < 1e12 && dxm > 1e-16 && dxm < 1e24 && std::isfinite(s.sigma) && ratio > 1e-8 &&
                ratio < 1e8) {
                est = dd / std::sqrt(dxm);
                valid = std::isfinite(est) && est > 1e-16;
            }
            if (valid) {
                const double target = old * std::exp(0.5 * std::log(est / old));
                if (std::isfinite(target) && target > 1e-16) {
                    const double r = std::fmax(0.60, std::fmin(target / old, 1.70));
                    s.sigma = old * r;
                }
            }
        }
        if (!std::isfinite(s.sigma) || s.sigma <= 1e-16 || s.sigma >= 1e16) {
            pcal_reset_sigma(s);
            rolled = true;
        } else {
            s.sigma = std::fmin(std::fmax(s.sigma, s.sigma_min), s.sigma_max);
            s.step = pcal_tau(s);
        }
        const double g = std::fabs(std::log10(std::fmax(s.rel_d, 1e-300) / std::fmax(s.rel_p, 1e-300)));
        if (g < s.best_sigma_gap) s.best_sigma_gap = g;
        if (rolled) {
            s.sigma = s.best_sigma;
            s.step = pcal_tau(s);
        }
    } else {
        const double ratio = s.rel_d / s.rel_p;
        const auto &rp = s.P.restart_params;
        if (pd > 1e-16 && dd > 1e-16 && pd < 1e12 && dd < 1e12 && ratio > 1e-8 && ratio < 1e8) {
            const double err = std::log(dd) - std::log(pd) - std::log(s.pw);
            s.pw_err_sum *= rp.i_smooth;
            s.pw_err_sum += err;
            const double de = err - s.pw_last_err;
            s.pw *= std::exp(rp.k_p * err + rp.k_i * s.pw_err_sum + rp.k_d * de);
            s.pw_last_err = err;
        } else {
            s.pw = s.best_pw;
            s.pw_err_sum = 0.0;
            s.pw_last_err = 0.0;
        }
        const double g = std::fabs(std::log10(s.rel_d / s.rel_p));
        if (g < s.best_gap) {
            s.best_gap = g;
            s.best_pw = s.pw;
        }
    }
    s.x0 = s.xT;
    s.x = s.xT;
    s.y0 = s.yT;
    s.y = s.yT;
    s.inner = 0;
    s.last_trial_fp = INFINITY;
}

#endif  // ORACLE_LIN_CP_AL / ORACLE_PLANTED_CP_AL

void init_state(S &s, const qp_problem_t *w)
{   // solver_state.cu:494-877 after rescale_problem (preconditioner.c)
    s.n = w->num_variables;
    s.m = w->num_constraints;
    s.r = w->num_rank_lowrank_obj;
    s.obj_const = w->objective_constant;
    s.A = from_component(w->constraint_matrix, s.m, s.n, w->constraint_matrix_num_nonzeros);
    s.Q = from_component(w->objective_sparse_matrix, s.n, s.n, w->objective_sparse_matrix_num_nonzeros);
    s.R = from_component(w->objective_lowrank_matrix, s.r, s.n, w->objective_lowrank_matrix_num_nonzeros);
    s.c.assign(w->objective_vector, w->objective_vector + s.n);
    s.lb.assign(w->variable_lower_bound, w->variable_lower_bound + s.n);
    s.ub.assign(w->variable_upper_bound, w->variable_upper_bound + s.n);
    s.l.assign(w->constraint_lower_bound, w->constraint_lower_bound + s.m);
    s.u.assign(w->constraint_upper_bound, w->constraint_upper_bound + s.m);
    const auto t0 = std::chrono::steady_clock::now();
    rescale(s);
    s.rescale_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    s.At = transpose(s.A);
    s.Rt = transpose(s.R);
    s.al_available = s.P.use_al_qp;
    s.use_al = s.P.use_al_qp;
#if defined(ORACLE_PLANTED_FA_CP)
    // planted solver_state.cu:498-504: the AL branch only with a nonzero sigma
    s.use_al = s.P.use_al_qp && s.P.al_sigma != 0.0;
    s.al_sigma = s.use_al ? s.P.al_sigma : 0.0;
#elif !defined(ORACLE_CP_AL_TYPES)
    s.al_sigma = s.P.al_sigma;
    s.best_al_sigma = s.P.al_sigma;
#endif
    for (Vec *v : {&s.x0, &s.x, &s.xT, &s.xR, &s.dual_prod, &s.dslack, &s.dres, &s.dx, &s.qx, &s.grad, &s.dir, &s.gbuf})
        v->assign(s.n, 0.0);
    for (Vec *v : {&s.y0, &s.y, &s.yT, &s.yR, &s.prim_prod, &s.al_u, &s.pslack, &s.pres, &s.dy}) v->assign(s.m, 0.0);
    s.rx.assign(s.r, 0.0);
    if (w->primal_start)
        for (int i = 0; i < s.n; ++i) s.x0[i] = s.x[i] = s.xT[i] = w->primal_start[i] * s.D[i] * s.b_scale;
    if (w->dual_start)
        for (int i = 0; i < s.m; ++i) s.y0[i] = s.y[i] = s.yT[i] = w->dual_start[i] * s.E[i] * s.c_scale;
    s.lfin.resize(s.m);
    s.ufin.resize(s.m);
    for (int i = 0; i < s.m; ++i) {
        s.lfin[i] = std::isfinite(s.l[i]) ? s.l[i] : 0.0;
        s.ufin[i] = std::isfinite(s.u[i]) ? s.u[i] : 0.0;
    }
    const bool linf = s.P.optimality_norm == NORM_TYPE_L_INF;
    double sq = 0.0, mx = 0.0;
    for (int i = 0; i < s.n; ++i) {
        if (linf) mx = std::max(mx, std::fabs(w->objective_vector[i]));
        else sq += w->objective_vector[i] * w->objective_vector[i];
    }
    s.c_norm = linf ? mx : std::sqrt(sq);
    sq = mx = 0.0;
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
    // quadratic term (solver_state.cu:220-303, 462-492)
    s.qtype = detect(s);
    if (s.qtype == FA_CP_DIAG_Q) {
        s.qdiag.assign(s.n, 0.0);
        for (int i = 0; i < s.n; ++i)
            for (int k = s.Q.rp[i]; k < s.Q.rp[i + 1]; ++k)
                if (s.Q.ci[k] < s.n) s.qdiag[s.Q.ci[k]] = s.Q.v[k] + 1e-12;
    }
    if (s.qtype == FA_CP_SPARSE_Q) {
        s.q_norm = q_max_eig(s);
        s.nonconvexity = q_min_eig(s, s.q_norm);
    } else if (s.qtype == FA_CP_DIAG_Q) {
        double mxe = 0.0, mne = 0.0;
        for (double d : s.qdiag) {
            mxe = std::fmax(mxe, std::fabs(d));
            mne = std::fmin(mne, d);
        }
        s.q_norm = mxe;
        s.nonconvexity = mne;
    }
#if defined(ORACLE_PLANTED_FA_CP)
    // planted FA_CP solver_state.cu:305-345: the inner loop for non-diagonal Q, and for diagonal Q in the AL branch
    const bool inner_loop = s.qtype != FA_CP_NON_Q && (s.qtype != FA_CP_DIAG_Q || s.use_al);
    s.has_inner = inner_loop;
    if (inner_loop) {
        s.inner_limit = 1000;
        s.inner_tol = s.inner_tol0 = 1e-3;
    }
#elif defined(ORACLE_CP_AL_TYPES)
    // LIN_CP_AL / PLANTED_CP_AL solver_state.cu:305-340: an inner loop only for the
    // non-diagonal Q types (the linearized update never runs it; the planted one's BB does)
    const bool simple = s.qtype == FA_CP_NON_Q || s.qtype == FA_CP_DIAG_Q;
    s.has_inner = !simple;
    if (!simple) {
        s.inner_limit = 1000;
        s.inner_tol = 1e-3;
    }
#else
    if (s.al_available && s.P.al_sigma <= s.P.sigma_zero_threshold) {
        s.use_al = false;
        s.al_sigma = 0.0;
    }
    // initialize_inner_solver, solver_state.cu:305-360
    s.has_inner = s.use_al || s.al_available;
    const bool simple = s.qtype == FA_CP_NON_Q || s.qtype == FA_CP_DIAG_Q;
    if (!simple) s.has_inner = true;
    if (!simple || s.al_available) {
        s.inner_limit = 1000;
        s.inner_tol = 1e-3;
    }
#endif
    s.is_lp = s.qtype == FA_CP_NON_Q;
}

}  // namespace

/* ---------------- public C API ---------------- */

#if defined(ORACLE_PLANTED_FA_CP)
ORACLE_API void planted_fa_cp_oracle_reset_rng(void)
#elif defined(ORACLE_LIN_CP_AL)
ORACLE_API void lin_cp_al_oracle_reset_rng(void)
#elif defined(ORACLE_PLANTED_CP_AL)
ORACLE_API void planted_cp_al_oracle_reset_rng(void)
#else
// Test hook, no counterpart in the reference: the Q eigenvalue estimates of
// initialize_quadratic_term_information on an n x n CSR matrix -- the power
// iteration of estimate_maximum_eigenvalue and the shifted one of
// estimate_minimum_eigenvalue (utils.cu:172-315) -- with the given iteration
// limit and tolerance, rand() reseeded with `seed` first.  out = {q_norm,
// minimum eigenvalue estimate}.
ORACLE_API void fa_cp_oracle_q_eigen(int n, const int *rp, const int *ci, const double *v, int max_iter, double tol,
                                     unsigned seed, double *out)
{
    S s;
    s.n = n;
    s.P.sv_max_iter = max_iter;
    s.P.sv_tol = tol;
    s.Q.rows = s.Q.cols = n;
    s.Q.rp.assign(rp, rp + n + 1);
    s.Q.ci.assign(ci, ci + rp[n]);
    s.Q.v.assign(v, v + rp[n]);
    srand(seed);
    out[0] = q_max_eig(s);
    out[1] = q_min_eig(s, out[0]);
}

ORACLE_API void fa_cp_oracle_reset_rng(void)
#endif
{
    g_gen.seed(1);
    g_dist.reset();
    srand(1);
}

#ifdef ORACLE_LIN_CP_AL
ORACLE_API void set_default_parameters(cp_al_parameters_t *p)  // LIN_CP_AL utils.cu:466-497
{
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
#elif defined(ORACLE_PLANTED_CP_AL)
ORACLE_API void set_default_parameters(cp_al_parameters_t *p)  // PLANTED_CP_AL utils.cu:481-513
{
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
#elif defined(ORACLE_PLANTED_FA_CP)
ORACLE_API void set_default_parameters(fa_cp_parameters_t *p)  // PLANTED_FA_CP utils.cu:471-512
{
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
ORACLE_API void set_default_parameters(fa_cp_parameters_t *p)  // utils.cu:494-537
{
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

static bool to_csr(const matrix_desc_t *d, int rows_expected, CsrComponent *out, int *nnz_out)
{   // fa_cp.c:60-270 with utils.cu:1135-1358 (zero_tolerance rules)
    const int m = d->m, n = d->n;
    std::vector<int> ri, cj;
    Vec vv;
    switch (d->fmt) {
    case matrix_dense: {
        const double tol = d->zero_tolerance > 0 ? d->zero_tolerance : 1e-12;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < n; ++j) {
                const double a = d->data.dense.A[(long)i * n + j];
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
        memcpy(out->row_ptr, d->data.csr.row_ptr, (size_t)(m + 1) * sizeof(int));
        if (nnz) {
            memcpy(out->col_ind, d->data.csr.col_ind, (size_t)nnz * sizeof(int));
            memcpy(out->val, d->data.csr.vals, (size_t)nnz * sizeof(double));
        }
        *nnz_out = nnz;
        return true;
    }
    case matrix_csc:
    case matrix_coo: {
        const bool csc = d->fmt == matrix_csc;
        const double tol = d->zero_tolerance > 0 ? d->zero_tolerance : 0.0;
        const int nnz = csc ? d->data.csc.nnz : d->data.coo.nnz;
        const double *vals = csc ? d->data.csc.vals : d->data.coo.vals;
        for (int k = 0, j = 0; k < nnz; ++k) {
            int r, c;
            if (csc) {
                while (j < n && k >= d->data.csc.col_ptr[j + 1]) ++j;
                r = d->data.csc.row_ind[k];
                c = j;
            } else {
                r = d->data.coo.row_ind[k];
                c = d->data.coo.col_ind[k];
            }
            if (r < 0 || r >= m || c < 0 || c >= n) return false;
            if (tol > 0 && std::fabs(vals[k]) <= tol) continue;
            ri.push_back(r);
            cj.push_back(c);
            vv.push_back(vals[k]);
        }
        break;
    }
    default: return false;
    }
    (void)rows_expected;
    const int nnz = (int)vv.size();
    out->row_ptr = (int *)calloc((size_t)m + 1, sizeof(int));
    out->col_ind = (int *)malloc((size_t)std::max(nnz, 1) * sizeof(int));
    out->val = (double *)malloc((size_t)std::max(nnz, 1) * sizeof(double));
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

static void fill_or_copy(double **dst, int n, const double *src, double fill)
{
    *dst = (double *)malloc((size_t)std::max(n, 1) * sizeof(double));
    if (src) memcpy(*dst, src, (size_t)n * sizeof(double));
    else
        for (int i = 0; i < n; ++i) (*dst)[i] = fill;
}

static void free_component(CsrComponent *c)
{
    if (!c) return;
    free(c->row_ptr);
    free(c->col_ind);
    free(c->val);
    free(c);
}

ORACLE_API void qp_problem_free(qp_problem_t *p)
{
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

ORACLE_API qp_problem_t *create_qp_problem(const double *c, const matrix_desc_t *Q, const matrix_desc_t *R,
                                           const matrix_desc_t *A, const double *con_lb, const double *con_ub,
                                           const double *var_lb, const double *var_ub, const double *obj_const)
{   // fa_cp.c:28-288
    int n, m = 0;
    if (A) {
        n = A->n;
        m = A->m;
    } else if (Q) {
        n = Q->n;
    } else if (R) {
        n = R->n;
    } else {
        fprintf(stderr, "[interface] Error: At least one matrix (A, Q, or R) must be provided.\n");
        return nullptr;
    }
    qp_problem_t *p = (qp_problem_t *)calloc(1, sizeof(qp_problem_t));
    p->num_variables = n;
    p->num_constraints = m;
    p->constraint_matrix = (CsrComponent *)calloc(1, sizeof(CsrComponent));
    p->objective_sparse_matrix = (CsrComponent *)calloc(1, sizeof(CsrComponent));
    p->objective_lowrank_matrix = (CsrComponent *)calloc(1, sizeof(CsrComponent));
    auto empty = [](CsrComponent *cc) { cc->row_ptr = (int *)calloc(1, sizeof(int)); };
    bool ok = true;
    if (A) ok = ok && to_csr(A, m, p->constraint_matrix, &p->constraint_matrix_num_nonzeros);
    else empty(p->constraint_matrix);
    if (Q) ok = ok && to_csr(Q, n, p->objective_sparse_matrix, &p->objective_sparse_matrix_num_nonzeros);
    else empty(p->objective_sparse_matrix);
    if (R) {
        p->num_rank_lowrank_obj = R->m;
        ok = ok && to_csr(R, R->m, p->objective_lowrank_matrix, &p->objective_lowrank_matrix_num_nonzeros);
    } else {
        empty(p->objective_lowrank_matrix);
    }
    if (!ok) {
        fprintf(stderr, "[interface] matrix conversion failed.\n");
        qp_problem_free(p);
        return nullptr;
    }
    p->objective_constant = obj_const ? *obj_const : 0.0;
    fill_or_copy(&p->objective_vector, n, c, 0.0);
    fill_or_copy(&p->variable_lower_bound, n, var_lb, -INFINITY);
    fill_or_copy(&p->variable_upper_bound, n, var_ub, INFINITY);
    fill_or_copy(&p->constraint_lower_bound, m, con_lb, -INFINITY);
    fill_or_copy(&p->constraint_upper_bound, m, con_ub, INFINITY);
    return p;
}

ORACLE_API void set_start_values(qp_problem_t *p, const double *primal, const double *dual)
{
    if (!p) return;
    free(p->primal_start);
    free(p->dual_start);
    p->primal_start = p->dual_start = nullptr;
    if (primal) {
        p->primal_start = (double *)malloc((size_t)std::max(p->num_variables, 1) * sizeof(double));
        memcpy(p->primal_start, primal, (size_t)p->num_variables * sizeof(double));
    }
    if (dual) {
        p->dual_start = (double *)malloc((size_t)std::max(p->num_constraints, 1) * sizeof(double));
        memcpy(p->dual_start, dual, (size_t)p->num_constraints * sizeof(double));
    }
}

#ifdef ORACLE_CP_AL_TYPES
ORACLE_API void cp_al_result_free(cp_al_result_t *r)
#else
ORACLE_API void fa_cp_result_free(fa_cp_result_t *r)
#endif
{
    if (!r) return;
    free(r->primal_solution);
    free(r->dual_solution);
    free(r->reduced_cost);
    free(r);
}

static result_t *result_from_state(S &s, const qp_problem_t *prob, double polish_time, int polish_iters);

#ifdef ORACLE_LIN_CP_AL
ORACLE_API cp_al_result_t *optimize(const cp_al_parameters_t *params, const qp_problem_t *prob)
{   // LIN_CP_AL solver.cu:34-114
    if (!params || !prob) return nullptr;
    S s;
    s.P = *params;
    init_state(s, prob);
    if (s.nonconvexity < 0) s.inner_limit = 1;
    lin_init_steps(s);
    const auto t0 = std::chrono::steady_clock::now();
    const int K = s.P.termination_evaluation_frequency;
    while (s.total < s.P.termination_criteria.iteration_limit) {
        const bool eval = s.total == 0 || s.total % K == 0;
        s.major = eval;
        if (eval || s.total % print_frequency(s.total) == 0) {
            compute_residual(s, s.P.optimality_norm);
            if (eval && s.total < 3 * K) infeasibility_info(s);
            s.solve_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            check_termination(s);
            if (s.reason != TERMINATION_REASON_UNSPECIFIED) break;
        }
        lin_update(s);
        lin_fixed_point_error(s);
        if (!std::isfinite(s.fp0)) s.fp0 = s.fp;
        if (eval && lin_should_restart(s)) {
            lin_restart(s);
            s.fp0 = s.fp;
        }
        lin_halpern(s);
        s.inner++;
        s.total++;
    }
    if (s.reason == TERMINATION_REASON_UNSPECIFIED) {
        s.reason = TERMINATION_REASON_ITERATION_LIMIT;
        compute_residual(s, s.P.optimality_norm);
    }
    double polish_time = 0.0;
    int polish_iters = 0;
    if (s.P.feasibility_polishing && s.reason != TERMINATION_REASON_DUAL_INFEASIBLE &&
        s.reason != TERMINATION_REASON_PRIMAL_INFEASIBLE)  // LIN_CP_AL solver.cu:102-106
        feasibility_polish(s, polish_time, polish_iters);
    return result_from_state(s, prob, polish_time, polish_iters);
}
#else
/* ---------------- trajectory CSV (solver.cu:35-71; the planted variants' :35-109) ---------------- */

bool trajectory_open(const S &s)
{   // initialize_trajectory_csv: the header row; false when the file cannot be written
    const char *path = s.P.trajectory_csv;
    if (!path || !path[0]) return true;
    FILE *f = fopen(path, "w");
    if (!f) {
        perror("Error opening trajectory CSV");
        return false;
    }
#if defined(ORACLE_PLANTED_FA_CP) || defined(ORACLE_PLANTED_CP_AL)
    fprintf(f, "iteration,time_sec,rel_primal,rel_dual,rel_gap,kkt_residual,fixed_point_residual,restart,sigma,"
               "primal_weight,al_sigma,lambda_est,primal_step,dual_step\n");
#else
    fprintf(f, "iteration,time_sec,rel_primal,rel_dual,rel_gap,kkt_residual,fixed_point_residual,restart,al_sigma,"
               "sigma_update_count\n");
#endif
    fclose(f);
    return true;
}

bool trajectory_row(const S &s, bool restart)
{   // append_trajectory_row: one row per checkpoint, before a planned restart is performed
    const char *path = s.P.trajectory_csv;
    if (!path || !path[0]) return true;
    FILE *f = fopen(path, "a");
    if (!f) {
        perror("Error appending trajectory CSV");
        return false;
    }
    const double kkt = std::fmax(s.rel_p, std::fmax(s.rel_d, s.rel_gap));
    const double fpr = s.total == 0 ? NAN : s.fp;
#if defined(ORACLE_PLANTED_FA_CP) || defined(ORACLE_PLANTED_CP_AL)
    double sigma = NAN, al_sigma = NAN, lambda = NAN, pstep = NAN, dstep = NAN;  // compute_trajectory_diagnostics
#ifdef ORACLE_PLANTED_CP_AL
    if (s.use_al) {
        sigma = s.sigma;
        lambda = s.lambda_est;
        pstep = s.step;
        dstep = s.sigma;
    } else if (s.pw > 0.0) {
        pstep = s.step / s.pw;
        dstep = s.step * s.pw;
    }
#else
    al_sigma = s.use_al ? s.al_sigma : NAN;
    if (s.pw > 0.0) {
        pstep = s.step / s.pw;
        dstep = s.step * s.pw;
    }
#endif
    fprintf(f, "%d,%.17e,%.17e,%.17e,%.17e,%.17e,%.17e,%d,%.17e,%.17e,%.17e,%.17e,%.17e,%.17e\n", s.total,
            s.solve_time, s.rel_p, s.rel_d, s.rel_gap, kkt, fpr, restart ? 1 : 0, sigma, s.pw, al_sigma, lambda, pstep,
            dstep);
#else
    fprintf(f, "%d,%.17e,%.17e,%.17e,%.17e,%.17e,%.17e,%d,%.17e,%d\n", s.total, s.solve_time, s.rel_p, s.rel_d,
            s.rel_gap, kkt, fpr, restart ? 1 : 0, s.al_sigma, s.sigma_update_count);
#endif
    fclose(f);
    return true;
}

ORACLE_API result_t *optimize(const params_t *params, const qp_problem_t *prob)
{   // solver.cu:74-168 (PLANTED_CP_AL solver.cu:111-204: the same loop with its own update,
    // fixed-point error, restart and step set-up)
    if (!params || !prob) return nullptr;
    S s;
    s.P = *params;
    init_state(s, prob);
    if (s.nonconvexity < 0) s.inner_limit = 1;
#ifdef ORACLE_PLANTED_CP_AL
    pcal_init_steps(s);
#else
    // initialize_step_size_and_primal_weight, :1022-1040
    if (s.A.nnz() == 0) s.step = 1.0;
    else s.step = 0.998 / max_singular_value(s);
    s.pw = s.P.bound_objective_rescaling ? 1.0 : (s.c_norm + 1.0) / (s.b_norm + 1.0);
    s.best_pw = s.pw;
#endif
    if (!trajectory_open(s)) return nullptr;  // the reference exits

    const auto t0 = std::chrono::steady_clock::now();
    const int K = s.P.termination_evaluation_frequency;
    bool do_restart = false;
    while (s.total < s.P.termination_criteria.iteration_limit) {
        const bool consider = s.major || s.total == 0;
        bool planned = false;
        if (consider || s.total % print_frequency(s.total) == 0) {
            compute_residual(s, s.P.optimality_norm);
            if (s.major && s.total < 3 * K) infeasibility_info(s);
            s.solve_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            check_termination(s);
            if (consider && s.reason == TERMINATION_REASON_UNSPECIFIED) {
#if !defined(ORACLE_CP_AL_TYPES) && !defined(ORACLE_PLANTED_FA_CP)
                choose_restart_candidate(s);  // solver.cu:112
#endif
                planned = should_restart(s);
            }
            if (!trajectory_row(s, planned)) return nullptr;
            if (s.reason != TERMINATION_REASON_UNSPECIFIED) break;
        }
        if (consider) {
            do_restart = planned;
#ifdef ORACLE_PLANTED_CP_AL
            if (do_restart) pcal_restart(s);
#else
            if (do_restart) perform_restart(s);
#endif
        }
        s.major = ((s.total + 1) % K) == 0;
#ifdef ORACLE_PLANTED_CP_AL
        pcal_update(s);
#else
        fa_cp_update(s);
#endif
        if (s.major || do_restart) {
#ifdef ORACLE_PLANTED_CP_AL
            pcal_fixed_point_error(s);
#else
            fixed_point_error(s);
#endif
            if (do_restart) {
                s.fp0 = s.fp;
                do_restart = false;
            }
        }
        halpern(s);
        s.inner++;
        s.total++;
    }
    if (s.reason == TERMINATION_REASON_UNSPECIFIED) {
        s.reason = TERMINATION_REASON_ITERATION_LIMIT;
        compute_residual(s, s.P.optimality_norm);
        if (!trajectory_row(s, false)) return nullptr;
    }
    double polish_time = 0.0;
    int polish_iters = 0;
    if (s.P.feasibility_polishing && s.reason != TERMINATION_REASON_DUAL_INFEASIBLE &&
        s.reason != TERMINATION_REASON_PRIMAL_INFEASIBLE)  // solver.cu:156-160 (planted :187-190)
        feasibility_polish(s, polish_time, polish_iters);
    return result_from_state(s, prob, polish_time, polish_iters);
}
#endif

static result_t *result_from_state(S &s, const qp_problem_t *prob, double polish_time, int polish_iters)
{   // create_result_from_state, fa_cp_core_op.cu:1171-1239 (LIN_CP_AL cp_al_core_op.cu:772-837)
    result_t *r = (result_t *)calloc(1, sizeof(result_t));
    spmv(s.At, s.yT.data(), s.dual_prod.data());
    r->primal_solution = (double *)malloc((size_t)std::max(s.n, 1) * sizeof(double));
    r->dual_solution = (double *)malloc((size_t)std::max(s.m, 1) * sizeof(double));
    r->reduced_cost = (double *)malloc((size_t)std::max(s.n, 1) * sizeof(double));
    for (int j = 0; j < s.n; ++j) {
        r->reduced_cost[j] = (s.c[j] - s.dual_prod[j]) * s.D[j] / s.c_scale;
        r->primal_solution[j] = s.xT[j] / s.D[j] / s.b_scale;
    }
    for (int i = 0; i < s.m; ++i) r->dual_solution[i] = s.yT[i] / s.E[i] / s.c_scale;
    r->num_variables = prob->num_variables;
    r->num_constraints = prob->num_constraints;
    r->num_nonzeros = prob->constraint_matrix_num_nonzeros;
    r->total_count = s.total;
    r->total_inner_count = (int)s.inner_total;
    r->rescaling_time_sec = s.rescale_time;
    r->cumulative_time_sec = s.solve_time;
    r->absolute_primal_residual = s.abs_p;
    r->relative_primal_residual = s.rel_p;
    r->absolute_dual_residual = s.abs_d;
    r->relative_dual_residual = s.rel_d;
    r->primal_objective_value = s.p_obj;
    r->dual_objective_value = s.d_obj;
    r->objective_gap = s.gap;
    r->relative_objective_gap = s.rel_gap;
    r->max_primal_ray_infeasibility = s.max_p_ray;
    r->max_dual_ray_infeasibility = s.max_d_ray;
    r->primal_ray_linear_objective = s.p_ray_obj;
    r->dual_ray_objective = s.d_ray_obj;
#if !defined(ORACLE_CP_AL_TYPES) && !defined(ORACLE_PLANTED_FA_CP)
    r->final_al_sigma = s.al_sigma;
    r->sigma_update_count = s.sigma_update_count;
    r->sigma_update_rule = s.al_available ? s.P.sigma_update_rule : FA_CP_SIGMA_UPDATE_FIXED;
#endif
    r->termination_reason = s.reason;
    r->feasibility_polishing_time = polish_time;
    r->feasibility_iteration = polish_iters;
    return r;
}

ORACLE_API result_t *solve_qp_problem(const qp_problem_t *prob, const params_t *params)
{   // fa_cp.c:352-376
    if (!prob) {
        fprintf(stderr, "[interface] solve_qp_problem: invalid arguments.\n");
        return nullptr;
    }
    params_t local;
    if (params) local = *params;
    else set_default_parameters(&local);
    return optimize(&local, prob);
}
