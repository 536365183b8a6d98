// gfx950 kernels of the FA_CP iteration (QP/FA_CP/src/fa_cp_core_op.cu),
// its projected Barzilai-Borwein inner solve, residual / fixed-point /
// infeasibility evaluations and the device-side rescaling.  FP64, int32 CSR.
//
// Matrices in HBM: A (m rows), At (n rows), Q and Qt (n rows, sparse part of
// the Hessian), R (r rows) and Rt (n rows, low-rank factor); each as separate
// row_ptr / col / val arrays.  Q.x means Q x + R^T (R x) as in the reference.
#pragma once

#include "../common/device_reduce.hpp"
#include "qp_variant.hpp"
#include "../common/spmv.hpp"

namespace pdal {
namespace qp {

// Per-iteration scalars, passed by value as kernel arguments.
struct StepScalars {
    double tau;       // primal step eta/pw (nonconvexity-adjusted)
    double sigma;     // dual step eta*pw
    double al_sigma;  // AL penalty (0 when the AL branch is off)
    double weight;    // Halpern weight (inner+1)/(inner+2)
    double rho;       // reflection coefficient
    double tau_res;   // eta/pw for the QP residual's recovered slack
};

// Projected-BB inner-solve state (fa_cp_core_op.cu:445-503), device resident.
struct BBState {
    double alpha;      // current step
    double dnorm;      // ||direction|| after the latest update
    int it;            // inner iteration counter (starts at 1)
    int done;          // loop finished
    long long total;   // accumulated inner iterations (reported as total_inner_count)
};

__device__ __forceinline__ double clampd(double v, double lo, double hi) { return fmax(lo, fmin(v, hi)); }

// ---------------------------------------------------------------------------
// Primal updates without an inner loop: At.y fused with the LP / diagonal-Q
// prox (kernels :93-154).  HALPERN folds the x Halpern step (:225-242) in when
// no fixed-point evaluation needs the pre-Halpern iterate this iteration.
struct SimplePrimalArgs {
    CsrView At;
    const double *__restrict__ y;
    const double *__restrict__ c, *__restrict__ lb, *__restrict__ ub;
    const double *__restrict__ qdiag;  // nullptr for the LP case
    const double *__restrict__ x0;
    double *__restrict__ x, *__restrict__ xR, *__restrict__ xT, *__restrict__ dslack, *__restrict__ dual_prod;
    StepScalars sc;
};

template <int G, bool MAJOR, bool DIAG, bool HALPERN>
static __global__ void __launch_bounds__(kBlock) simple_primal_kernel(SimplePrimalArgs a)
{
    const int tid = blockIdx.x * kBlock + threadIdx.x;
    const int row = tid / G;
    const int lane = tid & (G - 1);
    if (row >= a.At.rows) return;
    const double dp = row_dot<G>(a.At, row, lane, a.y);
    if (lane != 0) return;
    const double tau = a.sc.tau;
    const double xi = a.x[row];
    double t;
    if (DIAG) t = (xi - tau * (a.c[row] - dp)) / (1.0 + tau * a.qdiag[row]);
    else t = xi - tau * (a.c[row] - dp);
    const double p = fmax(a.lb[row], fmin(t, a.ub[row]));
    const double xr = 2.0 * p - xi;
    a.xR[row] = xr;
    a.dual_prod[row] = dp;
    if (MAJOR) {
        a.xT[row] = p;
        if (!DIAG) a.dslack[row] = (p - t) / tau;
    }
    if (HALPERN) {
        const double rho = a.sc.rho, w = a.sc.weight;
        const double rf = rho * xr + (1.0 - rho) * xi;
        a.x[row] = w * rf + (1.0 - w) * a.x0[row];
    }
}

// A.x_bar fused with the dual update (non-AL :156-179, AL :194-223) and,
// optionally, the y Halpern step.
struct DualArgs {
    CsrView A;
    const double *__restrict__ xR;
    const double *__restrict__ l, *__restrict__ u, *__restrict__ y0;
    double *__restrict__ y, *__restrict__ yR, *__restrict__ yT;
    StepScalars sc;
};

template <int G, bool MAJOR, bool AL, bool HALPERN>
static __global__ void __launch_bounds__(kBlock) dual_kernel(DualArgs a)
{
    const int tid = blockIdx.x * kBlock + threadIdx.x;
    const int row = tid / G;
    const int lane = tid & (G - 1);
    if (row >= a.A.rows) return;
    const double pp = row_dot<G>(a.A, row, lane, a.xR);
    if (lane != 0) return;
    const double sigma = a.sc.sigma;
    const double yi = a.y[row];
    double yr;
    if (AL) {
        double nd;
        if (kEqualityAL) {
            nd = yi + sigma * (a.l[row] - pp);
        } else {
            const double den = a.sc.al_sigma + sigma;
            const double sh = pp - yi / den;
            const double pr = fmax(a.l[row], fmin(sh, a.u[row]));
            nd = yi + sigma * (pr - pp);
        }
        if (MAJOR) a.yT[row] = nd;
        yr = 2.0 * nd - yi;
    } else {
        const double t = yi / sigma - pp;
        const double p = fmax(-a.u[row], fmin(t, -a.l[row]));
        if (MAJOR) {
            const double yt = (t - p) * sigma;
            a.yT[row] = yt;
            yr = 2.0 * yt - yi;
        } else {
            yr = 2.0 * (t - p) * sigma - yi;
        }
    }
    a.yR[row] = yr;
    if (HALPERN) {
        const double rho = a.sc.rho, w = a.sc.weight;
        const double rf = rho * yr + (1.0 - rho) * yi;
        a.y[row] = w * rf + (1.0 - w) * a.y0[row];
    }
}

// Halpern update for iterations with a fixed-point evaluation (:225-242).
static __global__ void __launch_bounds__(kBlock)
    halpern_kernel(double *__restrict__ x, const double *__restrict__ xR, const double *__restrict__ x0, int n,
                   double *__restrict__ y, const double *__restrict__ yR, const double *__restrict__ y0, int m,
                   StepScalars sc)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const double rho = sc.rho, w = sc.weight;
    if (i < n) {
        const double rf = rho * xR[i] + (1.0 - rho) * x[i];
        x[i] = w * rf + (1.0 - w) * x0[i];
    } else if (i < n + m) {
        const int r = i - n;
        const double rf = rho * yR[r] + (1.0 - rho) * y[r];
        y[r] = w * rf + (1.0 - w) * y0[r];
    }
}

// ---------------------------------------------------------------------------
// Q.x products (solver_state.cu:371-442).  MODE: 0 sparse Q, 1 R^T(Rx) from a
// precomputed rx, 2 Q x (precomputed in qx) + R^T rx, 5 R^T rx + Q x with the
// sparse part summed in the same pass (the BB gradient of the low-rank-plus-
// sparse form: one launch less per inner step).  Plain form writes qx.
template <int G, int MODE>
__device__ __forceinline__ double qrow(const CsrView &Qm, const CsrView &Rt, int row, int lane,
                                       const double *__restrict__ xv, const double *__restrict__ rx,
                                       const double *__restrict__ qx_in)
{
    if (MODE == 0) return row_dot<G>(Qm, row, lane, xv);
    const double t = row_dot<G>(Rt, row, lane, rx);
    if (MODE == 1) return t;
    if (MODE == 5) return t + row_dot<G>(Qm, row, lane, xv);
    return t + qx_in[row];
}

template <int G, int MODE>
static __global__ void __launch_bounds__(kBlock)
    qx_kernel(CsrView Qm, CsrView Rt, const double *__restrict__ xv, const double *__restrict__ rx,
              double *__restrict__ qx, int n, const BBState *__restrict__ bb)
{
    if (bb && bb->done) return;
    const int tid = blockIdx.x * kBlock + threadIdx.x;
    const int row = tid / G;
    const int lane = tid & (G - 1);
    if (row >= n) return;
    const double s = qrow<G, MODE>(Qm, Rt, row, lane, xv, rx, qx);
    if (lane == 0) qx[row] = s;
}

// y = M x gated by the BB loop flag (R.x, A.x, At.u inside the inner loop).
template <int G>
static __global__ void __launch_bounds__(kBlock)
    gated_spmv_kernel(CsrView M, const double *__restrict__ x, double *__restrict__ y,
                      const BBState *__restrict__ bb)
{
    if (bb && bb->done) return;
    const int tid = blockIdx.x * kBlock + threadIdx.x;
    const int row = tid / G;
    const int lane = tid & (G - 1);
    if (row >= M.rows) return;
    const double s = row_dot<G>(M, row, lane, x);
    if (lane == 0) y[row] = s;
}

// AL effective dual u = y - s(Ax - proj(Ax - y/s)) from A.x (:181-192, :275-303).
template <int G>
static __global__ void __launch_bounds__(kBlock)
    al_effective_dual_kernel(CsrView A, const double *__restrict__ xv, const double *__restrict__ y,
                             const double *__restrict__ l, const double *__restrict__ u,
                             double *__restrict__ pp_out, double *__restrict__ uo,
                             StepScalars sc, const BBState *__restrict__ bb)
{
    if (bb && bb->done) return;
    const int tid = blockIdx.x * kBlock + threadIdx.x;
    const int row = tid / G;
    const int lane = tid & (G - 1);
    if (row >= A.rows) return;
    const double pp = row_dot<G>(A, row, lane, xv);
    if (lane != 0) return;
    const double s = sc.al_sigma;
    const double yi = y[row];
    pp_out[row] = pp;
    if (kEqualityAL) {
        uo[row] = yi - s * (pp - l[row]);
    } else {
        const double sh = pp - yi / s;
        const double pr = fmax(l[row], fmin(sh, u[row]));
        uo[row] = yi - s * (pp - pr);
    }
}

// ---------------------------------------------------------------------------
// Projected BB inner loop.
struct BBArgs {
    const double *__restrict__ c, *__restrict__ lb, *__restrict__ ub;
    const double *__restrict__ x;      // outer iterate (fixed during the loop)
    const double *__restrict__ dual_prod;
    const double *__restrict__ qx;
    double *__restrict__ xT, *__restrict__ grad, *__restrict__ dir;
    double *__restrict__ parts;        // per-block partials of the kernel
    const double *__restrict__ prev;   // the previous kernel's per-block partials
    int nprev;                         // how many
    BBState *bb;
    StepScalars sc;
    int n;
    double tol;                        // inner stopping tolerance on ||direction||
    int limit;                         // inner iteration limit
};

// Sum of parts[0..nparts) in the fixed order of the single-block reductions
// (thread t over b = t, t + 256, ...; lane butterfly; wavefronts in order), so
// every block that runs it gets the same bits.  The result is in every thread.
__device__ __forceinline__ double block_sum_parts(const double *parts, int nparts)
{
    __shared__ double red[kWavesPerBlock];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nparts; b += kBlock) acc += parts[b];
    acc = wave_sum(acc);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = acc;
    __syncthreads();
    double r = red[0];
    for (int w = 1; w < kWavesPerBlock; ++w) r += red[w];
    __syncthreads();
    return r;
}

// The loop test after a direction update (:403-408): every block of the
// calling kernel reduces the direction partials of the previous kernel itself
// and reaches the same decision from the same bits (the counter it was
// advanced by that kernel and is only read here); block 0 records ||d|| and
// the flag.  Returns true when the loop has ended (the caller returns at once).
__device__ __forceinline__ bool bb_loop_test(const BBArgs &a)
{
    const double dn = sqrt(block_sum_parts(a.prev, a.nprev));
    const bool done = !(a.bb->it < a.limit) || dn <= a.tol;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.bb->dnorm = dn;
        if (done) a.bb->done = 1;
    }
    return done;
}

// bb_init (:381-399): gradient at x, first projected step with alpha = 1/(1/tau).
static __global__ void __launch_bounds__(kBlock) bb_init_kernel(BBArgs a)
{
    const double alpha = 1.0 / (1.0 / a.sc.tau);
    double q[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += gridDim.x * kBlock) {
        const double g = a.qx[i] + a.c[i] - a.dual_prod[i];
        const double xi = a.x[i];
        double nx = xi - alpha * g;
        nx = fmax(a.lb[i], fmin(nx, a.ub[i]));
        a.xT[i] = nx;
        a.grad[i] = g;
        const double d = nx - xi;
        a.dir[i] = d;
        q[0] += d * d;
    }
    block_reduce_store<1, 0>(q, a.parts + blockIdx.x);
}

// Gradient refresh (:401-415) fused with the product row: cur = Q.xT + c - dp +
// (xT - x)/tau; grad = cur; partial <dir, cur - grad_old>.  TEST: first run the
// loop test of the direction update before it (a.prev = its partials).
template <int G, int MODE, bool TEST>
static __global__ void __launch_bounds__(kBlock)
    bb_grad_kernel(BBArgs a, CsrView Qm, CsrView Rt, const double *__restrict__ rx, const double *__restrict__ qdiag)
{
    if (a.bb->done) return;
    if (TEST && bb_loop_test(a)) return;  // the previous step's loop test, folded in
    constexpr int RPB = kBlock / G;
    const double inv = 1.0 / a.sc.tau;
    double q[1] = {0.0};
    const int lane = threadIdx.x & (G - 1);
    for (int base = blockIdx.x * RPB; base < a.n; base += gridDim.x * RPB) {
        const int row = base + threadIdx.x / G;
        if (row < a.n) {
            double s;
            if (MODE == 3) s = 0.0;                      // no quadratic term
            else if (MODE == 4) s = qdiag[row] * a.xT[row];  // diagonal Q
            else s = qrow<G, MODE>(Qm, Rt, row, lane, a.xT, rx, a.qx);
            if (lane == 0) {
                const double last = a.grad[row];
                const double cur = s + a.c[row] - a.dual_prod[row] + inv * (a.xT[row] - a.x[row]);
                const double dg = cur - last;
                a.grad[row] = cur;
                q[0] += a.dir[row] * dg;
            }
        }
    }
    block_reduce_store<1, 0>(q, a.parts + blockIdx.x);
}

// Direction update (:417-430) with partial ||d||^2; the step alpha = ||d||^2 /
// <d, dg> (:412-413) is computed by every block from the gradient kernel's
// partials (a.prev), block 0 records it.
static __global__ void __launch_bounds__(kBlock) bb_dir_kernel(BBArgs a)
{
    if (a.bb->done) return;
    const double alpha = a.bb->dnorm * a.bb->dnorm / block_sum_parts(a.prev, a.nprev);
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // no block of this kernel reads these
        a.bb->alpha = alpha;
        a.bb->it += 1;  // inner_solver_iter++ after the direction update (:431)
    }
    double q[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += gridDim.x * kBlock) {
        const double cs = a.xT[i];
        double nx = cs - alpha * a.grad[i];
        nx = fmax(a.lb[i], fmin(nx, a.ub[i]));
        const double d = nx - cs;
        a.dir[i] = d;
        a.xT[i] = nx;
        q[0] += d * d;
    }
    block_reduce_store<1, 0>(q, a.parts + blockIdx.x);
}

// ||d|| and the loop test as a kernel of its own: FIRST after bb_init (it = 1);
// otherwise after the last direction update of a chunk (the direction kernel
// has advanced it).
template <bool FIRST>
static __global__ void __launch_bounds__(kBlock)
    bb_check_kernel(const double *__restrict__ parts, int nparts, BBState *bb, double tol, int limit)
{
    if (!FIRST && bb->done) return;
    __shared__ double red[kWavesPerBlock];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nparts; b += kBlock) acc += parts[b];
    acc = wave_sum(acc);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int w = 1; w < kWavesPerBlock; ++w) r += red[w];
        if (FIRST) {
            bb->it = 1;
            bb->done = 0;
        }
        bb->dnorm = sqrt(r);
        if (!(bb->it < limit) || bb->dnorm <= tol) bb->done = 1;
    }
}

// Final reflection (:432-443) and the inner-count accumulation; optionally the
// x Halpern step.
template <bool HALPERN>
static __global__ void __launch_bounds__(kBlock)
    bb_final_kernel(double *__restrict__ x, const double *__restrict__ xT, double *__restrict__ xR,
                    const double *__restrict__ x0, int n, BBState *bb, StepScalars sc)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i == 0) {
        bb->total += bb->it - 1;
    }
    if (i >= n) return;
    const double xi = x[i];
    const double xr = 2.0 * xT[i] - xi;
    xR[i] = xr;
    if (HALPERN) {
        const double rho = sc.rho, w = sc.weight;
        const double rf = rho * xr + (1.0 - rho) * xi;
        x[i] = w * rf + (1.0 - w) * x0[i];
    }
}

// ---------------------------------------------------------------------------
// Evaluations.  Partial layout: kEval stride doubles per block.
constexpr int kEvalStride = 12;

// Residual, rows of A (utils.cu:705-770, constraint part): A.xT, primal
// residual (sum sq, max abs) and the dual-objective contribution.
enum ResAQ { RA_PRES2, RA_CONTRIB, RA_PRESMAX, RA_N };
template <int G>
static __global__ void __launch_bounds__(kBlock)
    res_a_kernel(CsrView A, const double *__restrict__ xT, const double *__restrict__ yT,
                 const double *__restrict__ l, const double *__restrict__ u, const double *__restrict__ E,
                 double *__restrict__ parts)
{
    constexpr int RPB = kBlock / G;
    double q[RA_N] = {0, 0, 0};
    const int lane = threadIdx.x & (G - 1);
    for (int base = blockIdx.x * RPB; base < A.rows; base += gridDim.x * RPB) {
        const int row = base + threadIdx.x / G;
        if (row < A.rows) {
            const double pp = row_dot<G>(A, row, lane, xT);
            if (lane == 0) {
                const double lo = l[row], hi = u[row];
                const double pr = (pp - fmax(lo, fmin(pp, hi))) * E[row];
                q[RA_PRES2] += pr * pr;
                q[RA_PRESMAX] = fmax(q[RA_PRESMAX], fabs(pr));
                const double yt = yT[row];
                q[RA_CONTRIB] += fmax(yt, 0.0) * (isfinite(lo) ? lo : 0.0) + fmin(yt, 0.0) * (isfinite(hi) ? hi : 0.0);
            }
        }
    }
    block_reduce_store<2, 1>(q, parts + (size_t)blockIdx.x * kEvalStride);
}

// Residual, rows of At: At.yT, dual residual (LP: with the stored slack; QP:
// slack recovered from a projected step at tau_res, which is written back),
// objective dots and the QP dual-residual denominators.
enum ResAtQ { RT_DRES2, RT_CX, RT_SX, RT_XQX, RT_QXD2, RT_DPD2, RT_DRESMAX, RT_QXDMAX, RT_DPDMAX, RT_N };
template <int G, bool QP>
static __global__ void __launch_bounds__(kBlock)
    res_at_kernel(CsrView At, const double *__restrict__ yT, const double *__restrict__ xT,
                  const double *__restrict__ qx, const double *__restrict__ c, const double *__restrict__ lb,
                  const double *__restrict__ ub, const double *__restrict__ D, double *__restrict__ dslack,
                  StepScalars sc, double *__restrict__ parts)
{
    constexpr int RPB = kBlock / G;
    double q[RT_N] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int lane = threadIdx.x & (G - 1);
    const double tr = sc.tau_res;
    for (int base = blockIdx.x * RPB; base < At.rows; base += gridDim.x * RPB) {
        const int row = base + threadIdx.x / G;
        if (row < At.rows) {
            const double dp = row_dot<G>(At, row, lane, yT);
            if (lane == 0) {
                const double x = xT[row];
                double dres, ds;
                if (QP) {
                    const double g = qx[row] + c[row] - dp;
                    const double t = x - tr * g;
                    const double p = fmax(lb[row], fmin(ub[row], t));
                    ds = (p - t) / tr;
                    dres = (g - ds) * D[row];
                    dslack[row] = ds;
                    const double qd = qx[row] * D[row], dd = dp * D[row];
                    q[RT_XQX] += x * qx[row];
                    q[RT_QXD2] += qd * qd;
                    q[RT_DPD2] += dd * dd;
                    q[RT_QXDMAX] = fmax(q[RT_QXDMAX], fabs(qd));
                    q[RT_DPDMAX] = fmax(q[RT_DPDMAX], fabs(dd));
                } else {
                    ds = dslack[row];
                    dres = (c[row] - dp - ds) * D[row];
                }
                q[RT_DRES2] += dres * dres;
                q[RT_DRESMAX] = fmax(q[RT_DRESMAX], fabs(dres));
                q[RT_CX] += c[row] * x;
                q[RT_SX] += ds * x;
            }
        }
    }
    block_reduce_store<6, 3>(q, parts + (size_t)blockIdx.x * kEvalStride);
}

// Fixed-point error (:1042-1089): dx = xR - x, dy = yR - y (both kept for the
// infeasibility test), At.dy cross term and the two norms.
enum FpQ { FP_CROSS, FP_PN2, FP_N };
template <int G>
static __global__ void __launch_bounds__(kBlock)
    fp_at_kernel(CsrView At, const double *__restrict__ yR, const double *__restrict__ y,
                 const double *__restrict__ xR, const double *__restrict__ x, double *__restrict__ dx,
                 double *__restrict__ dual_prod, double *__restrict__ parts)
{
    constexpr int RPB = kBlock / G;
    double q[FP_N] = {0, 0};
    const int lane = threadIdx.x & (G - 1);
    for (int base = blockIdx.x * RPB; base < At.rows; base += gridDim.x * RPB) {
        const int row = base + threadIdx.x / G;
        if (row < At.rows) {
            double sd, sy;
            row_dot2<G>(At, row, lane, yR, y, sd, sy);
            if (lane == 0) {
                const double d = xR[row] - x[row];
                dx[row] = d;
                dual_prod[row] = sd;
                q[FP_CROSS] += sd * d;
                q[FP_PN2] += d * d;
            }
        }
    }
    block_reduce_store<2, 0>(q, parts + (size_t)blockIdx.x * kEvalStride);
}

// dy = a - b with partial ||dy||^2 (fixed-point and restart displacements).
static __global__ void __launch_bounds__(kBlock)
    diff_norm_kernel(const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ d, int len,
                     double *__restrict__ parts)
{
    double q[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < len; i += gridDim.x * kBlock) {
        const double t = a[i] - b[i];
        d[i] = t;
        q[0] += t * t;
    }
    block_reduce_store<1, 0>(q, parts + (size_t)blockIdx.x * kEvalStride);
}

// ||M x||^2 partials (normal motion at restart).
template <int G>
static __global__ void __launch_bounds__(kBlock)
    spmv_norm_kernel(CsrView M, const double *__restrict__ x, double *__restrict__ y, double *__restrict__ parts)
{
    constexpr int RPB = kBlock / G;
    double q[1] = {0.0};
    const int lane = threadIdx.x & (G - 1);
    for (int base = blockIdx.x * RPB; base < M.rows; base += gridDim.x * RPB) {
        const int row = base + threadIdx.x / G;
        if (row < M.rows) {
            const double s = row_dot<G>(M, row, lane, x);
            if (lane == 0) {
                y[row] = s;
                q[0] += s * s;
            }
        }
    }
    block_reduce_store<1, 0>(q, parts + (size_t)blockIdx.x * kEvalStride);
}

// Restart copy: anchor = current = T-output (:998-1010).
static __global__ void __launch_bounds__(kBlock)
    restart_copy_kernel(const double *__restrict__ xT, double *__restrict__ x0, double *__restrict__ x, int n,
                        const double *__restrict__ yT, double *__restrict__ y0, double *__restrict__ y, int m)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        x0[i] = xT[i];
        x[i] = xT[i];
    } else if (i < n + m) {
        const int r = i - n;
        y0[r] = yT[r];
        y[r] = yT[r];
    }
}

// ---------------------------------------------------------------------------
// Infeasibility information (utils.cu:1028-1132), run at most three times.
// Step 1: project the rays by the bounds, partial max |.|.
static __global__ void __launch_bounds__(kBlock)
    ray_project_kernel(double *__restrict__ dx, const double *__restrict__ lb, const double *__restrict__ ub, int n,
                       double *__restrict__ dy, const double *__restrict__ l, const double *__restrict__ u, int m,
                       double *__restrict__ parts)
{
    double q[2] = {0.0, 0.0};  // max |dx|, max |dy| (as maxima)
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        double d = dx[i];
        if (isfinite(lb[i])) d = fmax(d, 0.0);
        if (isfinite(ub[i])) d = fmin(d, 0.0);
        dx[i] = d;
        q[0] = fmax(q[0], fabs(d));
    }
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) {
        double d = dy[i];
        if (!isfinite(l[i])) d = fmin(d, 0.0);
        if (!isfinite(u[i])) d = fmax(d, 0.0);
        dy[i] = d;
        q[1] = fmax(q[1], fabs(d));
    }
    block_reduce_store<0, 2>(q, parts + (size_t)blockIdx.x * kEvalStride);
}

static __global__ void __launch_bounds__(kBlock) scale_kernel(double *__restrict__ v, double s, int len)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < len) v[i] *= s;
}

// Step 2 (after A.dx -> pp and At.dy -> dp): ray objective pieces and the two
// infeasibility vectors; dual_slack and primal_slack are overwritten as in the
// reference.  Sums: c.dx, sum pslack_obj, sum dslack_obj; maxima: primal and
// dual infeasibility.
enum RayQ { RY_CDX, RY_PSUM, RY_DSUM, RY_PMAX, RY_DMAX, RY_N };
static __global__ void __launch_bounds__(kBlock)
    ray_terms_kernel(const double *__restrict__ c, const double *__restrict__ dx, const double *__restrict__ dp,
                     const double *__restrict__ lb, const double *__restrict__ ub, const double *__restrict__ D,
                     double *__restrict__ dslack, int n, const double *__restrict__ dy,
                     const double *__restrict__ pp, const double *__restrict__ l, const double *__restrict__ u,
                     const double *__restrict__ E, double *__restrict__ pslack, int m, double *__restrict__ parts)
{
    double q[RY_N] = {0, 0, 0, 0, 0};
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        q[RY_CDX] += c[j] * dx[j];
        const double lbf = isfinite(lb[j]) ? lb[j] : 0.0, ubf = isfinite(ub[j]) ? ub[j] : 0.0;
        const double neg = -dp[j];
        q[RY_DSUM] += fmax(neg, 0.0) * lbf + fmin(neg, 0.0) * ubf;
        const double di = (fmax(0.0, neg) * (double)!isfinite(lb[j]) - fmin(0.0, neg) * (double)!isfinite(ub[j])) * D[j];
        dslack[j] = di;
        q[RY_DMAX] = fmax(q[RY_DMAX], fabs(di));
    }
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) {
        const double lo = l[i], hi = u[i];
        q[RY_PSUM] += fmax(dy[i], 0.0) * (isfinite(lo) ? lo : 0.0) + fmin(dy[i], 0.0) * (isfinite(hi) ? hi : 0.0);
        const double p = pp[i];
        const double pi = (fmax(0.0, -p) * (double)isfinite(lo) + fmax(0.0, p) * (double)isfinite(hi)) * E[i];
        pslack[i] = pi;
        q[RY_PMAX] = fmax(q[RY_PMAX], fabs(pi));
    }
    block_reduce_store<3, 2>(q, parts + (size_t)blockIdx.x * kEvalStride);
}

static __global__ void __launch_bounds__(kBlock) absmax_kernel(const double *__restrict__ v, int len, double *__restrict__ parts)
{
    double q[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < len; i += gridDim.x * kBlock) q[0] = fmax(q[0], fabs(v[i]));
    block_reduce_store<0, 1>(q, parts + (size_t)blockIdx.x * kEvalStride);
}

// ---------------------------------------------------------------------------
// Power iterations (utils.cu:65-315).
static __global__ void __launch_bounds__(kBlock)
    dot_sumsq_kernel(const double *__restrict__ a, const double *__restrict__ b, int len, double *__restrict__ parts)
{   // partials of <a,b> and ||b||^2
    double q[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < len; i += gridDim.x * kBlock) {
        q[0] += a[i] * b[i];
        q[1] += b[i] * b[i];
    }
    block_reduce_store<2, 0>(q, parts + (size_t)blockIdx.x * kEvalStride);
}

// ---------------------------------------------------------------------------
// Device rescaling (QP/FA_CP/src/preconditioner.c:92-289 moved to the GPU).
// Column maxima / power sums come from the transposed copies so every column
// is reduced by one thread in the reference's row order.
static __global__ void __launch_bounds__(kBlock)
    ruiz_factors_kernel(CsrView A, CsrView At, CsrView Qt, double *__restrict__ e, double *__restrict__ d)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < A.rows) {
        double mx = 0.0;
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const double v = fabs(A.v[k]);
            if (v > mx) mx = v;
        }
        e[i] = (mx < 1e-12) ? 1.0 : sqrt(mx);
    }
    if (i < At.rows) {
        double mx = 0.0;
        for (int k = At.rp[i]; k < At.rp[i + 1]; ++k) {
            const double v = fabs(At.v[k]);
            if (v > mx) mx = v;
        }
        if (Qt.rows > 0)
            for (int k = Qt.rp[i]; k < Qt.rp[i + 1]; ++k) {
                const double v = fabs(Qt.v[k]);
                if (v > mx) mx = v;
            }
        d[i] = (mx < 1e-12) ? 1.0 : sqrt(mx);
    }
}

// |v|^p with the exact special cases of the usual exponents.
__device__ __forceinline__ double abs_pow(double v, double p)
{
    v = fabs(v);
    if (p == 1.0) return v;
    if (p == 2.0) return v * v;
    if (p == 0.5) return sqrt(v);
    return pow(v, p);
}

static __global__ void __launch_bounds__(kBlock)
    pc_factors_kernel(CsrView A, CsrView At, CsrView Qt, double alpha, double *__restrict__ e, double *__restrict__ d)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < A.rows) {
        double acc = 0.0;
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) acc += abs_pow(A.v[k], alpha);
        e[i] = (acc < 1e-12) ? 1.0 : sqrt(acc);
    }
    if (i < At.rows) {
        double acc = 0.0;
        for (int k = At.rp[i]; k < At.rp[i + 1]; ++k) acc += abs_pow(At.v[k], 2.0 - alpha);
        if (Qt.rows > 0)
            for (int k = Qt.rp[i]; k < Qt.rp[i + 1]; ++k) acc += abs_pow(Qt.v[k], 2.0 - alpha);
        d[i] = (acc < 1e-12) ? 1.0 : sqrt(acc);
    }
}

// In-place division of a CSR matrix's values: MODE 0 by row_f[row]*col_f[col]
// (the two factors of an entry multiply commutatively, so A and At, Q and Qt
// receive identical bits), MODE 1 by col_f[col], MODE 2 by row_f[row].
template <int MODE>
static __global__ void __launch_bounds__(kBlock)
    divide_rows_kernel(CsrView M, double *__restrict__ v, const double *__restrict__ row_f,
                       const double *__restrict__ col_f)
{
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= M.rows) return;
    for (int k = M.rp[r]; k < M.rp[r + 1]; ++k) {
        if (MODE == 0) v[k] /= (row_f[r] * col_f[M.ci[k]]);
        else if (MODE == 1) v[k] /= col_f[M.ci[k]];
        else v[k] /= row_f[r];
    }
}

static __global__ void __launch_bounds__(kBlock)
    scale_vectors_kernel(double *__restrict__ c, double *__restrict__ lb, double *__restrict__ ub,
                         const double *__restrict__ d, double *__restrict__ cumD, int n, double *__restrict__ l,
                         double *__restrict__ u, const double *__restrict__ e, double *__restrict__ cumE, int m)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        c[i] /= d[i];
        ub[i] *= d[i];
        lb[i] *= d[i];
        cumD[i] *= d[i];
    } else if (i < n + m) {
        const int r = i - n;
        l[r] /= e[r];
        u[r] /= e[r];
        cumE[r] *= e[r];
    }
}

static __global__ void __launch_bounds__(kBlock)
    bound_obj_sums_kernel(const double *__restrict__ l, const double *__restrict__ u, int m,
                          const double *__restrict__ c, int n, double *__restrict__ parts)
{
    double q[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) {
        const double lo = l[i], hi = u[i];
        if (isfinite(lo) && lo != hi) q[0] += lo * lo;
        if (isfinite(hi)) q[0] += hi * hi;
    }
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) q[1] += c[j] * c[j];
    block_reduce_store<2, 0>(q, parts + (size_t)blockIdx.x * kEvalStride);
}

static __global__ void __launch_bounds__(kBlock)
    bound_obj_apply_kernel(double *__restrict__ l, double *__restrict__ u, int m, double *__restrict__ lb,
                           double *__restrict__ ub, double *__restrict__ c, int n, double bs, double cs)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < m) {
        l[i] *= bs;
        u[i] *= bs;
    } else if (i < m + n) {
        const int j = i - m;
        lb[j] *= bs;
        ub[j] *= bs;
        c[j] *= cs;
    }
}

// Finite entries -> 0, infinite ones kept: the dual polishing problem's
// bounds (zero_finite_value_vectors_kernel, fa_cp_core_op.cu:1530-1537).
static __global__ void __launch_bounds__(kBlock) zero_finite_kernel(double *__restrict__ v, int len)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < len && isfinite(v[i])) v[i] = 0.0;
}

static __global__ void __launch_bounds__(kBlock) fill_kernel(double *__restrict__ v, double val, int len)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < len) v[i] = val;
}

static __global__ void __launch_bounds__(kBlock) mul_all_kernel(double *__restrict__ v, long len, double f)
{
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i < len) v[i] *= f;
}

// Diagonal of a diagonal-only Q plus 1e-12 (solver_state.cu:462-492).
static __global__ void __launch_bounds__(kBlock) qdiag_kernel(CsrView Qm, int n, double *__restrict__ qdiag)
{
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= Qm.rows) return;
    for (int k = Qm.rp[r]; k < Qm.rp[r + 1]; ++k)
        if (Qm.ci[k] < n) qdiag[Qm.ci[k]] = Qm.v[k] + 1e-12;
}

static __global__ void __launch_bounds__(kBlock)
    diag_mul_kernel(const double *__restrict__ d, const double *__restrict__ x, double *__restrict__ y, int len)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < len) y[i] = d[i] * x[i];
}

// Shifted power step of the minimum-eigenvalue estimate (utils.cu:240-315):
// sh = lmax*v - Qv, partial <v, sh>.
static __global__ void __launch_bounds__(kBlock)
    shift_dot_kernel(const double *__restrict__ v, const double *__restrict__ av, double lmax,
                     double *__restrict__ sh, int len, double *__restrict__ parts)
{
    double q[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < len; i += gridDim.x * kBlock) {
        const double t = lmax * v[i] + (-av[i]);
        sh[i] = t;
        q[0] += v[i] * t;
    }
    block_reduce_store<1, 0>(q, parts + (size_t)blockIdx.x * kEvalStride);
}

// Power-iteration normalisation with the norm read from the device:
// dst = src * (1/sqrt(*sumsq)).
static __global__ void __launch_bounds__(kBlock)
    normalize_dev_kernel(const double *__restrict__ src, double *__restrict__ dst, int len,
                         const double *__restrict__ sumsq)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const double inv = 1.0 / sqrt(*sumsq);
    if (i < len) dst[i] = src[i] * inv;
}

// r = -(*sig)*a + b in place into a, partial ||r||^2 (utils.cu:65-170).
static __global__ void __launch_bounds__(kBlock)
    sv_resid_kernel(double *__restrict__ a, const double *__restrict__ b, int len, const double *__restrict__ sig,
                    double *__restrict__ parts)
{
    const double neg = -(*sig);
    double q[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < len; i += gridDim.x * kBlock) {
        const double r = neg * a[i] + b[i];
        a[i] = r;
        q[0] += r * r;
    }
    block_reduce_store<1, 0>(q, parts + (size_t)blockIdx.x * kEvalStride);
}

static __global__ void __launch_bounds__(kBlock)
    start_scale_kernel(double *__restrict__ x, const double *__restrict__ D, double bs, int n, double *__restrict__ y,
                       const double *__restrict__ E, double cs, int m)
{   // warm start into the scaled space (solver_state.cu:630-657)
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] = x[i] * D[i] * bs;
    else if (i < n + m) y[i - n] = y[i - n] * E[i - n] * cs;
}

static __global__ void __launch_bounds__(kBlock)
    result_kernel(const double *__restrict__ xT, const double *__restrict__ D, double bs, double *__restrict__ xo,
                  const double *__restrict__ c, const double *__restrict__ dp, double cs, double *__restrict__ rc,
                  int n, const double *__restrict__ yT, const double *__restrict__ E, double *__restrict__ yo, int m)
{   // compute_and_rescale_reduced_cost_kernel + rescale_solution_kernel
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        rc[i] = (c[i] - dp[i]) * D[i] / cs;
        xo[i] = xT[i] / D[i] / bs;
    } else if (i < n + m) {
        yo[i - n] = yT[i - n] / E[i - n] / cs;
    }
}

}  // namespace qp
}  // namespace pdal
