// Elementwise kernels of the linearized CP-AL iteration (RHR-Lin-CP-AL,
// QP/LIN_CP_AL/src/cp_al_core_op.cu) and of the planted-KKT CP-AL (RHR-CP-AL,
// QP/PLANTED_CP_AL/src/cp_al_core_op.cu).  The SpMVs around them (A.x, At.u,
// A.x_bar) and Q.x are the FA_CP path's.  Equality constraints: the right-hand
// side is the constraint lower bound l.
#pragma once

#include "fa_cp_kernels.hpp"

namespace pdal {
namespace qp {

// AL effective dual u = y - sigma (A x - l) (compute_al_effective_dual_kernel, :149-158).
static __global__ void __launch_bounds__(kBlock)
    lin_al_dual_kernel(const double *__restrict__ ax, const double *__restrict__ y, const double *__restrict__ l,
                       double sigma, double *__restrict__ u, int m)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < m) u[i] = y[i] - sigma * (ax[i] - l[i]);
}

// LP primal step with the T-output and dual slack every iteration
// (compute_lp_next_cp_al_primal_solution_major_kernel, :73-87).
static __global__ void __launch_bounds__(kBlock)
    lin_lp_primal_kernel(const double *__restrict__ x, const double *__restrict__ atu, const double *__restrict__ c,
                         const double *__restrict__ lb, const double *__restrict__ ub, double tau,
                         double *__restrict__ xT, double *__restrict__ xR, double *__restrict__ dslack, int n)
{
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const double xj = x[j];
    const double t = xj - tau * (c[j] - atu[j]);
    const double p = fmax(lb[j], fmin(t, ub[j]));
    xR[j] = 2.0 * p - xj;
    xT[j] = p;
    dslack[j] = (p - t) / tau;
}

// One projected gradient step on 1/2 x'Qx + c'x - u'A x
// (primal_gradient_descent_kernel_major, :353-369).
static __global__ void __launch_bounds__(kBlock)
    lin_grad_primal_kernel(const double *__restrict__ x, const double *__restrict__ qx, const double *__restrict__ c,
                           const double *__restrict__ atu, const double *__restrict__ lb,
                           const double *__restrict__ ub, double tau, double *__restrict__ xT,
                           double *__restrict__ xR, int n)
{
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const double xj = x[j];
    const double g = qx[j] + c[j] - atu[j];
    const double p = fmax(lb[j], fmin(xj - tau * g, ub[j]));
    xT[j] = p;
    xR[j] = 2.0 * p - xj;
}

// Equality dual step y + step (l - A x_bar) with its reflection
// (compute_next_cp_al_(al_)dual_solution_major_kernel, :135-147, :173-185); the
// T-output only when WRITE_T (the planted solver's non-major kernels, PLANTED_CP_AL
// :122-133, :160-171, leave it).
template <bool WRITE_T = true>
static __global__ void __launch_bounds__(kBlock)
    lin_dual_kernel(const double *__restrict__ y, const double *__restrict__ ax, const double *__restrict__ l,
                    double dstep, double *__restrict__ yT, double *__restrict__ yR, int m)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= m) return;
    const double yi = y[i];
    const double nd = yi + dstep * (l[i] - ax[i]);
    if (WRITE_T) yT[i] = nd;
    yR[i] = 2.0 * nd - yi;
}

// Planted CP-AL primal step for LP (DIAG = false) or diagonal Q in closed form,
// x+ = P((x - tau (c - At.u)) / (1 + tau q)); the T-output (and the LP dual
// slack) only in major-kernel iterations (PLANTED_CP_AL :59-120).
template <bool DIAG, bool MAJOR>
static __global__ void __launch_bounds__(kBlock)
    pcal_primal_kernel(const double *__restrict__ x, const double *__restrict__ atu, const double *__restrict__ c,
                       const double *__restrict__ lb, const double *__restrict__ ub, const double *__restrict__ qdiag,
                       double tau, double *__restrict__ xT, double *__restrict__ xR, double *__restrict__ dslack,
                       int n)
{
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const double xj = x[j];
    double t = xj - tau * (c[j] - atu[j]);
    if (DIAG) t = t / (1.0 + tau * qdiag[j]);
    const double p = fmax(lb[j], fmin(t, ub[j]));
    xR[j] = 2.0 * p - xj;
    if (MAJOR) {
        xT[j] = p;
        if (!DIAG) dslack[j] = (p - t) / tau;
    }
}

// Relaxed Halpern step z = w ((1 + rho) T - rho z) + (1 - w) z0 (halpern_update_kernel, :187-206).
static __global__ void __launch_bounds__(kBlock)
    lin_halpern_kernel(double *__restrict__ x, const double *__restrict__ xT, const double *__restrict__ x0, int n,
                       double *__restrict__ y, const double *__restrict__ yT, const double *__restrict__ y0, int m,
                       double w, double rho)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const double rl = (1.0 + rho) * xT[i] - rho * x[i];
        x[i] = w * rl + (1.0 - w) * x0[i];
    } else if (i < n + m) {
        const int r = i - n;
        const double rl = (1.0 + rho) * yT[r] - rho * y[r];
        y[r] = w * rl + (1.0 - w) * y0[r];
    }
}

}  // namespace qp
}  // namespace pdal
