// gfx950 kernels of the CP_AL iteration, evaluation, power iteration and
// rescaling.  All arithmetic is FP64 with int32 CSR indices.
//
// Layout in HBM: A (m rows) and its explicit transpose At (n rows) as three
// separate arrays each (row_ptr int32[rows+1], col int32[nnz], val f64[nnz]),
// so one SpMV streams exactly 12 B per nonzero plus 4 B per row.  Every
// solution vector is a dense FP64 array.
//
// SpMV shape: G lanes of a 64-wide wavefront own one row (G in {1..64},
// picked per matrix from the mean row length); lanes stride over the row,
// then a G-lane butterfly sums the partials and lane 0 of the group runs the
// fused epilogue (projection, reflection 2T-I, Halpern anchor average), so the
// products A.x and At.y never round-trip through HBM.
#pragma once

#include "../common/device_reduce.hpp"
#include "../common/sliced_csr.hpp"
#include "../common/spmv.hpp"

namespace pdal {
namespace lp {

struct IterScalars {  // device copy of the step schedule, written once per block
    double tau;       // primal step  eta / pw      (solver.cu:1738-1749)
    double sigma;     // dual step    eta * pw
    double beta;      // AL penalty   (al_penalty)
    int inner;        // inner_count at block start (d_inner_count)
    int pad;
};

struct PrimalArgs {
    CsrView At;                      // n rows
    const double *__restrict__ yv;   // y, or the AL predictor u when beta > 0
    const double *__restrict__ c;
    const double *__restrict__ lb;
    const double *__restrict__ ub;
    const double *__restrict__ x0;   // Halpern anchor
    double *__restrict__ x;          // current (in place)
    double *__restrict__ xR;         // reflected 2 x_T - x
    double *__restrict__ xT;         // major: T-output
    double *__restrict__ xH;         // major: Halpern average
    double *__restrict__ dslack;     // major: (x_T - x~)/tau
    const IterScalars *__restrict__ sc;
    double rho;                      // reflection coefficient
};

// At.y SpMV fused with the primal prox / projection / reflection / Halpern
// update.  Reference: compute_next_primal_solution(_major)_kernel,
// LP/CP_AL/src/solver.cu:2340-2397, launched after the At.y SpMV at :2561-2570.
template <bool MAJOR>
__device__ __forceinline__ void primal_update(const PrimalArgs &a, int row, double s, int k_offset)
{
    const double tau = a.sc->tau;
    const int k = a.sc->inner + k_offset;
    const double w = (double)k / (double)(k + 1);
    const double xj = a.x[row];
    const double temp = xj - tau * (a.c[row] - s);
    const double proj = fmax(a.lb[row], fmin(temp, a.ub[row]));
    const double x0 = a.x0[row];
    if (MAJOR) {
        a.xT[row] = proj;
        a.dslack[row] = (proj - temp) / tau;
        a.xH[row] = w * proj + (1.0 - w) * x0;
    }
    const double xr = 2.0 * proj - xj;
    a.xR[row] = xr;
    const double refl = a.rho * xr + (1.0 - a.rho) * xj;
    a.x[row] = w * refl + (1.0 - w) * x0;
}

template <int G, bool MAJOR>
static __global__ void __launch_bounds__(kBlock) primal_kernel(PrimalArgs a, int k_offset)
{
    const int tid = blockIdx.x * kBlock + threadIdx.x;
    const int row = tid / G;
    const int lane = tid & (G - 1);
    if (row >= a.At.rows) return;  // whole groups exit together (kBlock % G == 0)
    const double s = row_dot<G>(a.At, row, lane, a.yv);
    if (lane != 0) return;
    primal_update<MAJOR>(a, row, s, k_offset);
}

// The same pass over the column-sliced copy of At (one workgroup per row
// range, the gathered y slice resident in the XCD's L2; row sums in LDS).
template <int U, bool MAJOR>
static __global__ void __launch_bounds__(kSliceBlock) primal_sliced_kernel(PrimalArgs a, SlicedView S, int k_offset)
{
    extern __shared__ double acc[];
    sliced_row_sums<U>(S, blockIdx.x, a.yv, acc);
    const int rw = kSliceWavesPerGroup * S.rpw;
    for (int j = threadIdx.x; j < rw; j += kSliceBlock) {
        const long row = (long)blockIdx.x * rw + j;
        if (row < S.rows) primal_update<MAJOR>(a, (int)row, acc[j], k_offset);
    }
}

struct DualArgs {
    CsrView A;                       // m rows
    const double *__restrict__ xR;   // reflected primal (gathered)
    const double *__restrict__ l;
    const double *__restrict__ u;
    const double *__restrict__ y0;
    double *__restrict__ y;
    double *__restrict__ yT;
    double *__restrict__ yH;
    double *__restrict__ yR;
    const IterScalars *__restrict__ sc;
    double rho;
};

// A.x_bar SpMV fused with the dual prox on [l,u] with denominator sigma+beta,
// reflection and Halpern update.  Reference: compute_next_dual_solution(_major)
// _kernel, solver.cu:2399-2457, after the A.x_bar SpMV at :2617-2626.
template <bool MAJOR>
__device__ __forceinline__ void dual_update(const DualArgs &a, int row, double pp, int k_offset)
{
    const double sigma = a.sc->sigma;
    const double prox = sigma + a.sc->beta;
    const int k = a.sc->inner + k_offset;
    const double w = (double)k / (double)(k + 1);
    const double yi = a.y[row];
    const double zb = fmax(a.l[row], fmin(pp - yi / prox, a.u[row]));
    const double yn = yi - sigma * (pp - zb);
    const double y0 = a.y0[row];
    if (MAJOR) {
        a.yT[row] = yn;
        a.yH[row] = w * yn + (1.0 - w) * y0;
        a.yR[row] = 2.0 * yn - yi;
    }
    const double refl = a.rho * (2.0 * yn - yi) + (1.0 - a.rho) * yi;
    a.y[row] = w * refl + (1.0 - w) * y0;
}

template <int G, bool MAJOR>
static __global__ void __launch_bounds__(kBlock) dual_kernel(DualArgs a, int k_offset)
{
    const int tid = blockIdx.x * kBlock + threadIdx.x;
    const int row = tid / G;
    const int lane = tid & (G - 1);
    if (row >= a.A.rows) return;
    const double pp = row_dot<G>(a.A, row, lane, a.xR);
    if (lane != 0) return;
    dual_update<MAJOR>(a, row, pp, k_offset);
}

// The same pass over the column-sliced copy of A (x_bar slice in L2).
template <int U, bool MAJOR>
static __global__ void __launch_bounds__(kSliceBlock) dual_sliced_kernel(DualArgs a, SlicedView S, int k_offset)
{
    extern __shared__ double acc[];
    sliced_row_sums<U>(S, blockIdx.x, a.xR, acc);
    const int rw = kSliceWavesPerGroup * S.rpw;
    for (int j = threadIdx.x; j < rw; j += kSliceBlock) {
        const long row = (long)blockIdx.x * rw + j;
        if (row < S.rows) dual_update<MAJOR>(a, (int)row, acc[j], k_offset);
    }
}

struct PredictorArgs {
    CsrView A;
    const double *__restrict__ x;
    const double *__restrict__ y;
    const double *__restrict__ l;
    const double *__restrict__ u;
    double *__restrict__ out;  // predictor u (the reference stores it in reflected_dual)
    const IterScalars *__restrict__ sc;
};

// A.x SpMV fused with the box-AL dual predictor (only launched when beta > 0).
// Reference: compute_box_al_predictor_kernel, solver.cu:2459-2479, after :2533-2542.
__device__ __forceinline__ void predictor_update(const PredictorArgs &a, int row, double pp)
{
    const double beta = a.sc->beta;
    const double yi = a.y[row];
    if (beta <= 0.0) {
        a.out[row] = yi;
        return;
    }
    const double z = fmax(a.l[row], fmin(pp - yi / beta, a.u[row]));
    a.out[row] = yi - beta * (pp - z);
}

template <int G>
static __global__ void __launch_bounds__(kBlock) predictor_kernel(PredictorArgs a)
{
    const int tid = blockIdx.x * kBlock + threadIdx.x;
    const int row = tid / G;
    const int lane = tid & (G - 1);
    if (row >= a.A.rows) return;
    const double pp = row_dot<G>(a.A, row, lane, a.x);
    if (lane != 0) return;
    predictor_update(a, row, pp);
}

template <int U>
static __global__ void __launch_bounds__(kSliceBlock) predictor_sliced_kernel(PredictorArgs a, SlicedView S)
{
    extern __shared__ double acc[];
    sliced_row_sums<U>(S, blockIdx.x, a.x, acc);
    const int rw = kSliceWavesPerGroup * S.rpw;
    for (int j = threadIdx.x; j < rw; j += kSliceBlock) {
        const long row = (long)blockIdx.x * rw + j;
        if (row < S.rows) predictor_update(a, (int)row, acc[j]);
    }
}

// ---------------------------------------------------------------------------
// Evaluation block: fixed-point error (solver.cu:3611-3686), KKT residual
// (utils.cu:649-682, 812-917) and the restart displacement (solver.cu:2838-
// 2853), as two dual-right-hand-side SpMVs with fused reductions.

enum EvalAtQ { QAT_CROSS, QAT_PN2, QAT_DRES2, QAT_CX, QAT_LX, QAT_RX2, QAT_DRESMAX, QAT_N };
constexpr int kEvalAtSums = 6;
enum EvalAQ { QA_AX2, QA_DN2, QA_PRES2, QA_CONTRIB, QA_RY2, QA_PRESMAX, QA_N };
constexpr int kEvalASums = 5;
constexpr int kEvalStride = 8;  // partial row stride (doubles)

struct EvalAtArgs {
    CsrView At;
    const double *__restrict__ yR;
    const double *__restrict__ yT;
    const double *__restrict__ xR;
    const double *__restrict__ xT;
    const double *__restrict__ x0;
    const double *__restrict__ c;
    const double *__restrict__ dslack;
    const double *__restrict__ D;
    double *__restrict__ parts;
};

// Rows of At: (At dy)_j with dy = yR - yT and (At yT)_j in one pass.
template <int G>
static __global__ void __launch_bounds__(kBlock) eval_at_kernel(EvalAtArgs a)
{
    constexpr int RPB = kBlock / G;
    double q[QAT_N] = {0, 0, 0, 0, 0, 0, 0};
    const int lane = threadIdx.x & (G - 1);
    for (int base = blockIdx.x * RPB; base < a.At.rows; base += gridDim.x * RPB) {
        const int row = base + threadIdx.x / G;
        if (row < a.At.rows) {
            double s_dy, s_y;
            row_dot2<G>(a.At, row, lane, a.yR, a.yT, s_dy, s_y);
            if (lane == 0) {
                const double xt = a.xT[row];
                const double dx = a.xR[row] - xt;
                q[QAT_CROSS] += s_dy * dx;
                q[QAT_PN2] += dx * dx;
                const double dres = (a.c[row] - s_y - a.dslack[row]) * a.D[row];
                q[QAT_DRES2] += dres * dres;
                q[QAT_DRESMAX] = fmax(q[QAT_DRESMAX], fabs(dres));
                q[QAT_CX] += a.c[row] * xt;
                q[QAT_LX] += a.dslack[row] * xt;
                const double rx = xt - a.x0[row];
                q[QAT_RX2] += rx * rx;
            }
        }
    }
    block_reduce_store<kEvalAtSums, 1>(q, a.parts + (size_t)blockIdx.x * kEvalStride);
}

struct EvalAArgs {
    CsrView A;
    const double *__restrict__ xR;
    const double *__restrict__ xT;
    const double *__restrict__ yR;
    const double *__restrict__ yT;
    const double *__restrict__ y0;
    const double *__restrict__ l;
    const double *__restrict__ u;
    const double *__restrict__ E;
    double *__restrict__ parts;
};

// Rows of A: (A dx)_i with dx = xR - xT and (A xT)_i in one pass.
template <int G>
static __global__ void __launch_bounds__(kBlock) eval_a_kernel(EvalAArgs a)
{
    constexpr int RPB = kBlock / G;
    double q[QA_N] = {0, 0, 0, 0, 0, 0};
    const int lane = threadIdx.x & (G - 1);
    for (int base = blockIdx.x * RPB; base < a.A.rows; base += gridDim.x * RPB) {
        const int row = base + threadIdx.x / G;
        if (row < a.A.rows) {
            double s_dx, s_x;
            row_dot2<G>(a.A, row, lane, a.xR, a.xT, s_dx, s_x);
            if (lane == 0) {
                q[QA_AX2] += s_dx * s_dx;
                const double yt = a.yT[row];
                const double dy = a.yR[row] - yt;
                q[QA_DN2] += dy * dy;
                const double lo = a.l[row], hi = a.u[row];
                const double cl = fmax(lo, fmin(s_x, hi));
                const double pres = (s_x - cl) * a.E[row];
                q[QA_PRES2] += pres * pres;
                q[QA_PRESMAX] = fmax(q[QA_PRESMAX], fabs(pres));
                const double lf = isfinite(lo) ? lo : 0.0;
                const double uf = isfinite(hi) ? hi : 0.0;
                q[QA_CONTRIB] += fmax(yt, 0.0) * lf + fmin(yt, 0.0) * uf;
                const double ry = yt - a.y0[row];
                q[QA_RY2] += ry * ry;
            }
        }
    }
    block_reduce_store<kEvalASums, 1>(q, a.parts + (size_t)blockIdx.x * kEvalStride);
}

// Restart to the T-output: anchor = current = T (solver.cu:3056-3073).
static __global__ void __launch_bounds__(kBlock)
    restart_copy_kernel(const double *__restrict__ xT, double *__restrict__ x0, double *__restrict__ x, int n,
                        const double *__restrict__ yT, double *__restrict__ y0, double *__restrict__ y, int m)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const double v = xT[i];
        x0[i] = v;
        x[i] = v;
    } else if (i < n + m) {
        const int r = i - n;
        const double v = yT[r];
        y0[r] = v;
        y[r] = v;
    }
}

// Partials of ||A (xa - xb)||^2 (grid-stride, fixed grid): the sigma-mode
// restart's normal-motion term (solver.cu:2867-2884), the difference rounded
// first as the reference's materialised delta vector is.
template <int G>
static __global__ void __launch_bounds__(kBlock)
    delta_spmv_sumsq_kernel(CsrView A, const double *__restrict__ xa, const double *__restrict__ xb,
                            double *__restrict__ parts)
{
    constexpr int RPB = kBlock / G;
    double q[1] = {0.0};
    const int lane = threadIdx.x & (G - 1);
    for (int base = blockIdx.x * RPB; base < A.rows; base += gridDim.x * RPB) {
        const int row = base + threadIdx.x / G;
        if (row < A.rows) {
            double s_diff, s_b;
            row_dot2<G>(A, row, lane, xa, xb, s_diff, s_b);
            if (lane == 0) q[0] += s_diff * s_diff;
        }
    }
    block_reduce_store<1, 0>(q, parts + blockIdx.x);
}

// Restart-candidate score terms (compute_restart_delta_score, solver.cu:2663-2747)
// of the difference (candidate - reference) of a primal-dual pair, rounded
// first as the reference's materialised delta vectors are.  Over the rows of
// At: ||dx||^2 and <At dy, dx>; over the rows of A: ||A dx||^2 and ||dy||^2.
// Partials two per block (grid-stride, fixed grid).
template <int G>
static __global__ void __launch_bounds__(kBlock)
    score_at_kernel(CsrView At, const double *__restrict__ cy, const double *__restrict__ ry,
                    const double *__restrict__ cx, const double *__restrict__ rx, double *__restrict__ parts)
{
    constexpr int RPB = kBlock / G;
    double q[2] = {0.0, 0.0};
    const int lane = threadIdx.x & (G - 1);
    for (int base = blockIdx.x * RPB; base < At.rows; base += gridDim.x * RPB) {
        const int row = base + threadIdx.x / G;
        if (row < At.rows) {
            double s_diff, s_b;
            row_dot2<G>(At, row, lane, cy, ry, s_diff, s_b);
            if (lane == 0) {
                const double dx = cx[row] - rx[row];
                q[0] += dx * dx;
                q[1] += s_diff * dx;
            }
        }
    }
    block_reduce_store<2, 0>(q, parts + 2 * blockIdx.x);
}

template <int G>
static __global__ void __launch_bounds__(kBlock)
    score_a_kernel(CsrView A, const double *__restrict__ cx, const double *__restrict__ rx,
                   const double *__restrict__ cy, const double *__restrict__ ry, double *__restrict__ parts)
{
    constexpr int RPB = kBlock / G;
    double q[2] = {0.0, 0.0};
    const int lane = threadIdx.x & (G - 1);
    for (int base = blockIdx.x * RPB; base < A.rows; base += gridDim.x * RPB) {
        const int row = base + threadIdx.x / G;
        if (row < A.rows) {
            double s_diff, s_b;
            row_dot2<G>(A, row, lane, cx, rx, s_diff, s_b);
            if (lane == 0) {
                const double dy = cy[row] - ry[row];
                q[0] += s_diff * s_diff;
                q[1] += dy * dy;
            }
        }
    }
    block_reduce_store<2, 0>(q, parts + 2 * blockIdx.x);
}

// ||a - b||^2 over n entries and ||c - d||^2 over m entries: the restart
// distances from the anchor when the restart point is not the T output.
static __global__ void __launch_bounds__(kBlock)
    dist2_kernel(const double *__restrict__ a, const double *__restrict__ b, int n, const double *__restrict__ c,
                 const double *__restrict__ d, int m, double *__restrict__ parts)
{
    double q[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double t = a[i] - b[i];
        q[0] += t * t;
    }
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) {
        const double t = c[i] - d[i];
        q[1] += t * t;
    }
    block_reduce_store<2, 0>(q, parts + 2 * blockIdx.x);
}

// ---------------------------------------------------------------------------
// Power iteration for sigma_max(A)^2 on A A^T (utils.cu:87-228), device
// resident: every kernel returns at once after convergence or max_iter, so a
// captured chunk of iterations can be replayed until the host sees `done`.

struct PowerState {
    double vnorm;     // ||v|| of the current (unnormalised) iterate
    double sigma_sq;  // latest <v_hat, A A^T v_hat>
    double resid;     // ||A A^T v_hat - sigma_sq v_hat||
    int iter;         // completed iterations
    int done;         // residual < tol seen
};

__device__ __forceinline__ bool power_stopped(const PowerState *ps, int max_iter)
{
    return ps->done != 0 || ps->iter >= max_iter;
}

static __global__ void __launch_bounds__(kBlock)
    power_normalize_kernel(const double *__restrict__ v, double *__restrict__ vh, int m,
                           const PowerState *__restrict__ ps, int max_iter)
{
    if (power_stopped(ps, max_iter)) return;
    const double inv = 1.0 / ps->vnorm;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < m) vh[i] = v[i] * inv;
}

template <int G>
static __global__ void __launch_bounds__(kBlock)
    spmv_gated_kernel(CsrView M, const double *__restrict__ x, double *__restrict__ y,
                      const PowerState *__restrict__ ps, int max_iter)
{
    if (power_stopped(ps, max_iter)) return;
    const int tid = blockIdx.x * kBlock + threadIdx.x;
    const int row = tid / G;
    const int lane = tid & (G - 1);
    if (row >= M.rows) return;
    const double s = row_dot<G>(M, row, lane, x);
    if (lane == 0) y[row] = s;
}

// v = A w, partial <v_hat, v> per block (grid-stride, fixed grid).
template <int G>
static __global__ void __launch_bounds__(kBlock)
    power_av_dot_kernel(CsrView A, const double *__restrict__ w, const double *__restrict__ vh,
                        double *__restrict__ v, double *__restrict__ parts, const PowerState *__restrict__ ps,
                        int max_iter)
{
    if (power_stopped(ps, max_iter)) return;
    constexpr int RPB = kBlock / G;
    double q[1] = {0.0};
    const int lane = threadIdx.x & (G - 1);
    for (int base = blockIdx.x * RPB; base < A.rows; base += gridDim.x * RPB) {
        const int row = base + threadIdx.x / G;
        if (row < A.rows) {
            const double s = row_dot<G>(A, row, lane, w);
            if (lane == 0) {
                v[row] = s;
                q[0] += vh[row] * s;
            }
        }
    }
    block_reduce_store<1, 0>(q, parts + blockIdx.x);
}

// r = -sigma_sq * v_hat + v; partials of ||r||^2 and ||v||^2.
static __global__ void __launch_bounds__(kBlock)
    power_resid_kernel(const double *__restrict__ v, const double *__restrict__ vh, int m,
                       double *__restrict__ parts, const PowerState *__restrict__ ps, int max_iter)
{
    if (power_stopped(ps, max_iter)) return;
    const double neg = -ps->sigma_sq;
    double q[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) {
        const double vi = v[i];
        const double r = neg * vh[i] + vi;
        q[0] += r * r;
        q[1] += vi * vi;
    }
    block_reduce_store<2, 0>(q, parts + (size_t)blockIdx.x * 2);
}

// Single block: finish sigma_sq.
static __global__ void __launch_bounds__(kBlock)
    power_finish_dot_kernel(const double *__restrict__ parts, int nparts, PowerState *__restrict__ ps,
                            int max_iter)
{
    if (power_stopped(ps, max_iter)) return;
    __shared__ double red[kWavesPerBlock];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nparts; b += kBlock) acc += parts[b];
    acc = wave_sum(acc);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int w = 1; w < kWavesPerBlock; ++w) r += red[w];
        ps->sigma_sq = r;
    }
}

// Single block: finish ||r|| and ||v||, test convergence, count the iteration.
static __global__ void __launch_bounds__(kBlock)
    power_finish_resid_kernel(const double *__restrict__ parts, int nparts, PowerState *__restrict__ ps,
                              int max_iter, double tol)
{
    if (power_stopped(ps, max_iter)) return;
    __shared__ double red[2][kWavesPerBlock];
    double a = 0.0, b = 0.0;
    for (int p = threadIdx.x; p < nparts; p += kBlock) {
        a += parts[(size_t)p * 2];
        b += parts[(size_t)p * 2 + 1];
    }
    a = wave_sum(a);
    b = wave_sum(b);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        red[0][threadIdx.x / kWave] = a;
        red[1][threadIdx.x / kWave] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ra = red[0][0], rb = red[1][0];
        for (int w = 1; w < kWavesPerBlock; ++w) {
            ra += red[0][w];
            rb += red[1][w];
        }
        ps->resid = sqrt(ra);
        ps->vnorm = sqrt(rb);
        ps->iter += 1;
        if (ps->resid < tol) ps->done = 1;
    }
}

static __global__ void __launch_bounds__(kBlock)
    sumsq_partials_kernel(const double *__restrict__ v, int len, double *__restrict__ parts)
{
    double q[1] = {0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < len; i += gridDim.x * kBlock) q[0] += v[i] * v[i];
    block_reduce_store<1, 0>(q, parts + blockIdx.x);
}

// ---------------------------------------------------------------------------
// Rescaling (LP/CP_AL/src/preconditioner.cu:477-736), one thread per row.

static __global__ void __launch_bounds__(kBlock) row_absmax_kernel(CsrView M, double *__restrict__ out)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= M.rows) return;
    double mx = 0.0;
    for (int k = M.rp[i]; k < M.rp[i + 1]; ++k) {
        double v = fabs(M.v[k]);
        if (!isfinite(v)) v = 0.0;
        if (v > mx) mx = v;
    }
    out[i] = mx;
}

__device__ __forceinline__ double pow_fast(double v, double p)
{
    if (p == 2.0) return v * v;
    if (p == 1.0) return v;
    if (p == 0.5) return sqrt(v);
    return pow(v, p);
}

static __global__ void __launch_bounds__(kBlock) row_powsum_kernel(CsrView M, double degree, double *__restrict__ out)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= M.rows) return;
    double acc = 0.0;
    for (int k = M.rp[i]; k < M.rp[i + 1]; ++k) {
        double v = fabs(M.v[k]);
        if (!isfinite(v)) v = 0.0;
        acc += pow_fast(v, degree);
    }
    out[i] = acc;
}

static __global__ void __launch_bounds__(kBlock)
    clamp_sqrt_accum_kernel(double *__restrict__ sf, double *__restrict__ inv, double *__restrict__ cum, int len)
{
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= len) return;
    const double v = sf[t];
    const double s = (v < 1e-12) ? 1.0 : sqrt(v);
    cum[t] *= s;
    sf[t] = s;
    inv[t] = 1.0 / s;
}

static __global__ void __launch_bounds__(kBlock)
    cr_log_update_kernel(CsrView M, const double *__restrict__ other, double *__restrict__ log_scale,
                         double *__restrict__ log_delta)
{   // log_delta (may be null): this pass's change of each log scale
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= M.rows) return;
    const int b = M.rp[r], e = M.rp[r + 1];
    if (e - b <= 0) {
        if (log_delta) log_delta[r] = -log_scale[r];
        log_scale[r] = 0.0;
        return;
    }
    double acc = 0.0;
    for (int k = b; k < e; ++k) {
        double v = fabs(M.v[k]);
        if (!isfinite(v) || v < 1e-300) v = 1e-300;
        acc += -log(v) - other[M.ci[k]];
    }
    const double next = acc / (double)(e - b);
    if (log_delta) log_delta[r] = next - log_scale[r];
    log_scale[r] = next;
}

// Largest change of a log edge weight, |row change + column change|, over the
// nonzeros of M in one Curtis-Reid pass: per-block maxima into parts[block]
// (preconditioner.cu:63-80 and the max reduction at :222-235).
static __global__ void __launch_bounds__(kBlock)
    cr_edge_change_kernel(CsrView M, const double *__restrict__ row_delta, const double *__restrict__ col_delta,
                          double *__restrict__ parts)
{
    double q[1] = {0.0};
    for (int r = blockIdx.x * kBlock + threadIdx.x; r < M.rows; r += gridDim.x * kBlock) {
        const double dr = row_delta[r];
        for (int k = M.rp[r]; k < M.rp[r + 1]; ++k) {
            const double ch = fabs(dr + col_delta[M.ci[k]]);
            if (isfinite(ch) && ch > q[0]) q[0] = ch;
        }
    }
    block_reduce_store<0, 1>(q, parts + blockIdx.x);
}

static __global__ void __launch_bounds__(kBlock)
    cr_exp_kernel(double *__restrict__ ls, double *__restrict__ inv, double *__restrict__ cum, int len)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= len) return;
    double f = exp(ls[i]);
    f = fmin(fmax(f, 1e-30), 1e30);
    const double s = 1.0 / f;
    ls[i] = s;
    inv[i] = f;
    cum[i] *= s;
}

// c *= 1/d, bounds *= d, x0 *= d  (scale_variables_kernel)
static __global__ void __launch_bounds__(kBlock)
    scale_vars_kernel(double *__restrict__ c, double *__restrict__ lb, double *__restrict__ ub,
                      double *__restrict__ x0, const double *__restrict__ d, const double *__restrict__ inv_d, int n)
{
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const double dj = d[j];
    c[j] *= inv_d[j];
    lb[j] *= dj;
    ub[j] *= dj;
    x0[j] *= dj;
}

// bounds *= 1/e, y0 *= e  (scale_constraints_kernel)
static __global__ void __launch_bounds__(kBlock)
    scale_cons_kernel(double *__restrict__ l, double *__restrict__ u, double *__restrict__ y0,
                      const double *__restrict__ e, const double *__restrict__ inv_e, int m)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= m) return;
    const double ie = inv_e[i];
    l[i] *= ie;
    u[i] *= ie;
    y0[i] *= e[i];
}

// v[k] *= inv_row[row] * inv_col[col]: the product of the two inverse factors is
// commutative, so A and At receive bit-identical scale factors without a
// transpose map (scale_csr_nnz_kernel, preconditioner.cu:513-531).
static __global__ void __launch_bounds__(kBlock)
    scale_rows_kernel(const int *__restrict__ rp, const int *__restrict__ ci, double *__restrict__ v, int rows,
                      const double *__restrict__ inv_row, const double *__restrict__ inv_col)
{
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= rows) return;
    const double ir = inv_row[r];
    for (int k = rp[r]; k < rp[r + 1]; ++k) v[k] *= inv_col[ci[k]] * ir;
}

static __global__ void __launch_bounds__(kBlock)
    bound_contrib_partials_kernel(const double *__restrict__ l, const double *__restrict__ u, int m,
                                  const double *__restrict__ c, int n, double *__restrict__ parts)
{   // compute_bound_contrib_kernel (:699-729) and ||c||^2 in one pass
    double q[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) {
        const double L = l[i], U = u[i];
        const bool fL = isfinite(L), fU = isfinite(U);
        double acc = 0.0;
        if (fL && (!fU || fabs(L - U) > 1e-12)) acc += L * L;
        if (fU) acc += U * U;
        q[0] += acc;
    }
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) q[1] += c[j] * c[j];
    block_reduce_store<2, 0>(q, parts + (size_t)blockIdx.x * 2);
}

static __global__ void __launch_bounds__(kBlock)
    bound_obj_scale_kernel(double *__restrict__ l, double *__restrict__ u, double *__restrict__ y0, int m,
                           double *__restrict__ lb, double *__restrict__ ub, double *__restrict__ c,
                           double *__restrict__ x0, int n, double cs, double os)
{   // scale_bounds_kernel + scale_objective_kernel
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < m) {
        l[i] *= cs;
        u[i] *= cs;
        y0[i] *= os;
    } else if (i < m + n) {
        const int j = i - m;
        lb[j] *= cs;
        ub[j] *= cs;
        c[j] *= os;
        x0[j] *= cs;
    }
}

// Finite entries -> 0, infinite ones kept: the dual polishing problem's
// bounds (zero_finite_value_vectors_kernel, solver.cu:4199-4207).
static __global__ void __launch_bounds__(kBlock) zero_finite_kernel(double *__restrict__ v, int len)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < len && isfinite(v[i])) v[i] = 0.0;
}

static __global__ void __launch_bounds__(kBlock) fill_kernel(double *__restrict__ x, double v, int len)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < len) x[i] = v;
}

// Result: reduced cost with infinite-bound clamps, then unscale x and y
// (compute_and_rescale_reduced_cost_kernel :1768-1793, rescale_solution_kernel :2481-2500).
template <int G>
static __global__ void __launch_bounds__(kBlock)
    reduced_cost_kernel(CsrView At, const double *__restrict__ yT, const double *__restrict__ c,
                        const double *__restrict__ D, const double *__restrict__ lb, const double *__restrict__ ub,
                        double c_scale, double *__restrict__ rc)
{
    const int tid = blockIdx.x * kBlock + threadIdx.x;
    const int row = tid / G;
    const int lane = tid & (G - 1);
    if (row >= At.rows) return;
    const double s = row_dot<G>(At, row, lane, yT);
    if (lane != 0) return;
    double r = (c[row] - s) * D[row] / c_scale;
    if (!isfinite(lb[row])) r = fmin(r, 0.0);
    if (!isfinite(ub[row])) r = fmax(r, 0.0);
    rc[row] = r;
}

static __global__ void __launch_bounds__(kBlock)
    unscale_kernel(const double *__restrict__ xT, const double *__restrict__ D, double b_scale,
                   double *__restrict__ xo, int n, const double *__restrict__ yT, const double *__restrict__ E,
                   double c_scale, double *__restrict__ yo, int m)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) xo[i] = xT[i] / D[i] / b_scale;
    else if (i < n + m) yo[i - n] = yT[i - n] / E[i - n] / c_scale;
}

}  // namespace lp
}  // namespace pdal
