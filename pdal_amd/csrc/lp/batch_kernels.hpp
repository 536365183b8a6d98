// gfx950 kernels of the batched CP_AL solver: many independent LPs solved
// together as one block-diagonal system.
//
// Layout in HBM: the instances' A matrices are concatenated into one CSR with
// column offsets (block diagonal, rows of instance i contiguous), At is its
// stable transpose (so also block diagonal), and every vector is the
// concatenation of the instances' vectors.  A block table maps each 256-thread
// workgroup to exactly one instance and a run of that instance's rows, so a
// workgroup reads its instance's step sizes once, skips the whole run when the
// instance has stopped, and writes one partial per evaluation quantity that a
// per-instance segmented reduction combines in a fixed order.
//
// Per-instance control (fixed-point error, KKT residual, termination,
// adaptive restart with the PID primal weight and the step schedule; the
// single-instance logic of LP/CP_AL/src/solver.cu:1197-1736, 2821-3126,
// 3611-3686 and utils.cu:275-359, 812-917) runs on the device, one thread per
// instance, so a whole block of iterations plus its evaluation and restart is
// one hipGraph replay with no host round trip.
#pragma once

#include "cp_al_types.h"

#include "../common/device_reduce.hpp"
#include "../common/spmv.hpp"

namespace pdal {
namespace lp {
namespace batch {

constexpr int kEvalSlots = 16;  // per-instance evaluation quantities (13 used)
enum EvalSlot {
    E_CROSS, E_PN2, E_DRES2, E_CX, E_LX, E_RX2, E_DRESMAX,  // rows of At (6 sums, 1 max)
    E_AX2 = 8, E_DN2, E_PRES2, E_CONTRIB, E_RY2, E_PRESMAX  // rows of A (5 sums, 1 max)
};
constexpr int kAtSums = 6, kAtQ = 7, kASums = 5, kAQ = 6;
constexpr int kPartStride = 8;
// Rows per lane group per workgroup: a table entry covers (256/G) * kGroupRows
// rows.  Measured on the 4096-instance family: 1 is fastest (2.24 ms for the
// At.y pass against 2.62 ms at 4) -- the pass is bound by random-gather
// throughput, and a lane group walking several rows in sequence only lowers
// the number of gathers in flight.
constexpr int kGroupRows = 1;

// Workgroup -> (instance, first row) for a row-parallel kernel over A or At;
// the workgroup covers (256/G) * kGroupRows rows from there.
//
// With xcd set, workgroup b takes table entry block(): the dispatcher deals
// workgroups to the 8 XCDs in turn (b mod 8), so the entries are cut into 8
// contiguous ranges and XCD x walks range x.  An instance's rows then run on
// one or two XCDs and its gathered vector is filled into those L2 caches
// only, not into all eight.  Entry ranges and partial-sum slots stay in table
// order, so per-instance reductions are unchanged.
struct BlockTable {
    const int *__restrict__ inst;
    const int *__restrict__ row0;
    const int *__restrict__ row_end;  // per instance: one past its last row
    int xcd;

    __device__ __forceinline__ int block() const
    {
        if (!xcd) return blockIdx.x;
        const int nb = gridDim.x, per = nb >> 3, rem = nb & 7;
        const int x = blockIdx.x & 7, k = blockIdx.x >> 3;  // range x holds per + (x < rem) entries
        return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + k;
    }
};

// Device-resident state of one instance.
struct InstState {
    double tau, sigma, beta;           // step schedule read by the iteration kernels
    double step, pw, best_pw, pw_err_sum, pw_last_err, best_gap, lambda_est, al_lambda;
    double fp, fp0, last_trial_fp;
    double abs_p, rel_p, abs_d, rel_d, p_obj, d_obj, gap, rel_gap;
    double b_scale, c_scale, c_norm, b_norm, obj_const;
    int inner, total, charged, reason;
    int active, restart, pad0, pad1;
};

// Run-wide constants of the control logic (environment overrides resolved on the host).
struct ControlParams {
    double step_margin;               // CP_AL_STEP_MARGIN (0.998^2)
    double suff, nec, art;            // restart thresholds
    double k_p, k_i, k_d, i_smooth;   // PID
    double eps_feas, eps_opt;
    int iteration_limit, K, linf, pad;
};

struct PrimalArgs {
    CsrView At;
    BlockTable bt;
    const double *__restrict__ yv;
    const double *__restrict__ c, *__restrict__ lb, *__restrict__ ub, *__restrict__ x0;
    double *__restrict__ x, *__restrict__ xR, *__restrict__ xT, *__restrict__ xH, *__restrict__ dslack;
    const InstState *__restrict__ st;
    double rho;
};

// At.y fused with the primal update (solver.cu:2340-2397); one instance per workgroup.
template <int G, bool MAJOR>
static __global__ void __launch_bounds__(kBlock) primal_kernel(PrimalArgs a, int k_offset)
{
    const int tb = a.bt.block();
    const int inst = a.bt.inst[tb];
    const InstState &s = a.st[inst];
    if (!s.active) return;
    const int lane = threadIdx.x & (G - 1);
    const int end = a.bt.row_end[inst];
    const int base = a.bt.row0[tb] + threadIdx.x / G;
    const double tau = s.tau;
    const int k = s.inner + k_offset;
    const double w = (double)k / (double)(k + 1);
    for (int u = 0; u < kGroupRows; ++u) {
    const int row = base + u * (kBlock / G);
    if (row >= end) break;
    const double sdot = row_dot<G>(a.At, row, lane, a.yv);
    if (lane != 0) continue;
    const double xj = a.x[row];
    const double temp = xj - tau * (a.c[row] - sdot);
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
}

struct DualArgs {
    CsrView A;
    BlockTable bt;
    const double *__restrict__ xR;
    const double *__restrict__ l, *__restrict__ u, *__restrict__ y0;
    double *__restrict__ y, *__restrict__ yT, *__restrict__ yH, *__restrict__ yR;
    const InstState *__restrict__ st;
    double rho;
};

// A.x_bar fused with the dual update (solver.cu:2399-2457).
template <int G, bool MAJOR>
static __global__ void __launch_bounds__(kBlock) dual_kernel(DualArgs a, int k_offset)
{
    const int tb = a.bt.block();
    const int inst = a.bt.inst[tb];
    const InstState &s = a.st[inst];
    if (!s.active) return;
    const int lane = threadIdx.x & (G - 1);
    const int end = a.bt.row_end[inst];
    const int base = a.bt.row0[tb] + threadIdx.x / G;
    const double sigma = s.sigma;
    const double prox = sigma + s.beta;
    const int k = s.inner + k_offset;
    const double w = (double)k / (double)(k + 1);
    for (int u = 0; u < kGroupRows; ++u) {
    const int row = base + u * (kBlock / G);
    if (row >= end) break;
    const double pp = row_dot<G>(a.A, row, lane, a.xR);
    if (lane != 0) continue;
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
}

// ---------------------------------------------------------------------------
// LDS-staged iteration kernels: one 1024-thread workgroup per instance.  The
// instance's segment of the gathered vector (y for At rows, x_bar for A rows)
// is copied into LDS with coalesced loads, then every row of the instance
// gathers from LDS instead of issuing one random global load per nonzero.
// gfx950 gives a workgroup up to 160 KB of LDS: 20480 doubles.  The matrix,
// the row-wise vectors and the epilogue are those of the table kernels.

constexpr int kStagedBlock = 1024;

struct StagedList {
    const int *__restrict__ inst;      // instance per workgroup
    const long *__restrict__ row_off;  // N+1: the instance's rows of the walked matrix
    const long *__restrict__ seg_off;  // N+1: its segment of the gathered vector
};

// Two rows of one lane group at a time (their loads interleave: more memory
// requests in flight per wave).  Each row is still summed in lane-strided
// order over [b, e) and folded by the G-lane butterfly, as in row_dot.
template <int G>
__device__ __forceinline__ void range_dot2(const CsrView &M, int b1, int e1, int b2, int e2, int lane,
                                           const double *seg, int base1, int base2, double &s1, double &s2)
{
    double a = 0.0, c = 0.0;
    int k1 = b1 + lane, k2 = b2 + lane;
    for (; k1 < e1 && k2 < e2; k1 += G, k2 += G) {
        const double v1 = M.v[k1], v2 = M.v[k2];
        const int c1 = M.ci[k1], c2 = M.ci[k2];
        a += v1 * seg[c1 - base1];
        c += v2 * seg[c2 - base2];
    }
    for (; k1 < e1; k1 += G) a += M.v[k1] * seg[M.ci[k1] - base1];
    for (; k2 < e2; k2 += G) c += M.v[k2] * seg[M.ci[k2] - base2];
    s1 = group_sum<G>(a);
    s2 = group_sum<G>(c);
}

__device__ __forceinline__ int stage_segment(const StagedList &sl, int inst, const double *__restrict__ src,
                                             double *seg)
{
    const int g0 = (int)sl.seg_off[inst], glen = (int)(sl.seg_off[inst + 1] - g0);
    for (int t = threadIdx.x; t < glen; t += kStagedBlock) seg[t] = src[g0 + t];
    __syncthreads();
    return g0;
}

// The fused epilogues of primal_kernel and dual_kernel for one row.
template <bool MAJOR>
__device__ __forceinline__ void primal_row_update(const PrimalArgs &a, int row, double sdot, double tau, double w)
{
    const double xj = a.x[row];
    const double temp = xj - tau * (a.c[row] - sdot);
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

template <bool MAJOR>
__device__ __forceinline__ void dual_row_update(const DualArgs &a, int row, double pp, double sigma, double prox,
                                                double w)
{
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
static __global__ void __launch_bounds__(kStagedBlock) primal_staged_kernel(PrimalArgs a, StagedList sl, int k_offset)
{
    extern __shared__ double seg[];
    const int inst = sl.inst[blockIdx.x];
    const InstState &s = a.st[inst];
    if (!s.active) return;
    const int g0 = stage_segment(sl, inst, a.yv, seg);
    const int lane = threadIdx.x & (G - 1);
    const int end = (int)sl.row_off[inst + 1];
    const double tau = s.tau;
    const int k = s.inner + k_offset;
    const double w = (double)k / (double)(k + 1);
    constexpr int kStride = kStagedBlock / G;
    for (int row = (int)sl.row_off[inst] + threadIdx.x / G; row < end; row += 2 * kStride) {
        const int row2 = row + kStride;
        const bool two = row2 < end;
        double d1, d2;
        range_dot2<G>(a.At, a.At.rp[row], a.At.rp[row + 1], two ? a.At.rp[row2] : 0, two ? a.At.rp[row2 + 1] : 0,
                      lane, seg, g0, g0, d1, d2);
        if (lane == 0) {
            primal_row_update<MAJOR>(a, row, d1, tau, w);
            if (two) primal_row_update<MAJOR>(a, row2, d2, tau, w);
        }
    }
}

template <int G, bool MAJOR>
static __global__ void __launch_bounds__(kStagedBlock) dual_staged_kernel(DualArgs a, StagedList sl, int k_offset)
{
    extern __shared__ double seg[];
    const int inst = sl.inst[blockIdx.x];
    const InstState &s = a.st[inst];
    if (!s.active) return;
    const int g0 = stage_segment(sl, inst, a.xR, seg);
    const int lane = threadIdx.x & (G - 1);
    const int end = (int)sl.row_off[inst + 1];
    const double sigma = s.sigma;
    const double prox = sigma + s.beta;
    const int k = s.inner + k_offset;
    const double w = (double)k / (double)(k + 1);
    constexpr int kStride = kStagedBlock / G;
    for (int row = (int)sl.row_off[inst] + threadIdx.x / G; row < end; row += 2 * kStride) {
        const int row2 = row + kStride;
        const bool two = row2 < end;
        double d1, d2;
        range_dot2<G>(a.A, a.A.rp[row], a.A.rp[row + 1], two ? a.A.rp[row2] : 0, two ? a.A.rp[row2 + 1] : 0, lane,
                      seg, g0, g0, d1, d2);
        if (lane == 0) {
            dual_row_update<MAJOR>(a, row, d1, sigma, prox, w);
            if (two) dual_row_update<MAJOR>(a, row2, d2, sigma, prox, w);
        }
    }
}

// A.x_bar for an instance whose x_bar segment fills up to two LDS loads
// (gfx950: n up to 40960).  Its rows hold ascending column indices, and
// split[row] is the first nonzero of the second column half.  Pass 1 stages
// the first half and stores every row's partial sum in part[row]; pass 2
// stages the second half, adds its partial and runs the dual epilogue.  A row
// is summed as (first-half sum) + (second-half sum), each in lane-strided
// order: the same terms as the table kernel in a different association.
template <int G, bool MAJOR>
static __global__ void __launch_bounds__(kStagedBlock)
    dual_split_kernel(DualArgs a, StagedList sl, const int *__restrict__ split, double *__restrict__ part,
                      int k_offset)
{
    extern __shared__ double seg[];
    const int inst = sl.inst[blockIdx.x];
    const InstState &s = a.st[inst];
    if (!s.active) return;
    const int c0 = (int)sl.seg_off[inst], n = (int)(sl.seg_off[inst + 1] - c0), h = (n + 1) / 2;
    const int lane = threadIdx.x & (G - 1);
    const int r0 = (int)sl.row_off[inst], end = (int)sl.row_off[inst + 1];
    constexpr int kStride = kStagedBlock / G;
    for (int t = threadIdx.x; t < h; t += kStagedBlock) seg[t] = a.xR[c0 + t];
    __syncthreads();
    for (int row = r0 + threadIdx.x / G; row < end; row += 2 * kStride) {
        const int row2 = row + kStride;
        const bool two = row2 < end;
        double p1, p2;
        range_dot2<G>(a.A, a.A.rp[row], split[row], two ? a.A.rp[row2] : 0, two ? split[row2] : 0, lane, seg, c0,
                      c0, p1, p2);
        if (lane == 0) {
            part[row] = p1;
            if (two) part[row2] = p2;
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n - h; t += kStagedBlock) seg[t] = a.xR[c0 + h + t];
    __syncthreads();
    const double sigma = s.sigma;
    const double prox = sigma + s.beta;
    const int k = s.inner + k_offset;
    const double w = (double)k / (double)(k + 1);
    for (int row = r0 + threadIdx.x / G; row < end; row += 2 * kStride) {
        const int row2 = row + kStride;
        const bool two = row2 < end;
        double q1, q2;
        range_dot2<G>(a.A, split[row], a.A.rp[row + 1], two ? split[row2] : 0, two ? a.A.rp[row2 + 1] : 0, lane,
                      seg, c0 + h, c0 + h, q1, q2);
        if (lane == 0) {
            dual_row_update<MAJOR>(a, row, part[row] + q1, sigma, prox, w);
            if (two) dual_row_update<MAJOR>(a, row2, part[row2] + q2, sigma, prox, w);
        }
    }
}

struct PredictorArgs {
    CsrView A;
    BlockTable bt;
    const double *__restrict__ x, *__restrict__ y, *__restrict__ l, *__restrict__ u;
    double *__restrict__ out;
    const InstState *__restrict__ st;
};

// Box-AL dual predictor (solver.cu:2459-2479), launched when beta > 0.
template <int G>
static __global__ void __launch_bounds__(kBlock) predictor_kernel(PredictorArgs a)
{
    const int tb = a.bt.block();
    const int inst = a.bt.inst[tb];
    const InstState &s = a.st[inst];
    if (!s.active) return;
    const int lane = threadIdx.x & (G - 1);
    const int end = a.bt.row_end[inst];
    const int base = a.bt.row0[tb] + threadIdx.x / G;
    const double beta = s.beta;
    for (int u = 0; u < kGroupRows; ++u) {
        const int row = base + u * (kBlock / G);
        if (row >= end) break;
        const double pp = row_dot<G>(a.A, row, lane, a.x);
        if (lane != 0) continue;
        const double yi = a.y[row];
        if (beta <= 0.0) {
            a.out[row] = yi;
            continue;
        }
        const double z = fmax(a.l[row], fmin(pp - yi / beta, a.u[row]));
        a.out[row] = yi - beta * (pp - z);
    }
}

// ---------------------------------------------------------------------------
// Evaluation: one partial row (kPartStride doubles) per workgroup.

struct EvalAtArgs {
    CsrView At;
    BlockTable bt;
    const double *__restrict__ yR, *__restrict__ yT, *__restrict__ xR, *__restrict__ xT, *__restrict__ x0;
    const double *__restrict__ c, *__restrict__ dslack, *__restrict__ D;
    const InstState *__restrict__ st;
    double *__restrict__ parts;
};

template <int G>
static __global__ void __launch_bounds__(kBlock) eval_at_kernel(EvalAtArgs a)
{
    const int tb = a.bt.block();
    const int inst = a.bt.inst[tb];
    if (!a.st[inst].active) return;
    double q[kAtQ] = {0, 0, 0, 0, 0, 0, 0};
    const int lane = threadIdx.x & (G - 1);
    const int end = a.bt.row_end[inst];
    const int base = a.bt.row0[tb] + threadIdx.x / G;
    for (int u = 0; u < kGroupRows; ++u) {
        const int row = base + u * (kBlock / G);
        if (row >= end) break;
        double s_dy, s_y;
        row_dot2<G>(a.At, row, lane, a.yR, a.yT, s_dy, s_y);
        if (lane == 0) {
            const double xt = a.xT[row];
            const double dx = a.xR[row] - xt;
            q[E_CROSS] += s_dy * dx;
            q[E_PN2] += dx * dx;
            const double dres = (a.c[row] - s_y - a.dslack[row]) * a.D[row];
            q[E_DRES2] += dres * dres;
            q[6] = fmax(q[6], fabs(dres));
            q[E_CX] += a.c[row] * xt;
            q[E_LX] += a.dslack[row] * xt;
            const double rx = xt - a.x0[row];
            q[E_RX2] += rx * rx;
        }
    }
    block_reduce_store<kAtSums, 1>(q, a.parts + (size_t)tb * kPartStride);
}

struct EvalAArgs {
    CsrView A;
    BlockTable bt;
    const double *__restrict__ xR, *__restrict__ xT, *__restrict__ yR, *__restrict__ yT, *__restrict__ y0;
    const double *__restrict__ l, *__restrict__ u, *__restrict__ E;
    const InstState *__restrict__ st;
    double *__restrict__ parts;
};

template <int G>
static __global__ void __launch_bounds__(kBlock) eval_a_kernel(EvalAArgs a)
{
    const int tb = a.bt.block();
    const int inst = a.bt.inst[tb];
    if (!a.st[inst].active) return;
    double q[kAQ] = {0, 0, 0, 0, 0, 0};
    const int lane = threadIdx.x & (G - 1);
    const int end = a.bt.row_end[inst];
    const int base = a.bt.row0[tb] + threadIdx.x / G;
    for (int u = 0; u < kGroupRows; ++u) {
        const int row = base + u * (kBlock / G);
        if (row >= end) break;
        double s_dx, s_x;
        row_dot2<G>(a.A, row, lane, a.xR, a.xT, s_dx, s_x);
        if (lane == 0) {
            q[0] += s_dx * s_dx;
            const double yt = a.yT[row];
            const double dy = a.yR[row] - yt;
            q[1] += dy * dy;
            const double lo = a.l[row], hi = a.u[row];
            const double cl = fmax(lo, fmin(s_x, hi));
            const double pres = (s_x - cl) * a.E[row];
            q[2] += pres * pres;
            q[5] = fmax(q[5], fabs(pres));
            const double lf = isfinite(lo) ? lo : 0.0;
            const double uf = isfinite(hi) ? hi : 0.0;
            q[3] += fmax(yt, 0.0) * lf + fmin(yt, 0.0) * uf;
            const double ry = yt - a.y0[row];
            q[4] += ry * ry;
        }
    }
    block_reduce_store<kASums, 1>(q, a.parts + (size_t)tb * kPartStride);
}

// One workgroup per instance: combine its workgroups' partial rows in index
// order into evalq[inst * kEvalSlots + slot0 ..].
static __global__ void __launch_bounds__(kBlock)
    seg_reduce_kernel(const double *__restrict__ parts, const int *__restrict__ first, const int *__restrict__ count,
                      int nsum, int nq, int slot0, double *__restrict__ evalq, const InstState *__restrict__ st)
{
    const int inst = blockIdx.x;
    if (!st[inst].active) return;
    __shared__ double red[kWavesPerBlock];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int b0 = first[inst], nb = count[inst];
    for (int k = 0; k < nq; ++k) {
        const bool is_sum = k < nsum;
        double acc = 0.0;
        for (int b = threadIdx.x; b < nb; b += kBlock) {
            const double v = parts[(size_t)(b0 + b) * kPartStride + k];
            acc = is_sum ? acc + v : fmax(acc, v);
        }
        acc = is_sum ? wave_sum(acc) : wave_max(acc);
        if (lane == 0) red[wave] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = red[0];
            for (int w = 1; w < kWavesPerBlock; ++w) r = is_sum ? r + red[w] : fmax(r, red[w]);
            evalq[(size_t)inst * kEvalSlots + slot0 + k] = r;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Per-instance control, one thread per instance.

__device__ __forceinline__ void refresh_step_schedule(InstState &s, double margin)
{   // solver.cu:3103-3126 (alpha mode)
    s.al_lambda = fmin(fmax(s.al_lambda, 0.0), 0.95);
    const double eta_sq = margin * (1.0 - s.al_lambda) / fmax(s.lambda_est, 1e-12);
    s.step = sqrt(fmax(eta_sq, 1e-24));
    const double dual_step = s.step * fmax(s.pw, 1e-12);
    s.beta = dual_step * s.al_lambda / fmax(1.0 - s.al_lambda, 1e-12);
    s.tau = s.step / s.pw;
    s.sigma = s.step * s.pw;
}

__device__ __forceinline__ double fixed_point_error(const InstState &s, const double *q)
{   // solver.cu:3658-3686 (alpha-mode branch)
    const double dn = sqrt(q[E_DN2]), pn = sqrt(q[E_PN2]), an = sqrt(q[E_AX2]);
    const double movement = pn * pn * s.pw + dn * dn / s.pw;
    const double interaction = 2 * s.step * q[E_CROSS];
    const double aug = s.step * s.beta * an * an;
    return sqrt(fmax(movement + interaction - aug, 0.0));
}

static __global__ void __launch_bounds__(kBlock)
    control_kernel(InstState *__restrict__ st, const double *__restrict__ evalq, int count, ControlParams cp)
{
    const int inst = blockIdx.x * kBlock + threadIdx.x;
    if (inst >= count) return;
    InstState s = st[inst];
    if (!s.active) return;
    const double *q = evalq + (size_t)inst * kEvalSlots;
    s.fp = fixed_point_error(s, q);
    // compute_residual, utils.cu:812-917
    if (cp.linf) {
        s.abs_p = q[E_PRESMAX];
        s.abs_d = q[E_DRESMAX];
    } else {
        s.abs_p = sqrt(q[E_PRES2]) / s.b_scale;
        s.abs_d = sqrt(q[E_DRES2]);
    }
    s.abs_d /= s.c_scale;
    const double bc = s.b_scale * s.c_scale;
    s.p_obj = q[E_CX] / bc + s.obj_const;
    s.d_obj = (q[E_LX] + q[E_CONTRIB]) / bc + s.obj_const;
    s.rel_p = s.abs_p / (1.0 + s.b_norm);
    s.rel_d = s.abs_d / (1.0 + s.c_norm);
    s.gap = fabs(s.p_obj - s.d_obj);
    s.rel_gap = s.gap / (1.0 + fabs(s.p_obj) + fabs(s.d_obj));
    s.inner += cp.K;
    s.total += cp.K;
    s.charged += cp.K;
    // check_termination_criteria, utils.cu:299-317
    int reason = TERMINATION_REASON_UNSPECIFIED;
    if (s.rel_d < cp.eps_feas && s.rel_p < cp.eps_feas && s.rel_gap < cp.eps_opt)
        reason = TERMINATION_REASON_OPTIMAL;
    else if (s.total >= cp.iteration_limit)
        reason = TERMINATION_REASON_ITERATION_LIMIT;
    if (reason == TERMINATION_REASON_UNSPECIFIED && s.charged >= cp.iteration_limit)
        reason = TERMINATION_REASON_ITERATION_LIMIT;
    s.restart = 0;
    if (reason != TERMINATION_REASON_UNSPECIFIED) {
        s.reason = reason;
        s.active = 0;
        st[inst] = s;
        return;
    }
    // should_do_adaptive_restart, utils.cu:319-359
    bool r = false;
    if (s.total == cp.K) {
        r = true;
    } else if (s.total > cp.K) {
        if (s.fp <= cp.suff * s.fp0) r = true;
        if (s.fp <= cp.nec * s.fp0 && s.fp > s.last_trial_fp) r = true;
        if (s.inner >= cp.art * s.total) r = true;
    }
    s.last_trial_fp = s.fp;
    if (r) {  // perform_restart, solver.cu:2821-3101 (target: T-output, PID primal weight)
        const double pd = sqrt(q[E_RX2]), dd = sqrt(q[E_RY2]);
        const double ratio = s.rel_d / fmax(s.rel_p, 1e-300);
        if (pd > 1e-16 && dd > 1e-16 && pd < 1e12 && dd < 1e12 && ratio > 1e-8 && ratio < 1e8) {
            const double err = log(dd) - log(pd) - log(s.pw);
            s.pw_err_sum *= cp.i_smooth;
            s.pw_err_sum += err;
            const double de = err - s.pw_last_err;
            s.pw *= exp(cp.k_p * err + cp.k_i * s.pw_err_sum + cp.k_d * de);
            s.pw_last_err = err;
        } else {
            s.pw = s.best_pw;
            s.pw_err_sum = 0.0;
            s.pw_last_err = 0.0;
        }
        const double g = fabs(log10(fmax(s.rel_d, 1e-300) / fmax(s.rel_p, 1e-300)));
        if (g < s.best_gap) {
            s.best_gap = g;
            s.best_pw = s.pw;
        }
        refresh_step_schedule(s, cp.step_margin);
        s.restart = 1;
        s.inner = 0;
        s.last_trial_fp = INFINITY;
        s.fp0 = fixed_point_error(s, q);  // the next block's fp0 (solver.cu:1279-1284)
    }
    st[inst] = s;
}

// Restart copy for instances that restarted: anchor = current = T-output.
static __global__ void __launch_bounds__(kBlock)
    restart_copy_kernel(const int *__restrict__ col_inst, const double *__restrict__ xT, double *__restrict__ x0,
                        double *__restrict__ x, long n, const int *__restrict__ row_inst,
                        const double *__restrict__ yT, double *__restrict__ y0, double *__restrict__ y, long m,
                        const InstState *__restrict__ st)
{
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        if (!st[col_inst[i]].restart) return;
        const double v = xT[i];
        x0[i] = v;
        x[i] = v;
    } else if (i < n + m) {
        const long r = i - n;
        if (!st[row_inst[r]].restart) return;
        const double v = yT[r];
        y0[r] = v;
        y[r] = v;
    }
}

// Restart copy over the rows of a block table (only the instances still in
// the table are visited): threads [0, 256/G) of a workgroup copy its rows.
template <int G>
static __global__ void __launch_bounds__(kBlock)
    restart_copy_table_kernel(BlockTable bt, const double *__restrict__ src, double *__restrict__ anchor,
                              double *__restrict__ cur, const InstState *__restrict__ st)
{
    const int tb = bt.block();
    const int inst = bt.inst[tb];
    if (!st[inst].restart) return;
    const int end = bt.row_end[inst];
    for (int r = threadIdx.x; r < (kBlock / G) * kGroupRows; r += kBlock) {
        const int row = bt.row0[tb] + r;
        if (row >= end) break;
        const double v = src[row];
        anchor[row] = v;
        cur[row] = v;
    }
}

static __global__ void __launch_bounds__(kBlock)
    active_count_kernel(const InstState *__restrict__ st, int count, int *__restrict__ out)
{
    __shared__ int red[kBlock];
    int a = 0;
    for (int i = threadIdx.x; i < count; i += kBlock) a += st[i].active;
    red[threadIdx.x] = a;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// Plain y = M x through a block table (lane-count autotuning).
template <int G>
static __global__ void __launch_bounds__(kBlock)
    table_spmv_kernel(CsrView M, BlockTable bt, const double *__restrict__ x, double *__restrict__ y)
{
    const int tb = bt.block();
    const int inst = bt.inst[tb];
    const int lane = threadIdx.x & (G - 1);
    const int end = bt.row_end[inst];
    const int base = bt.row0[tb] + threadIdx.x / G;
    for (int u = 0; u < kGroupRows; ++u) {
        const int row = base + u * (kBlock / G);
        if (row >= end) break;
        const double s = row_dot<G>(M, row, lane, x);
        if (lane == 0) y[row] = s;
    }
}

// ---------------------------------------------------------------------------
// Batched power iteration for sigma_max(A_i)^2 (utils.cu:87-228).

struct PowerState {
    double vnorm, sigma_sq;
    int iter, done;
};

// Per instance: vnorm = ||v0|| (one workgroup per instance).
static __global__ void __launch_bounds__(kBlock)
    power_init_kernel(const long *__restrict__ row_off, const double *__restrict__ v, PowerState *__restrict__ ps)
{
    const int inst = blockIdx.x;
    __shared__ double red[kWavesPerBlock];
    double a = 0.0;
    for (long i = row_off[inst] + threadIdx.x; i < row_off[inst + 1]; i += kBlock) a += v[i] * v[i];
    a = wave_sum(a);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int k = 1; k < kWavesPerBlock; ++k) r += red[k];
        ps[inst].vnorm = sqrt(r);
    }
}

static __global__ void __launch_bounds__(kBlock)
    power_normalize_kernel(const int *__restrict__ row_inst, const double *__restrict__ v, double *__restrict__ vh,
                           long m, const PowerState *__restrict__ ps, int max_iter)
{
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= m) return;
    const PowerState &p = ps[row_inst[i]];
    if (p.done || p.iter >= max_iter) return;
    vh[i] = v[i] * (1.0 / p.vnorm);
}

// w = At vh over the instances still iterating.
template <int G>
static __global__ void __launch_bounds__(kBlock)
    power_at_kernel(CsrView At, BlockTable bt, const double *__restrict__ vh, double *__restrict__ w,
                    const PowerState *__restrict__ ps, int max_iter)
{
    const int tb = bt.block();
    const int inst = bt.inst[tb];
    if (ps[inst].done || ps[inst].iter >= max_iter) return;
    const int lane = threadIdx.x & (G - 1);
    const int end = bt.row_end[inst];
    const int base = bt.row0[tb] + threadIdx.x / G;
    for (int u = 0; u < kGroupRows; ++u) {
        const int row = base + u * (kBlock / G);
        if (row >= end) break;
        const double s = row_dot<G>(At, row, lane, vh);
        if (lane == 0) w[row] = s;
    }
}

// v = A w with partial <vh, v> per workgroup.
template <int G>
static __global__ void __launch_bounds__(kBlock)
    power_a_kernel(CsrView A, BlockTable bt, const double *__restrict__ w, const double *__restrict__ vh,
                   double *__restrict__ v, double *__restrict__ parts, const PowerState *__restrict__ ps, int max_iter)
{
    const int tb = bt.block();
    const int inst = bt.inst[tb];
    if (ps[inst].done || ps[inst].iter >= max_iter) return;
    double q[1] = {0.0};
    const int lane = threadIdx.x & (G - 1);
    const int end = bt.row_end[inst];
    const int base = bt.row0[tb] + threadIdx.x / G;
    for (int u = 0; u < kGroupRows; ++u) {
        const int row = base + u * (kBlock / G);
        if (row >= end) break;
        const double s = row_dot<G>(A, row, lane, w);
        if (lane == 0) {
            v[row] = s;
            q[0] += vh[row] * s;
        }
    }
    block_reduce_store<1, 0>(q, parts + (size_t)tb * kPartStride);
}

// Per instance: sigma_sq = sum of its partials (one workgroup per instance).
static __global__ void __launch_bounds__(kBlock)
    power_sigma_kernel(const double *__restrict__ parts, const int *__restrict__ first, const int *__restrict__ cnt,
                       PowerState *__restrict__ ps, int max_iter)
{
    const int inst = blockIdx.x;
    if (ps[inst].done || ps[inst].iter >= max_iter) return;
    __shared__ double red[kWavesPerBlock];
    double acc = 0.0;
    for (int b = threadIdx.x; b < cnt[inst]; b += kBlock) acc += parts[(size_t)(first[inst] + b) * kPartStride];
    acc = wave_sum(acc);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int k = 1; k < kWavesPerBlock; ++k) r += red[k];
        ps[inst].sigma_sq = r;
    }
}

// Per instance (one workgroup each, its rows [r0, r1)): r = -sigma_sq vh + v,
// ||r|| and ||v||; convergence test and iteration count.
static __global__ void __launch_bounds__(kBlock)
    power_resid_kernel(const long *__restrict__ row_off, const double *__restrict__ v, const double *__restrict__ vh,
                       PowerState *__restrict__ ps, int max_iter, double tol)
{
    const int inst = blockIdx.x;
    PowerState p = ps[inst];
    if (p.done || p.iter >= max_iter) return;
    __shared__ double red[2][kWavesPerBlock];
    const double neg = -p.sigma_sq;
    double a = 0.0, b = 0.0;
    for (long i = row_off[inst] + threadIdx.x; i < row_off[inst + 1]; i += kBlock) {
        const double vi = v[i];
        const double r = neg * vh[i] + vi;
        a += r * r;
        b += vi * vi;
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
        for (int k = 1; k < kWavesPerBlock; ++k) {
            ra += red[0][k];
            rb += red[1][k];
        }
        p.iter += 1;
        if (sqrt(ra) < tol) p.done = 1;
        p.vnorm = sqrt(rb);
        ps[inst] = p;
    }
}

// ---------------------------------------------------------------------------
// Per-instance bound-objective scaling (preconditioner.cu:325-377) and
// segmented helpers.

static __global__ void __launch_bounds__(kBlock)
    seg_bound_sums_kernel(const long *__restrict__ row_off, const long *__restrict__ col_off,
                          const double *__restrict__ l, const double *__restrict__ u, const double *__restrict__ c,
                          double *__restrict__ out)
{
    const int inst = blockIdx.x;
    __shared__ double red[2][kWavesPerBlock];
    double a = 0.0, b = 0.0;
    for (long i = row_off[inst] + threadIdx.x; i < row_off[inst + 1]; i += kBlock) {
        const double L = l[i], U = u[i];
        const bool fL = isfinite(L), fU = isfinite(U);
        double acc = 0.0;
        if (fL && (!fU || fabs(L - U) > 1e-12)) acc += L * L;
        if (fU) acc += U * U;
        a += acc;
    }
    for (long j = col_off[inst] + threadIdx.x; j < col_off[inst + 1]; j += kBlock) b += c[j] * c[j];
    a = wave_sum(a);
    b = wave_sum(b);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        red[0][threadIdx.x / kWave] = a;
        red[1][threadIdx.x / kWave] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ra = red[0][0], rb = red[1][0];
        for (int k = 1; k < kWavesPerBlock; ++k) {
            ra += red[0][k];
            rb += red[1][k];
        }
        out[2 * inst] = 1.0 / (sqrt(ra) + 1.0);      // b_scale
        out[2 * inst + 1] = 1.0 / (sqrt(rb) + 1.0);  // c_scale
    }
}

static __global__ void __launch_bounds__(kBlock)
    seg_bound_apply_kernel(const int *__restrict__ row_inst, double *__restrict__ l, double *__restrict__ u,
                           double *__restrict__ y0, long m, const int *__restrict__ col_inst, double *__restrict__ lb,
                           double *__restrict__ ub, double *__restrict__ c, double *__restrict__ x0, long n,
                           const double *__restrict__ scales)
{
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i < m) {
        const int k = row_inst[i];
        const double cs = scales[2 * k], os = scales[2 * k + 1];
        l[i] *= cs;
        u[i] *= cs;
        y0[i] *= os;
    } else if (i < m + n) {
        const long j = i - m;
        const int k = col_inst[j];
        const double cs = scales[2 * k], os = scales[2 * k + 1];
        lb[j] *= cs;
        ub[j] *= cs;
        c[j] *= os;
        x0[j] *= cs;
    }
}

// Norms of the unscaled objective and constraint bounds per instance
// (solver.cu:1999-2067): out[2i] = ||c_i||, out[2i+1] = ||b_i|| over the
// finite l_r with l_r != u_r and the finite u_r; 2-norm, or max |.| when linf.
// One workgroup per instance, fixed summation order.
static __global__ void __launch_bounds__(kBlock)
    seg_norms_kernel(const long *__restrict__ row_off, const long *__restrict__ col_off, const double *__restrict__ l,
                     const double *__restrict__ u, const double *__restrict__ c, int linf, double *__restrict__ out)
{
    const int inst = blockIdx.x;
    __shared__ double red[2][kWavesPerBlock];
    double a = 0.0, b = 0.0;
    for (long j = col_off[inst] + threadIdx.x; j < col_off[inst + 1]; j += kBlock) {
        const double v = c[j];
        a = linf ? fmax(a, fabs(v)) : a + v * v;
    }
    for (long i = row_off[inst] + threadIdx.x; i < row_off[inst + 1]; i += kBlock) {
        const double lo = l[i], hi = u[i];
        if (isfinite(lo) && lo != hi) b = linf ? fmax(b, fabs(lo)) : b + lo * lo;
        if (isfinite(hi)) b = linf ? fmax(b, fabs(hi)) : b + hi * hi;
    }
    a = linf ? wave_max(a) : wave_sum(a);
    b = linf ? wave_max(b) : wave_sum(b);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        red[0][threadIdx.x / kWave] = a;
        red[1][threadIdx.x / kWave] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ra = red[0][0], rb = red[1][0];
        for (int k = 1; k < kWavesPerBlock; ++k) {
            ra = linf ? fmax(ra, red[0][k]) : ra + red[0][k];
            rb = linf ? fmax(rb, red[1][k]) : rb + red[1][k];
        }
        out[2 * inst] = linf ? ra : sqrt(ra);
        out[2 * inst + 1] = linf ? rb : sqrt(rb);
    }
}

// The recorded scaling passes applied again to new vectors of the same
// matrices: per pass, c *= 1/d, variable bounds *= d (scale_vars_kernel) and
// constraint bounds *= 1/e (scale_cons_kernel), in the passes' order, so the
// result has the bits a full rescale would give.  Factor tables are pass-major
// with `period` entries per pass (one instance's when all share the matrix).
static __global__ void __launch_bounds__(kBlock)
    replay_scaling_kernel(double *__restrict__ c, double *__restrict__ lb, double *__restrict__ ub, long n,
                          const double *__restrict__ fd, const double *__restrict__ finv_d, long pn,
                          double *__restrict__ l, double *__restrict__ u, long m, const double *__restrict__ finv_e,
                          long pm, int passes)
{
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const long k = i % pn;
        double cj = c[i], lo = lb[i], hi = ub[i];
        for (int p = 0; p < passes; ++p) {
            const double dj = fd[p * pn + k];
            cj *= finv_d[p * pn + k];
            lo *= dj;
            hi *= dj;
        }
        c[i] = cj;
        lb[i] = lo;
        ub[i] = hi;
    } else if (i < n + m) {
        const long r = i - n, k = r % pm;
        double lo = l[r], hi = u[r];
        for (int p = 0; p < passes; ++p) {
            const double ie = finv_e[p * pm + k];
            lo *= ie;
            hi *= ie;
        }
        l[r] = lo;
        u[r] = hi;
    }
}

// Warm start into the scaled space: x * D * b_scale, y * E * c_scale.
static __global__ void __launch_bounds__(kBlock)
    scale_start_kernel(const int *__restrict__ col_inst, const double *__restrict__ xs, const double *__restrict__ D,
                       double *__restrict__ x, long n, const int *__restrict__ row_inst, const double *__restrict__ ys,
                       const double *__restrict__ E, double *__restrict__ y, long m, const InstState *__restrict__ st)
{
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        x[i] = xs[i] * D[i] * st[col_inst[i]].b_scale;
    } else if (i < n + m) {
        const long r = i - n;
        y[r] = ys[r] * E[r] * st[row_inst[r]].c_scale;
    }
}

// Result: reduced cost (with infinite-bound clamps) and unscaled x, y.
template <int G>
static __global__ void __launch_bounds__(kBlock)
    result_x_kernel(CsrView At, BlockTable bt, const double *__restrict__ yT, const double *__restrict__ xT,
                    const double *__restrict__ c, const double *__restrict__ D, const double *__restrict__ lb,
                    const double *__restrict__ ub, const InstState *__restrict__ st, double *__restrict__ rc,
                    double *__restrict__ xo)
{
    const int tb = bt.block();
    const int inst = bt.inst[tb];
    const int lane = threadIdx.x & (G - 1);
    const int end = bt.row_end[inst];
    const int base = bt.row0[tb] + threadIdx.x / G;
    const double cs = st[inst].c_scale, bs = st[inst].b_scale;
    for (int u = 0; u < kGroupRows; ++u) {
        const int row = base + u * (kBlock / G);
        if (row >= end) break;
        const double s = row_dot<G>(At, row, lane, yT);
        if (lane != 0) continue;
        double r = (c[row] - s) * D[row] / cs;
        if (!isfinite(lb[row])) r = fmin(r, 0.0);
        if (!isfinite(ub[row])) r = fmax(r, 0.0);
        rc[row] = r;
        xo[row] = xT[row] / D[row] / bs;
    }
}

static __global__ void __launch_bounds__(kBlock)
    result_y_kernel(const int *__restrict__ row_inst, const double *__restrict__ yT, const double *__restrict__ E,
                    const InstState *__restrict__ st, double *__restrict__ yo, long m)
{
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i < m) yo[i] = yT[i] / E[i] / st[row_inst[i]].c_scale;
}

}  // namespace batch
}  // namespace lp
}  // namespace pdal
