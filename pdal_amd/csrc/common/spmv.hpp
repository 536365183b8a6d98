// Shared FP64 CSR SpMV building blocks for gfx950 (wave64).
//
// G lanes of a wavefront own one row (G a power of two, 1..64): lanes stride
// over the row's nonzeros, a G-lane butterfly sums the partials, and lane 0 of
// the group holds the row's dot product for a fused epilogue.  The summation
// order depends only on G, so results are run-to-run deterministic.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "device_reduce.hpp"

namespace pdal {

struct CsrView {
    const int *__restrict__ rp;
    const int *__restrict__ ci;
    const double *__restrict__ v;
    int rows;
};

template <int G>
__device__ __forceinline__ double row_dot(const CsrView &M, int row, int lane, const double *__restrict__ x)
{
    double s = 0.0;
    const int b = M.rp[row], e = M.rp[row + 1];
    for (int k = b + lane; k < e; k += G) s += M.v[k] * x[M.ci[k]];
    return group_sum<G>(s);
}

template <int G>
__device__ __forceinline__ void row_dot2(const CsrView &M, int row, int lane, const double *__restrict__ xa,
                                         const double *__restrict__ xb, double &s_diff, double &s_b)
{   // s_diff = sum a*(xa - xb) (the difference is rounded first, as a materialised
    // delta vector would be); s_b = sum a*xb.
    double d = 0.0, t = 0.0;
    const int b = M.rp[row], e = M.rp[row + 1];
    for (int k = b + lane; k < e; k += G) {
        const int c = M.ci[k];
        const double a = M.v[k];
        const double xbv = xb[c];
        d += a * (xa[c] - xbv);
        t += a * xbv;
    }
    s_diff = group_sum<G>(d);
    s_b = group_sum<G>(t);
}

// Plain y = M x.
template <int G>
static __global__ void __launch_bounds__(kBlock) spmv_kernel(CsrView M, const double *__restrict__ x, double *__restrict__ y)
{
    const int tid = blockIdx.x * kBlock + threadIdx.x;
    const int row = tid / G;
    const int lane = tid & (G - 1);
    if (row >= M.rows) return;
    const double s = row_dot<G>(M, row, lane, x);
    if (lane == 0) y[row] = s;
}


// Host: call f(std::integral_constant<int, G>{}) for the runtime lane count G.
template <class F>
void dispatch_group(int G, F &&f)
{
    switch (G) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    case 32: f(std::integral_constant<int, 32>{}); break;
    default: f(std::integral_constant<int, 64>{}); break;
    }
}

}  // namespace pdal
