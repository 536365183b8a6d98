// Deterministic wavefront / block reductions for gfx950 (wave64).
//
// Every reduction here has a fixed combination order that depends only on the
// launch shape, never on timing: a lane butterfly inside each 64-wide
// wavefront (commutative pairwise adds, so every lane ends with the same bits),
// then the block's wavefronts combined in index order by thread 0, then the
// per-block partials combined by one block in index order.  No FP64 atomics.
#pragma once

#include <hip/hip_runtime.h>

namespace pdal {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

// Sum over a G-lane group (G a power of two <= 64), result in every lane.
template <int G>
__device__ __forceinline__ double group_sum(double v)
{
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
    return v;
}

__device__ __forceinline__ double wave_sum(double v) { return group_sum<kWave>(v); }

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    return v;
}

// Reduce q[0..NSUM) as sums and q[NSUM..NSUM+NMAX) as maxima over a 256-thread
// block; thread 0 writes the NSUM+NMAX results to out[0..).
template <int NSUM, int NMAX>
__device__ __forceinline__ void block_reduce_store(double (&q)[NSUM + NMAX], double *__restrict__ out)
{
    constexpr int NQ = NSUM + NMAX;
    __shared__ double red[kWavesPerBlock][NQ];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        const double r = (k < NSUM) ? wave_sum(q[k]) : wave_max(q[k]);
        if (lane == 0) red[wave][k] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            double r = red[0][k];
            for (int w = 1; w < kWavesPerBlock; ++w) r = (k < NSUM) ? r + red[w][k] : fmax(r, red[w][k]);
            out[k] = r;
        }
    }
}

// Combine `nparts` rows of `stride` partials (row b = block b) into out[0..NQ):
// entries [0,nsum) summed, [nsum,nq) maxed.  Launch with one 256-thread block.
static __global__ void __launch_bounds__(kBlock)
    reduce_partials_kernel(const double *__restrict__ parts, int nparts, int stride, int nsum, int nq,
                           double *__restrict__ out)
{
    __shared__ double red[kWavesPerBlock];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    for (int k = 0; k < nq; ++k) {
        const bool is_sum = k < nsum;
        double acc = 0.0;
        for (int b = threadIdx.x; b < nparts; b += kBlock) {
            const double v = parts[(size_t)b * stride + k];
            acc = is_sum ? acc + v : fmax(acc, v);
        }
        acc = is_sum ? wave_sum(acc) : wave_max(acc);
        if (lane == 0) red[wave] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = red[0];
            for (int w = 1; w < kWavesPerBlock; ++w) r = is_sum ? r + red[w] : fmax(r, red[w]);
            out[k] = r;
        }
        __syncthreads();
    }
}

}  // namespace pdal
