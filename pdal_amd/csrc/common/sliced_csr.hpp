// Column-sliced CSR: the SpMV layout for matrices whose gathered vector is far
// larger than one XCD's L2 (4 MB on gfx950).  The host plan and the storage
// order are described in sliced_plan.hpp.
//
// With uniform random column indices, every 8-byte gather of a plain CSR pass
// misses the L2 and pulls a whole line through the fabric, so such a pass runs
// at the fabric's random-line rate instead of the HBM stream rate.  Here each
// workgroup owns a range of rows and walks the column slices in order; in
// slice s it multiplies only the nonzeros of its rows whose column lies in
// slice s.  All workgroups are resident at once and sweep the slices together,
// so the slice being gathered (1 MB) sits in each XCD's L2 and the gathers hit
// there.
//
// A wavefront streams its range of nonzeros (sorted by slice, then row) 64 at
// a time, one per lane (coalesced loads, no idle lanes whatever the row
// lengths), with the loads of the next U chunks issued before the current ones
// are summed; a segmented prefix sum over the lanes (equal rows are adjacent)
// gives each row's partial, which the row's last lane adds to the row's
// running sum in LDS.  Only this wavefront touches those rows and it takes its
// chunks in order, so the summation order of every row is fixed: chunks in
// order, the same lane tree inside a chunk.  The fused epilogue then reads
// the row sums from LDS.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_reduce.hpp"
#include "hip_util.hpp"
#include "sliced_plan.hpp"

namespace pdal {

static_assert(kSliceWaveLanes == kWave, "the plan's wavefront width is gfx950's");


// Device storage of a sliced matrix.
struct DevSliced {
    DBuf<uint32_t> ci;
    DBuf<int> seg;
    DBuf<double> v;
    int rows = 0, slices = 0, rpw = 0, groups = 0, colbits = 0;
    long slice_cols = 0;
    bool ready = false;
    SlicedView view() const { return SlicedView{ci.get(), v.get(), seg.get(), rows, slices, rpw, colbits}; }
    size_t lds_bytes() const { return (size_t)kSliceWavesPerGroup * rpw * sizeof(double); }
};

// Whole-wavefront lane shifts by one lane (DPP wave_shr:1 / wave_shl:1, VALU
// moves instead of LDS permutes): lane i receives lane i - 1 (resp. i + 1);
// the lane shifted in from outside the wavefront receives `fill`.  All 64
// lanes must be active.
__device__ __forceinline__ int lane_prev(int v, int fill) { return __builtin_amdgcn_update_dpp(fill, v, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ int lane_next(int v, int fill) { return __builtin_amdgcn_update_dpp(fill, v, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ double lane_prev(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = lane_prev((int)(b & 0xffffffffll), 0), hi = lane_prev((int)(b >> 32), 0);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Row sums of workgroup p's rows into acc (LDS, kSliceWavesPerGroup * rpw
// doubles; acc[w * rpw + j] is row p * 4 * rpw + w * rpw + j).  U chunks of 64
// nonzeros have their loads in flight together.  Ends with a workgroup
// barrier, so every thread may read any row's sum afterwards.
template <int U>
__device__ __forceinline__ void sliced_row_sums(const SlicedView &M, int p, const double *__restrict__ x,
                                                double *acc)
{
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(t / kWave);
    const int rw = kSliceWavesPerGroup * M.rpw;
    for (int j = t; j < rw; j += kSliceBlock) acc[j] = 0.0;
    __syncthreads();
    double *accw = acc + (size_t)w * M.rpw;
    const uint32_t mask = (M.colbits >= 32) ? 0xffffffffu : ((1u << M.colbits) - 1u);
    const int b = __builtin_amdgcn_readfirstlane(M.seg[(size_t)p * kSliceWavesPerGroup + w]);
    const int e = __builtin_amdgcn_readfirstlane(M.seg[(size_t)p * kSliceWavesPerGroup + w + 1]);
    uint32_t pk[U];
    double pr[U];
    auto load = [&](int c0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = c0 + u * kWave + lane;
            pk[u] = k < e ? M.ci[k] : 0u;
            pr[u] = k < e ? M.v[k] : 0.0;
        }
    };
    if (b < e) load(b);
    for (int c0 = b; c0 < e; c0 += kWave * U) {
        uint32_t ck[U];
        double cv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ck[u] = pk[u];
            cv[u] = pr[u];
        }
        if (c0 + kWave * U < e) load(c0 + kWave * U);  // next chunks in flight while these are summed
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (c0 + u * kWave + lane < e) cv[u] *= x[ck[u] & mask];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (c0 + u * kWave >= e) break;  // wave-uniform: no nonzeros left
            // lanes past the range end sit at the top of the chunk with row
            // -1: a run of their own, never stored
            const int row = (c0 + u * kWave + lane < e) ? (int)(ck[u] >> M.colbits) : -1;
            double v = cv[u];
            // runs of equal adjacent rows; a row can recur in a later run when
            // the chunk crosses into the next column slice, so runs are told
            // apart by their first lane, not by the row
            const int next = lane_next(row, -1);
            const int prev = lane_prev(row, -1);
            const unsigned long long starts = __ballot(lane == 0 || row != prev);
            const int first = 63 - __builtin_clzll(starts & (~0ull >> (kWave - 1 - lane)));
            // segmented inclusive sum over the lanes: after the step of offset
            // d every lane holds the sum of its run over the last 2d lanes; it
            // stops once no run is longer than that (runs are mostly one or
            // two nonzeros long, so the first step, a DPP shift, is the rule)
            {
                const double ov = lane_prev(v);
                const bool same = lane - 1 >= first;
                if (same) v += ov;
                const unsigned long long sm = __ballot(same);
                if ((sm & (sm << 1)) != 0ull)
                    for (int d = 2; d < kWave; d <<= 1) {
                        const double ow = __shfl_up(v, d, kWave);
                        const bool same_d = lane - d >= first;
                        if (same_d) v += ow;
                        const unsigned long long sd = __ballot(same_d);
                        if ((sd & (sd << d)) == 0ull) break;
                    }
            }
            // the last lane of each run adds the run to the row's sum; where
            // the rows restart (a descent) a row may end a run on both sides,
            // so the pieces between descents store one after the other
            const bool end = row >= 0 && (lane == kWave - 1 || next != row);
            unsigned long long dm = __ballot(lane > 0 && row >= 0 && row < prev);
            int lo = 0;
            while (true) {
                const int hi = dm ? __builtin_ctzll(dm) : kWave;
                if (end && lane >= lo && lane < hi) accw[row] += v;
                if (!dm) break;
                lo = hi;
                dm &= dm - 1;
            }
        }
    }
    __syncthreads();
}

// Gather the values into the stored order (after any in-place rescaling).
static __global__ void __launch_bounds__(kBlock)
    sliced_permute_kernel(const int *__restrict__ perm, long nnz, const double *__restrict__ v,
                          double *__restrict__ v_s)
{
    for (long k = (long)blockIdx.x * kBlock + threadIdx.x; k < nnz; k += (long)gridDim.x * kBlock) v_s[k] = v[perm[k]];
}

// Plain y = M x over the sliced layout (autotuning and tests).
template <int U>
static __global__ void __launch_bounds__(kSliceBlock) sliced_spmv_kernel(SlicedView M, const double *__restrict__ x,
                                                                        double *__restrict__ y)
{
    extern __shared__ double acc[];
    sliced_row_sums<U>(M, blockIdx.x, x, acc);
    const int rw = kSliceWavesPerGroup * M.rpw;
    for (int j = threadIdx.x; j < rw; j += kSliceBlock) {
        const long row = (long)blockIdx.x * rw + j;
        if (row < M.rows) y[row] = acc[j];
    }
}

// Host: call f(std::integral_constant<int, U>{}) for the chunks in flight (2, 4 or 8).
template <class F>
void dispatch_unroll(int u, F &&f)
{
    switch (u) {
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    default: f(std::integral_constant<int, 4>{}); break;
    }
}

}  // namespace pdal
