// Host-side planning of the column-sliced CSR layout (sliced_csr.hpp holds
// the device side).  Plain C++ with no HIP dependency, so the plan can be
// checked on a CPU-only machine.
//
// Layout: columns are cut into S slices of `slice_cols` columns; rows into
// groups of kSliceWavesPerGroup * rpw rows (one 256-thread workgroup per
// group, wavefront w owning the group's rows [w * rpw, (w + 1) * rpw)).  The
// nonzeros are stored in the order (group, wavefront, slice, row, CSR order
// within the row): the nonzeros one wavefront multiplies form one contiguous
// range seg[g * W + w] .. seg[g * W + w + 1], sorted by slice, so every
// wavefront streams its range from the first column slice to the last and all
// of them sweep the columns together.  Each stored column index carries the
// row of its nonzero inside the wavefront's rows in its high bits: packed =
// col | (row_in_wave << colbits), so the layout adds no bytes per nonzero to
// the CSR arrays (and drops the row pointers).
#pragma once

#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

namespace pdal {

constexpr int kSliceWaveLanes = 64;                         // wavefront width of gfx950
constexpr int kSliceBlock = 256;                            // threads per workgroup of a sliced pass
constexpr int kSliceWavesPerGroup = kSliceBlock / kSliceWaveLanes;
constexpr long kSliceDefaultCols = 64 * 1024;               // columns per slice: 512 KB of FP64
constexpr int kSliceTargetWaves = 2048;                     // wavefronts of a pass: 8 per CU on 256 CUs
constexpr int kSliceMaxRowsPerWave = 1024;                  // 32 KB of row sums in LDS per workgroup

struct SlicedView {
    const uint32_t *__restrict__ ci;  // packed column | row-in-wavefront << colbits
    const double *__restrict__ v;
    const int *__restrict__ seg;      // range starts of the wavefronts, groups * W + 1 entries
    int rows;                         // R
    int slices;                       // S
    int rpw;                          // rows per wavefront
    int colbits;                      // bits of the column in a packed index
};

struct SlicedPlan {
    int rows = 0, cols = 0, slices = 0, rpw = 0, groups = 0, colbits = 0;
    long slice_cols = 0;
    std::vector<int> perm;      // perm[pos] = CSR index of the nonzero stored at pos
    std::vector<uint32_t> ci;   // packed column indices in stored order
    std::vector<int> seg;
};


inline int bits_for(long v)
{   // bits needed to hold the values 0 .. v - 1 (at least 1)
    int b = 1;
    while (b < 31 && (1L << b) < v) ++b;
    return b;
}

// Rows per wavefront: `want` when it is a power of two in 1 ..
// kSliceMaxRowsPerWave, else a power of two (64 .. kSliceMaxRowsPerWave)
// giving about kSliceTargetWaves wavefronts; halved until the row index fits
// beside the column bits of a `cols`-column matrix.  Measured on a 2e6 x 4e6
// random LP, the passes run fastest with every wavefront resident and each
// one owning as many rows as possible: the fewer ranges sweep the column
// slices closer together, so more gathers hit the L2.
inline int sliced_rows_per_wave(int rows, int cols, int want = 0)
{
    int rpw = 64;
    if (want > 0 && want <= kSliceMaxRowsPerWave && (want & (want - 1)) == 0) {
        rpw = want;
    } else {
        const long need = ((long)rows + kSliceTargetWaves - 1) / kSliceTargetWaves;
        while (rpw < need && rpw < kSliceMaxRowsPerWave) rpw *= 2;
    }
    while (rpw > 1 && bits_for(cols) + bits_for(rpw) > 32) rpw /= 2;
    return rpw;
}

// Build the plan for an R x C CSR matrix.  Returns false (plan unusable) when
// the row index inside a wavefront does not fit beside the column bits, or
// the column indices are out of range.
inline bool build_sliced_plan(int R, int C, const int *rp, const int *ci, long slice_cols, int rpw, SlicedPlan &P)
{
    if (R <= 0 || C <= 0 || slice_cols <= 0 || rpw <= 0) return false;
    P.rows = R;
    P.cols = C;
    P.rpw = rpw;
    P.colbits = bits_for(C);
    if (P.colbits + bits_for(rpw) > 32) return false;
    const long rw = (long)kSliceWavesPerGroup * rpw;
    P.groups = (int)(((long)R + rw - 1) / rw);
    P.slice_cols = slice_cols;
    P.slices = (int)(((long)C + slice_cols - 1) / slice_cols);
    const int S = P.slices, W = kSliceWavesPerGroup;
    const long nnz = rp[R];
    P.perm.assign((size_t)nnz, 0);
    P.ci.assign((size_t)nnz, 0);
    P.seg.assign((size_t)P.groups * W + 1, 0);
    P.seg.back() = (int)nnz;
    // every group's nonzeros are a permutation of its own CSR range, so the
    // groups are planned in parallel
    auto plan_groups = [&](int g0, int g1, bool &ok) {
        std::vector<long> cnt((size_t)S * W), pos((size_t)S * W);
        for (int g = g0; g < g1; ++g) {
            const long r0 = (long)g * rw, r1 = std::min<long>(R, r0 + rw);
            std::fill(cnt.begin(), cnt.end(), 0);
            for (long r = r0; r < r1; ++r) {
                const int w = (int)((r - r0) / rpw);
                for (int k = rp[r]; k < rp[r + 1]; ++k) {
                    if (ci[k] < 0 || ci[k] >= C) {
                        ok = false;
                        return;
                    }
                    cnt[(size_t)(ci[k] / slice_cols) * W + w]++;
                }
            }
            long at = rp[r0];
            for (int w = 0; w < W; ++w) {
                P.seg[(size_t)g * W + w] = (int)at;
                for (int s = 0; s < S; ++s) {
                    pos[(size_t)s * W + w] = at;
                    at += cnt[(size_t)s * W + w];
                }
            }
            if (at != rp[r1]) {
                ok = false;
                return;
            }
            for (long r = r0; r < r1; ++r) {
                const int w = (int)((r - r0) / rpw);
                const uint32_t lr = (uint32_t)((r - r0) - (long)w * rpw);
                for (int k = rp[r]; k < rp[r + 1]; ++k) {
                    const long q = pos[(size_t)(ci[k] / slice_cols) * W + w]++;
                    P.perm[(size_t)q] = k;
                    P.ci[(size_t)q] = (uint32_t)ci[k] | (lr << P.colbits);
                }
            }
        }
    };
    const int nth = std::max(1, std::min<int>(16, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    std::vector<char> oks(nth, 1);
    const int per = (P.groups + nth - 1) / nth;
    for (int t = 0; t < nth; ++t) {
        const int g0 = t * per, g1 = std::min(P.groups, g0 + per);
        if (g0 >= g1) break;
        th.emplace_back([&, t, g0, g1] {
            bool ok = true;
            plan_groups(g0, g1, ok);
            oks[t] = ok;
        });
    }
    for (auto &t : th) t.join();
    for (char o : oks)
        if (!o) return false;
    return true;
}

}  // namespace pdal
