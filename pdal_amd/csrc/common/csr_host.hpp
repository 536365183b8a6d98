// Host-side CSR helpers shared by the CP_AL and FA_CP drivers: a stable
// transpose, device-resident CSR storage and the SpMV lane-count heuristic.
#pragma once

#include <cstdlib>
#include <vector>

#include "hip_util.hpp"
#include "spmv.hpp"

namespace pdal {

inline int blocks_for(long work) { return work > 0 ? (int)((work + kBlock - 1) / kBlock) : 1; }

struct HostCsr {
    int rows = 0, cols = 0;
    std::vector<int> rp, ci;
    std::vector<double> v;
    int nnz() const { return rp.empty() ? 0 : rp[rows]; }
};

// Stable CSR transpose: rows stay ascending inside each column, the order the
// reference gets from cusparseCsr2cscEx2 (LP/CP_AL/src/solver.cu:1859-1890) and
// from its host transposes (QP/FA_CP/src/solver_state.cu:155-218).
inline HostCsr transpose_csr(int m, int n, const int *rp, const int *ci, const double *v)
{
    HostCsr T;
    T.rows = n;
    T.cols = m;
    const int nnz = rp[m];
    T.rp.assign((size_t)n + 1, 0);
    for (int k = 0; k < nnz; ++k) T.rp[ci[k] + 1]++;
    for (int j = 0; j < n; ++j) T.rp[j + 1] += T.rp[j];
    T.ci.resize(nnz);
    T.v.resize(nnz);
    std::vector<int> next(T.rp.begin(), T.rp.end() - 1);
    for (int i = 0; i < m; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int p = next[ci[k]]++;
            T.ci[p] = i;
            T.v[p] = v[k];
        }
    return T;
}

struct DevCsr {
    DBuf<int> rp, ci;
    DBuf<double> v;
    int rows = 0, nnz = 0, G = 4;
    CsrView view() const { return CsrView{rp.get(), ci.get(), v.get(), rows}; }
};

// Lanes per row: a power of two giving about two nonzeros per lane, unless the
// environment variable `env` (when non-null) names a valid lane count.
inline int pick_group(const char *env, long nnz, int rows)
{
    if (env)
        if (const char *raw = getenv(env)) {
            const int g = atoi(raw);
            if (g == 1 || g == 2 || g == 4 || g == 8 || g == 16 || g == 32 || g == 64) return g;
        }
    const double mean = rows > 0 ? (double)nnz / rows : 0.0;
    int g = 1;
    while (g < 64 && 2.0 * g <= mean / 2.0) g *= 2;
    return g;
}

}  // namespace pdal
