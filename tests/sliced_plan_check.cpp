// CPU check of the column-sliced CSR plan (pdal_amd/csrc/common/sliced_plan.hpp).
//
// For random CSR matrices it builds the plan and replays the index arithmetic
// of the device routine sliced_row_sums lane by lane: wavefront ranges from
// seg, chunks of 64 nonzeros (4 chunks per step), the packed column
// and row, the segmented prefix sum over the 64 lanes and the store of each
// row's partial by the row's last lane.  It asserts that every position a lane
// loads lies inside the matrix, that the positions cover every nonzero exactly
// once, that each row index lies in the wavefront's rows and each column in
// the slice, that perm is a permutation, that no valid packed index equals the
// 0xffffffff sentinel, and that every row sum matches the CSR row sum (to
// rounding: the lane tree adds in another order).
//
// Build: g++ -O2 -std=c++17 -I pdal_amd/csrc/common tests/sliced_plan_check.cpp
// Prints "OK <cases>" and exits 0, or prints the first failure and exits 1.
#include "sliced_plan.hpp"

#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

using namespace pdal;

static int fail(const char *what, long a, long b)
{
    printf("FAIL %s (%ld, %ld)\n", what, a, b);
    return 1;
}

struct Csr {
    int R, C;
    std::vector<int> rp, ci;
    std::vector<double> v;
};

static Csr random_csr(std::mt19937_64 &g, int R, int C, double mean_len, bool sorted, int long_row)
{
    Csr M{R, C, {0}, {}, {}};
    std::poisson_distribution<int> len(mean_len);
    std::uniform_int_distribution<int> col(0, C - 1);
    std::normal_distribution<double> val;
    for (int r = 0; r < R; ++r) {
        int L = (r % 97 == 5) ? 0 : len(g);  // some empty rows
        if (r == long_row) L = 600;           // a long row: several chunks in one slice
        std::vector<int> cs;
        for (int k = 0; k < L; ++k) cs.push_back(col(g));
        if (sorted) std::sort(cs.begin(), cs.end());
        for (int c : cs) {
            M.ci.push_back(c);
            M.v.push_back(val(g));
        }
        M.rp.push_back((int)M.ci.size());
    }
    return M;
}

static int check(const Csr &M, long slice_cols, int rpw)
{
    SlicedPlan P;
    if (!build_sliced_plan(M.R, M.C, M.rp.data(), M.ci.data(), slice_cols, rpw, P)) return fail("plan refused", M.R, M.C);
    const long nnz = M.rp[M.R];
    const int W = kSliceWavesPerGroup, L = kSliceWaveLanes, U = 4;
    if ((long)P.perm.size() != nnz || (long)P.ci.size() != nnz) return fail("array sizes", (long)P.perm.size(), nnz);
    if ((long)P.seg.size() != (long)P.groups * W + 1) return fail("seg size", (long)P.seg.size(), 0);
    std::vector<char> seen(nnz, 0), visited(nnz, 0);
    std::vector<double> v_s(nnz);
    for (long k = 0; k < nnz; ++k) {
        if (P.perm[k] < 0 || P.perm[k] >= nnz) return fail("perm range", k, P.perm[k]);
        if (seen[P.perm[k]]++) return fail("perm repeats", k, P.perm[k]);
        v_s[k] = M.v[P.perm[k]];
    }
    const uint32_t mask = (1u << P.colbits) - 1u;
    std::mt19937_64 g(7);
    std::normal_distribution<double> nd;
    std::vector<double> x(M.C);
    for (auto &e : x) e = nd(g);
    std::vector<double> acc((size_t)W * P.rpw);
    for (int p = 0; p < P.groups; ++p) {
        std::fill(acc.begin(), acc.end(), 0.0);
        for (int w = 0; w < W; ++w) {
                const size_t idx = (size_t)p * W + w;
                const long b = P.seg[idx], e = P.seg[idx + 1];
                if (b > e || b < 0 || e > nnz) return fail("range bounds", b, e);
                int last_slice = 0;
                for (long c0 = b; c0 < e; c0 += (long)L * U)
                    for (int u = 0; u < U; ++u) {
                        if (c0 + (long)u * L >= e) break;
                        int row[64];
                        double val[64];
                        for (int lane = 0; lane < L; ++lane) {
                            const long k = c0 + (long)u * L + lane;
                            if (k >= e) {
                                row[lane] = -1;
                                val[lane] = 0.0;
                                continue;
                            }
                            if (visited[k]++) return fail("position visited twice", k, p);
                            const uint32_t pk = P.ci[k];
                            const long col = pk & mask;
                            row[lane] = (int)(pk >> P.colbits);
                            if (row[lane] < 0 || row[lane] >= P.rpw) return fail("row outside the wavefront", k, row[lane]);
                            if (col >= M.C) return fail("column out of range", k, col);
                            if (col / slice_cols < last_slice) return fail("range not sorted by slice", k, col);
                            last_slice = (int)(col / slice_cols);
                            val[lane] = v_s[k] * x[col];
                        }
                        int first[64];  // first lane of each lane's run of equal adjacent rows
                        for (int lane = 0; lane < L; ++lane)
                            first[lane] = (lane == 0 || row[lane] != row[lane - 1]) ? lane : first[lane - 1];
                        for (int d = 1; d < L; d <<= 1) {  // lanes read the values of the previous step
                            double nv[64];
                            unsigned long long sm = 0;
                            for (int lane = 0; lane < L; ++lane) {
                                const bool same = lane - d >= first[lane];
                                nv[lane] = same ? val[lane] + val[lane - d] : val[lane];
                                if (same) sm |= 1ull << lane;
                            }
                            for (int lane = 0; lane < L; ++lane) val[lane] = nv[lane];
                            if ((sm & (sm << d)) == 0ull) break;
                        }
                        // run ends store piece by piece between descents (rows restart
                        // when the chunk crosses into the next slice); within a piece
                        // no row may end two runs (the store of one instruction)
                        int lo = 0;
                        while (lo < L) {
                            int hi = lo + 1;
                            while (hi < L && !(row[hi] >= 0 && row[hi] < row[hi - 1])) ++hi;
                            std::vector<int> rows_here;
                            for (int lane = lo; lane < hi; ++lane)
                                if (row[lane] >= 0 && (lane == L - 1 || row[lane + 1] != row[lane])) {
                                    for (int r : rows_here)
                                        if (r == row[lane]) return fail("row stored twice by one instruction", r, p);
                                    rows_here.push_back(row[lane]);
                                    acc[(size_t)w * P.rpw + row[lane]] += val[lane];
                                }
                            lo = hi;
                        }
                    }
            }
        for (int j = 0; j < W * P.rpw; ++j) {
            const long r = (long)p * W * P.rpw + j;
            if (r >= M.R) {
                if (acc[j] != 0.0) return fail("padding row has a sum", r, p);
                continue;
            }
            double ref = 0.0, mag = 0.0;
            for (int k = M.rp[r]; k < M.rp[r + 1]; ++k) {
                ref += M.v[k] * x[M.ci[k]];
                mag += std::fabs(M.v[k] * x[M.ci[k]]);
            }
            if (std::fabs(acc[j] - ref) > 1e-12 * (mag + 1e-300)) return fail("row sum differs", r, p);
        }
    }
    for (long k = 0; k < nnz; ++k)
        if (!visited[k]) return fail("position never visited", k, nnz);
    return 0;
}

int main()
{
    std::mt19937_64 g(12345);
    int cases = 0;
    struct Case {
        int R, C;
        double len;
        long slice_cols;
        int rpw;
        bool sorted;
        int long_row;
    };
    const Case cs[] = {
        {1, 1, 1.0, 1, 1, true, -1},           {300, 700, 6.0, 64, 1, true, 17},
        {1000, 257, 12.0, 16, 2, false, 999},  {5000, 20000, 8.0, 1024, 64, true, -1},
        {5000, 20000, 8.0, 3000, 8, false, 0}, {70000, 9000, 3.0, 2048, 512, true, 65535},
        {513, 100000, 30.0, 4096, 16, true, 256}, {20000, 40000, 12.0, 1500, 64, true, 3},
        {2000, 1000, 5.0, 1000, 64, false, 700},  // one slice: the long row spans ten chunks
        {600, 2000000, 3.0, 1 << 21, 2048, true, 40},  // 21 column bits + 11 row bits: all 32 used
    };
    for (const Case &c : cs) {
        Csr M = random_csr(g, c.R, c.C, c.len, c.sorted, c.long_row);
        if (int rc = check(M, c.slice_cols, c.rpw)) {
            printf("case R=%d C=%d slice_cols=%ld rpw=%d\n", c.R, c.C, c.slice_cols, c.rpw);
            return rc;
        }
        ++cases;
    }
    // the row index must fit beside the column bits
    {
        Csr M{1, 1 << 27, {0, 1}, {5}, {1.0}};
        SlicedPlan P;
        if (build_sliced_plan(M.R, M.C, M.rp.data(), M.ci.data(), 1 << 20, 64, P)) return fail("32-bit overflow planned", 0, 0);
        ++cases;
    }
    printf("OK %d\n", cases);
    return 0;
}
