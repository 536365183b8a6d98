// LP presolve and postsolve (see presolve.hpp for the reductions and the
// reference interface they sit behind).
#include "presolve.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace pdal {
namespace lp {

namespace {

constexpr double kFeasTol = 1e-9;  // bound comparisons, relative to 1 + |value|

bool le_tol(double a, double b) { return a <= b + kFeasTol * (1.0 + std::fmax(std::fabs(a), std::fabs(b))); }

double *copy_of(const std::vector<double> &v)
{
    double *p = (double *)malloc(std::max<size_t>(v.size(), 1) * sizeof(double));
    if (!v.empty()) memcpy(p, v.data(), v.size() * sizeof(double));
    return p;
}

}  // namespace

Presolved::~Presolved()
{
    if (!reduced) return;
    free(reduced->constraint_matrix_row_pointers);
    free(reduced->constraint_matrix_col_indices);
    free(reduced->constraint_matrix_values);
    free(reduced->variable_lower_bound);
    free(reduced->variable_upper_bound);
    free(reduced->objective_vector);
    free(reduced->constraint_lower_bound);
    free(reduced->constraint_upper_bound);
    free(reduced->primal_start);
    free(reduced->dual_start);
    free(reduced);
}

const char *presolve_status_name(int status)
{   // get_presolve_status_str, presolve.c:16-31
    switch (status) {
    case kPresolveUnchanged: return "UNCHANGED";
    case kPresolveReduced: return "REDUCED";
    case kPresolveInfeasible: return "INFEASIBLE";
    case kPresolveUnbndOrInfeas: return "INFEASIBLE_OR_UNBOUNDED";
    default: return "UNKNOWN_STATUS";
    }
}

void presolve(const lp_problem_t *p, Presolved &out)
{
    const auto t0 = std::chrono::steady_clock::now();
    const int m = p->num_constraints, n = p->num_variables;
    const int *rp = p->constraint_matrix_row_pointers;
    const int *ci = p->constraint_matrix_col_indices;
    const double *av = p->constraint_matrix_values;
    out.n = n;
    out.m = m;
    out.col_value.assign((size_t)n, 0.0);

    std::vector<double> lo(p->constraint_lower_bound, p->constraint_lower_bound + m);
    std::vector<double> hi(p->constraint_upper_bound, p->constraint_upper_bound + m);
    std::vector<double> lb(p->variable_lower_bound, p->variable_lower_bound + n);
    std::vector<double> ub(p->variable_upper_bound, p->variable_upper_bound + n);
    const double *c = p->objective_vector;
    double obj_const = p->objective_constant;

    // column-wise copy of A: for each column the (row, value) pairs
    std::vector<int> cp((size_t)n + 1, 0), cr(rp[m]);
    std::vector<double> cv(rp[m]);
    for (int k = 0; k < rp[m]; ++k) cp[ci[k] + 1]++;
    for (int j = 0; j < n; ++j) cp[j + 1] += cp[j];
    {
        std::vector<int> pos(cp.begin(), cp.end() - 1);
        for (int i = 0; i < m; ++i)
            for (int k = rp[i]; k < rp[i + 1]; ++k) {
                cr[pos[ci[k]]] = i;
                cv[pos[ci[k]]++] = av[k];
            }
    }
    std::vector<char> row_on((size_t)m, 1), col_on((size_t)n, 1);
    std::vector<int> rcount((size_t)m), ccount((size_t)n);
    for (int i = 0; i < m; ++i) rcount[i] = rp[i + 1] - rp[i];
    for (int j = 0; j < n; ++j) ccount[j] = cp[j + 1] - cp[j];
    // a stored zero is not a nonzero of the structure
    for (int i = 0; i < m; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (av[k] == 0.0) {
                rcount[i]--;
                ccount[ci[k]]--;
            }

    int status = kPresolveUnchanged;
    auto remove_col = [&](int j, double v, PresolveStep::Kind kind) {
        for (int k = cp[j]; k < cp[j + 1]; ++k) {
            const int i = cr[k];
            if (!row_on[i] || cv[k] == 0.0) continue;
            if (std::isfinite(lo[i])) lo[i] -= cv[k] * v;
            if (std::isfinite(hi[i])) hi[i] -= cv[k] * v;
            rcount[i]--;
        }
        obj_const += c[j] * v;
        col_on[j] = 0;
        out.col_value[j] = v;
        PresolveStep s;
        s.kind = kind;
        s.col = j;
        s.value = v;
        out.steps.push_back(s);
    };
    bool changed = true;
    while (changed && !(status & (kPresolveInfeasible | kPresolveUnbndOrInfeas))) {
        changed = false;
        for (int i = 0; i < m && !(status & kPresolveInfeasible); ++i) {
            if (!row_on[i] || rcount[i] > 1) continue;
            if (rcount[i] == 0) {  // empty row: 0 must lie in [lo, hi]
                if (!le_tol(lo[i], 0.0) || !le_tol(0.0, hi[i])) {
                    status |= kPresolveInfeasible;
                    break;
                }
                row_on[i] = 0;
                PresolveStep s;
                s.kind = PresolveStep::EmptyRow;
                s.row = i;
                out.steps.push_back(s);
                changed = true;
                continue;
            }
            int j = -1;  // singleton row: lo <= a x_j <= hi becomes a bound of x_j
            double a = 0.0;
            for (int k = rp[i]; k < rp[i + 1]; ++k)
                if (col_on[ci[k]] && av[k] != 0.0) {
                    j = ci[k];
                    a = av[k];
                }
            if (j < 0) continue;  // (the count says one; nothing found: keep the row)
            double nl = a > 0.0 ? lo[i] / a : hi[i] / a, nu = a > 0.0 ? hi[i] / a : lo[i] / a;
            if (!std::isfinite(a > 0.0 ? lo[i] : hi[i])) nl = -INFINITY;
            if (!std::isfinite(a > 0.0 ? hi[i] : lo[i])) nu = INFINITY;
            PresolveStep s;
            s.kind = PresolveStep::SingletonRow;
            s.row = i;
            s.col = j;
            s.a = a;
            s.lb0 = lb[j];
            s.ub0 = ub[j];
            if (nl > lb[j]) lb[j] = nl;
            if (nu < ub[j]) ub[j] = nu;
            if (!le_tol(lb[j], ub[j])) {
                status |= kPresolveInfeasible;
                break;
            }
            if (lb[j] > ub[j]) lb[j] = ub[j] = 0.5 * (lb[j] + ub[j]);  // equal within the tolerance
            s.lb1 = lb[j];
            s.ub1 = ub[j];
            out.steps.push_back(s);
            row_on[i] = 0;
            ccount[j]--;
            changed = true;
        }
        if (status & kPresolveInfeasible) break;
        for (int j = 0; j < n; ++j) {
            if (!col_on[j]) continue;
            if (lb[j] == ub[j] && std::isfinite(lb[j])) {  // fixed column
                remove_col(j, lb[j], PresolveStep::FixedCol);
                changed = true;
            } else if (ccount[j] == 0) {  // empty column: the bound the objective prefers
                double v;
                if (c[j] > 0.0) v = lb[j];
                else if (c[j] < 0.0) v = ub[j];
                else v = std::fmin(std::fmax(0.0, lb[j]), ub[j]);
                if (!std::isfinite(v)) {
                    status |= kPresolveUnbndOrInfeas;
                    break;
                }
                remove_col(j, v, PresolveStep::EmptyCol);
                changed = true;
            }
        }
    }

    // the reduced problem from the rows and columns left
    std::vector<int> new_col((size_t)n, -1);
    for (int j = 0; j < n; ++j)
        if (col_on[j]) {
            new_col[j] = (int)out.cols.size();
            out.cols.push_back(j);
        }
    for (int i = 0; i < m; ++i)
        if (row_on[i]) out.rows.push_back(i);
    const int rn = (int)out.cols.size(), rm = (int)out.rows.size();
    if (!(status & (kPresolveInfeasible | kPresolveUnbndOrInfeas))) {
        if (rn < n || rm < m) status |= kPresolveReduced;
        if (rn > 0) {
            lp_problem_t *r = (lp_problem_t *)calloc(1, sizeof(lp_problem_t));
            r->num_variables = rn;
            r->num_constraints = rm;
            r->objective_constant = obj_const;
            std::vector<double> rc(rn), rlb(rn), rub(rn), rlo(rm), rhi(rm);
            for (int t = 0; t < rn; ++t) {
                rc[t] = c[out.cols[t]];
                rlb[t] = lb[out.cols[t]];
                rub[t] = ub[out.cols[t]];
            }
            for (int t = 0; t < rm; ++t) {
                rlo[t] = lo[out.rows[t]];
                rhi[t] = hi[out.rows[t]];
            }
            r->objective_vector = copy_of(rc);
            r->variable_lower_bound = copy_of(rlb);
            r->variable_upper_bound = copy_of(rub);
            r->constraint_lower_bound = copy_of(rlo);
            r->constraint_upper_bound = copy_of(rhi);
            std::vector<int> nrp((size_t)rm + 1, 0), nci;
            std::vector<double> nv;
            for (int t = 0; t < rm; ++t) {
                const int i = out.rows[t];
                for (int k = rp[i]; k < rp[i + 1]; ++k)
                    if (col_on[ci[k]] && av[k] != 0.0) {
                        nci.push_back(new_col[ci[k]]);
                        nv.push_back(av[k]);
                    }
                nrp[t + 1] = (int)nci.size();
            }
            r->constraint_matrix_num_nonzeros = (int)nci.size();
            r->constraint_matrix_row_pointers = (int *)malloc(((size_t)rm + 1) * sizeof(int));
            memcpy(r->constraint_matrix_row_pointers, nrp.data(), ((size_t)rm + 1) * sizeof(int));
            r->constraint_matrix_col_indices = (int *)malloc(std::max<size_t>(nci.size(), 1) * sizeof(int));
            if (!nci.empty()) memcpy(r->constraint_matrix_col_indices, nci.data(), nci.size() * sizeof(int));
            r->constraint_matrix_values = copy_of(nv);
            out.reduced = r;
        }
    }
    out.status = status;
    out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void postsolve(const Presolved &ps, const lp_problem_t *p, const double *x_red, const double *y_red, double *x,
               double *y, double *z)
{
    const int m = ps.m, n = ps.n;
    const int *rp = p->constraint_matrix_row_pointers;
    const int *ci = p->constraint_matrix_col_indices;
    const double *av = p->constraint_matrix_values;
    const double *c = p->objective_vector;
    std::vector<char> y_known((size_t)m, 0);
    for (int j = 0; j < n; ++j) x[j] = ps.col_value[j];
    for (int i = 0; i < m; ++i) y[i] = 0.0;
    if (x_red)
        for (size_t t = 0; t < ps.cols.size(); ++t) x[ps.cols[t]] = x_red[t];
    if (y_red)
        for (size_t t = 0; t < ps.rows.size(); ++t) {
            y[ps.rows[t]] = y_red[t];
            y_known[ps.rows[t]] = 1;
        }
    // column-wise access to the rows a column meets
    std::vector<int> cp((size_t)n + 1, 0), cr(rp[m]);
    std::vector<double> cv(rp[m]);
    for (int k = 0; k < rp[m]; ++k) cp[ci[k] + 1]++;
    for (int j = 0; j < n; ++j) cp[j + 1] += cp[j];
    {
        std::vector<int> pos(cp.begin(), cp.end() - 1);
        for (int i = 0; i < m; ++i)
            for (int k = rp[i]; k < rp[i + 1]; ++k) {
                cr[pos[ci[k]]] = i;
                cv[pos[ci[k]]++] = av[k];
            }
    }
    for (auto it = ps.steps.rbegin(); it != ps.steps.rend(); ++it) {
        const PresolveStep &s = *it;
        switch (s.kind) {
        case PresolveStep::EmptyRow:
            y[s.row] = 0.0;
            y_known[s.row] = 1;
            break;
        case PresolveStep::FixedCol:
        case PresolveStep::EmptyCol:
            x[s.col] = s.value;
            break;
        case PresolveStep::SingletonRow: {
            // the column's reduced cost with the duals known so far; the row takes
            // it over when the bound it implied is the active one
            const int j = s.col;
            double zj = c[j];
            for (int k = cp[j]; k < cp[j + 1]; ++k)
                if (y_known[cr[k]] && cr[k] != s.row) zj -= cv[k] * y[cr[k]];
            const bool at_row_lb = s.lb1 > s.lb0 && x[j] <= s.lb1 && zj > 0.0;
            const bool at_row_ub = s.ub1 < s.ub0 && x[j] >= s.ub1 && zj < 0.0;
            y[s.row] = (at_row_lb || at_row_ub) ? zj / s.a : 0.0;
            y_known[s.row] = 1;
            break;
        }
        }
    }
    // reduced cost of the original problem, clamped as the reference does (presolve.c:208-217)
    for (int j = 0; j < n; ++j) z[j] = c[j];
    for (int i = 0; i < m; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) z[ci[k]] -= av[k] * y[i];
    for (int j = 0; j < n; ++j) {
        if (!std::isfinite(p->variable_lower_bound[j])) z[j] = std::fmin(z[j], 0.0);
        if (!std::isfinite(p->variable_upper_bound[j])) z[j] = std::fmax(z[j], 0.0);
    }
}

}  // namespace lp
}  // namespace pdal
