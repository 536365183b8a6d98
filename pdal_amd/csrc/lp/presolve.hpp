// LP presolve and postsolve of the plain CP_AL path (host code).
//
// Reference interface: LP/CP_AL/src/presolve.c:61-120 (pslp_presolve: the
// reduced problem, or "solved during presolve" when it is infeasible or has
// no columns left), :122-181 (create_result_from_presolve), :183-226
// (pslp_postsolve: the original-size x, y and reduced cost, reduced cost
// clamped at 0 on infinite bounds), called from solver.cu:1183-1194 and
// :1726-1730.  The reference links the third-party PSLP v0.0.4, which is not
// part of the reference tree; the reductions here are this build's own, the
// classical ones whose postsolve is exact:
//   * empty rows (infeasible unless 0 lies in [l, u]; dual 0);
//   * fixed columns (lb = ub: substituted into the row bounds and the
//     objective constant);
//   * empty columns (set to the bound the objective prefers; an infinite one
//     makes the problem unbounded or infeasible);
//   * singleton rows (turned into bounds of their column; the dual of the
//     row is the column's reduced cost when the bound the row implied is the
//     active one).
// They are applied until none applies.  Status values are bit flags, as the
// reference tests them (`status & INFEASIBLE`).
#pragma once

#include <vector>

#include "cp_al_types.h"

namespace pdal {
namespace lp {

enum PresolveStatus { kPresolveUnchanged = 0, kPresolveReduced = 1, kPresolveInfeasible = 2, kPresolveUnbndOrInfeas = 4 };

struct PresolveStep {  // one reduction, undone in reverse order
    enum Kind { EmptyRow, SingletonRow, FixedCol, EmptyCol } kind;
    int row = -1, col = -1;
    double a = 0.0;              // SingletonRow: the row's coefficient
    double value = 0.0;          // FixedCol / EmptyCol: the column's value
    double lb0 = 0.0, ub0 = 0.0; // SingletonRow: the column's bounds before the row's
    double lb1 = 0.0, ub1 = 0.0; // SingletonRow: the column's bounds after it
};

struct Presolved {
    int status = kPresolveUnchanged;
    double seconds = 0.0;
    lp_problem_t *reduced = nullptr;  // owned; null when nothing is left to solve
    std::vector<int> rows, cols;      // reduced index -> original index
    std::vector<PresolveStep> steps;
    std::vector<double> col_value;    // values of removed columns (original index)
    int n = 0, m = 0;                 // original sizes
    Presolved() = default;
    Presolved(const Presolved &) = delete;
    Presolved &operator=(const Presolved &) = delete;
    ~Presolved();
};

// Reduce `p` (not modified).
void presolve(const lp_problem_t *p, Presolved &out);

// Original-size solution from the reduced one (x_red, y_red may be null when
// no reduced problem was solved): x, y, reduced cost z = c - A^T y with the
// reference's clamping on infinite bounds.  Arrays hold n, m, n doubles.
void postsolve(const Presolved &ps, const lp_problem_t *p, const double *x_red, const double *y_red, double *x,
               double *y, double *z);

const char *presolve_status_name(int status);

}  // namespace lp
}  // namespace pdal
