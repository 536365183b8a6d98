"""LP presolve of the plain CP_AL path (pdal_amd/csrc/lp/presolve.cpp) behind the
reference's presolve interface (LP/CP_AL/src/presolve.c:61-226, solver.cu:
1183-1194 and 1726-1730).  The reference links PSLP v0.0.4, which is not in the
reference tree, so which reductions run is this build's choice: empty rows,
fixed columns, empty columns and singleton rows, with an exact postsolve.

These CPU tests cover the problems the presolve decides by itself (no device
solve): infeasible rows and bounds, an unbounded empty column, and problems it
reduces to nothing, whose postsolved primal, dual and reduced cost are checked
in numpy (primal feasibility, c - A^T y - z = 0, complementary signs of z and
y at their active bounds) and against HiGHS (scipy.optimize.linprog).  The
device solves of reduced problems are tests/test_presolve_gpu.py.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.optimize as so
import scipy.sparse as sp

INF = np.inf


def solve(lib, A, c, l, u, lb, ub, obj_const=0.0):
    p = lib.default_params(presolve=True)
    return lib.solve(A=sp.csr_matrix(A), c=np.asarray(c, float), con_lb=np.asarray(l, float),
                     con_ub=np.asarray(u, float), var_lb=np.asarray(lb, float), var_ub=np.asarray(ub, float),
                     obj_const=obj_const, params=p)


def kkt_check(A, c, l, u, lb, ub, x, y, z, tol=1e-12):
    A = sp.csr_matrix(A)
    ax = A @ x
    assert np.all(ax >= l - tol) and np.all(ax <= u + tol)
    assert np.all(x >= lb - tol) and np.all(x <= ub + tol)
    np.testing.assert_allclose(c - A.T @ y, z, atol=tol)
    # a positive reduced cost sits at a finite lower bound, a negative one at an upper bound
    assert np.all((z <= tol) | (np.abs(x - lb) <= tol))
    assert np.all((z >= -tol) | (np.abs(x - ub) <= tol))
    # a positive row dual sits at the row's lower bound, a negative one at its upper bound
    assert np.all((y <= tol) | (np.abs(ax - l) <= tol))
    assert np.all((y >= -tol) | (np.abs(ax - u) <= tol))


def highs(A, c, l, u, lb, ub):
    A = sp.csr_matrix(A)
    A_ub = sp.vstack([A[np.isfinite(u)], -A[np.isfinite(l)]]) if A.shape[0] else None
    b_ub = np.concatenate([u[np.isfinite(u)], -l[np.isfinite(l)]]) if A.shape[0] else None
    r = so.linprog(c, A_ub=A_ub, b_ub=b_ub, bounds=list(zip(lb, ub)), method="highs")
    return r


def test_empty_row_outside_its_bounds_is_infeasible(gpu_lp):
    A = sp.csr_matrix(np.array([[1.0, 2.0], [0.0, 0.0]]))
    r = solve(gpu_lp, A, [1, 1], [1, 1], [2, 3], [0, 0], [INF, INF])
    assert r.status == "PRIMAL_INFEASIBLE" and r.extra["presolve_status"] == 2
    assert np.isinf(r.relative_primal_residual)
    # the same row inside its bounds is just dropped (the problem then reaches the device; not here)


def test_conflicting_singleton_rows_are_infeasible(gpu_lp):
    A = np.array([[1.0, 0.0], [2.0, 0.0], [1.0, 1.0]])
    r = solve(gpu_lp, A, [1, 1], [2, -INF, 0], [INF, 2, 5], [0, 0], [INF, INF])  # x1 >= 2 and 2 x1 <= 2
    assert r.status == "PRIMAL_INFEASIBLE" and r.extra["presolve_status"] == 2


def test_unbounded_empty_column(gpu_lp):
    A = np.array([[1.0, 0.0], [1.0, 0.0]])  # column 2 is empty and its cost pulls it to +inf
    r = solve(gpu_lp, A, [1, -1], [1, 1], [1, 1], [0, 0], [INF, INF])
    assert r.status == "INFEASIBLE_OR_UNBOUNDED" and r.extra["presolve_status"] == 4


@pytest.mark.parametrize("seed", range(4))
def test_problem_reduced_to_nothing_matches_highs(gpu_lp, seed):
    """Singleton rows (some tighter than the variable bounds, some looser),
    fixed columns, empty rows and empty columns: the presolve removes
    everything; its x, y, z pass the KKT conditions and its objective is
    HiGHS's."""
    rng = np.random.default_rng(seed)
    n = 12
    rows, l, u = [], [], []
    lb = rng.uniform(-2, 0, n)
    ub = rng.uniform(1, 3, n)
    lb[0] = ub[0] = 0.7           # fixed column
    ub[1] = INF                   # bounded from above only by a row
    lb[2] = -INF
    c = rng.standard_normal(n)
    c[1] = -abs(c[1])             # pulls column 1 up, against its row bound
    c[2] = abs(c[2])              # pulls column 2 down, against its row bound
    for j in range(1, n - 2):     # one singleton row per column 1..n-3 (columns n-2, n-1 stay empty)
        a = rng.choice([-1.0, 1.0]) * rng.uniform(0.5, 2.0)
        lo_j, hi_j = rng.uniform(-1.5, -0.2), rng.uniform(0.2, 1.5)
        row = np.zeros(n)
        row[j] = a
        rows.append(row)
        l.append(a * lo_j if a > 0 else a * hi_j)
        u.append(a * hi_j if a > 0 else a * lo_j)
    rows.append(np.zeros(n))     # an empty row with 0 in its range
    l.append(-1.0)
    u.append(2.0)
    fix = np.zeros(n)            # a row on the fixed column and one other: empty once both are gone
    fix[0], fix[3] = 1.0, 0.0
    rows.append(fix)
    l.append(0.7 - 1.0)
    u.append(0.7 + 1.0)
    A = np.array(rows)
    l, u = np.array(l), np.array(u)
    r = solve(gpu_lp, A, c, l, u, lb, ub, obj_const=1.5)
    assert r.status == "OPTIMAL" and r.extra["presolve_status"] == 1
    assert r.extra["num_reduced_variables"] == 0
    kkt_check(A, c, l, u, lb, ub, r.x, r.y, r.reduced_cost)
    h = highs(A, c, l, u, lb, ub)
    assert h.status == 0
    assert r.primal_objective == pytest.approx(h.fun + 1.5, rel=1e-12, abs=1e-12)
    assert r.dual_objective == r.primal_objective
