"""LP presolve with device solves of the reduced problem (see tests/test_presolve.py
for the reductions and the problems the presolve decides alone).

A planted LP is extended with the structures the presolve removes: columns
fixed at their planted value, empty columns with a positive cost, rows that
hold no nonzero, and singleton rows that bound a column on both sides around
its planted value.  The presolved device solve must report the reduced sizes,
reach the planted optimum within the bar the KKT test implies and the oracle's
objective on the unreduced problem within 1e-6, and its postsolved x, y,
reduced cost must satisfy the original problem's KKT conditions (recomputed
in numpy).  A problem the presolve leaves unchanged solves to the same bits as
with presolve off.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from pdal_amd.generators import kkt_objective_bar, planted_lp

pytestmark = pytest.mark.gpu


def extended(seed):
    inst = planted_lp(300, 600, 8, seed=seed)
    rng = np.random.default_rng(seed + 77)
    A = sp.csr_matrix(inst.A)
    m, n = A.shape
    xs, ys = inst.x_star.copy(), inst.y_star.copy()
    c = inst.c.copy()
    lb, ub = np.zeros(n), np.full(n, np.inf)
    fixed = rng.choice(n, 20, replace=False)
    lb[fixed] = ub[fixed] = xs[fixed]
    # two empty columns with positive cost (optimal at 0), two empty rows with 0 inside their range
    A = sp.hstack([A, sp.csr_matrix((m, 2))]).tocsr()
    c = np.concatenate([c, [0.5, 1.5]])
    lb, ub = np.concatenate([lb, [0.0, 0.0]]), np.concatenate([ub, [np.inf, np.inf]])
    xs = np.concatenate([xs, [0.0, 0.0]])
    rows = [sp.csr_matrix((2, n + 2))]
    l, u = [inst.b, np.array([-1.0, 0.0])], [inst.b, np.array([1.0, 3.0])]
    # singleton rows bounding 15 other columns on both sides around their planted values (inactive)
    free_cols = np.setdiff1d(np.arange(n), fixed)
    sing = rng.choice(free_cols, 15, replace=False)
    a = rng.choice([-1.0, 1.0], 15) * rng.uniform(0.5, 2.0, 15)
    S = sp.csr_matrix((a, (np.arange(15), sing)), shape=(15, n + 2))
    rows.append(S)
    lo_x, hi_x = xs[sing] - 0.5, xs[sing] + 0.5
    l.append(np.where(a > 0, a * lo_x, a * hi_x))
    u.append(np.where(a > 0, a * hi_x, a * lo_x))
    A_full = sp.vstack([A] + rows).tocsr()
    l_full, u_full = np.concatenate(l), np.concatenate(u)
    ys_full = np.concatenate([ys, np.zeros(2 + 15)])
    opt = float(c @ xs)
    return dict(A=A_full, c=c, con_lb=l_full, con_ub=u_full, var_lb=lb, var_ub=ub), xs, ys_full, opt, \
        len(fixed), 2, 2, 15


@pytest.mark.parametrize("seed", [0, 1])
def test_presolved_solve(oracle_lp, gpu_lp, monkeypatch, seed):
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    args, xs, ys, opt, n_fixed, n_empty_cols, n_empty_rows, n_sing = extended(seed)
    A, c, l, u, lb, ub = (args[k] for k in ("A", "c", "con_lb", "con_ub", "var_lb", "var_ub"))
    m, n = A.shape
    tol = 1e-8
    print(f"[case] presolved solve seed={seed} m={m} n={n}", flush=True)
    gpu_lp.reset_rng()
    r = gpu_lp.solve(**args, params=gpu_lp.default_params(presolve=True, eps_optimal_relative=tol,
                                                          eps_feasible_relative=tol))
    assert r.status == "OPTIMAL"
    assert r.extra["presolve_status"] == 1 and r.extra["presolve_time"] > 0
    assert r.extra["num_reduced_variables"] == n - n_fixed - n_empty_cols
    assert r.extra["num_reduced_constraints"] == m - n_empty_rows - n_sing
    # the postsolved triple on the original problem
    x, y, z = r.x, r.y, r.reduced_cost
    assert x.shape == (n,) and y.shape == (m,)
    np.testing.assert_array_equal(x[lb == ub], lb[lb == ub])
    ax = A @ x
    viol = np.maximum(l - ax, 0) + np.maximum(ax - u, 0)
    assert np.linalg.norm(viol) <= 10 * tol * (1 + np.linalg.norm(np.where(np.isfinite(u), u, 0)))
    assert x.min() >= -1e-12
    lam = c - A.T @ y
    np.testing.assert_allclose(z, np.maximum(lam, 0.0) * (ub == np.inf) + lam * (ub < np.inf), atol=1e-12)
    assert np.linalg.norm(np.minimum(lam, 0.0)[ub == np.inf]) <= 10 * tol * (1 + np.linalg.norm(c))
    # objectives: planted (through the KKT bar), and the oracle on the unreduced problem
    bar = kkt_objective_bar(r.primal_objective, r.dual_objective, xs, ys, l[np.isfinite(l)], c, tol)
    assert abs(r.primal_objective - opt) <= bar
    assert abs(float(c @ x) - r.primal_objective) <= 1e-9 * abs(opt)
    oracle_lp.reset_rng()
    o = oracle_lp.solve(**args, params=oracle_lp.default_params(eps_optimal_relative=tol,
                                                                eps_feasible_relative=tol))
    assert o.status == "OPTIMAL"
    assert abs(r.primal_objective - o.primal_objective) <= 1e-6 * abs(o.primal_objective)


def test_unchanged_problem_is_the_plain_solve(gpu_lp, monkeypatch):
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    monkeypatch.setenv("PDAL_SPMV_G_A", "8")  # lane counts pinned: the autotuner picks by timing
    monkeypatch.setenv("PDAL_SPMV_G_AT", "4")
    inst = planted_lp(300, 600, 8, seed=4)
    out = []
    for pre in (True, False):
        gpu_lp.reset_rng()
        out.append(gpu_lp.solve(**inst.lp_args(), params=gpu_lp.default_params(
            presolve=pre, eps_optimal_relative=1e-8, eps_feasible_relative=1e-8)))
    a, b = out
    assert a.extra["presolve_status"] == 0 and b.extra["presolve_time"] == 0
    assert a.iterations == b.iterations
    np.testing.assert_array_equal(a.x, b.x)
    np.testing.assert_array_equal(a.y, b.y)
    np.testing.assert_allclose(a.reduced_cost, b.reduced_cost, rtol=0, atol=1e-12 * np.linalg.norm(inst.c))
