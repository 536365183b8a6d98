"""Degenerate LPs on the plain CP_AL path against the oracle and closed forms:
no constraint rows at all (a box LP: every variable at the bound its cost
prefers, a free zero-cost variable at 0), rows without a nonzero (zero
matrix), a single variable, and a zero objective (the primal objective
of every iterate is exactly 0).  An empty matrix skips the power iteration
and takes step scale 1 (LP/CP_AL/src/solver.cu:3538-3544).  Iterates are
held to the oracle within 1e-10 relative, like the other parity tests."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


ITER_RTOL = 1e-10


def rel_err(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def both(oracle_lp, gpu_lp, args, **params):
    out = []
    for lib in (oracle_lp, gpu_lp):
        lib.reset_rng()
        print(f"[case] lib={lib.path.name} m={args['A'].shape[0]} n={args['A'].shape[1]}", flush=True)
        out.append(lib.solve(**args, params=lib.default_params(**params)))
    return out


def test_box_lp_without_rows(oracle_lp, gpu_lp, monkeypatch):
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    c = np.array([1.0, -1.0, 2.0, 0.0, -3.0])
    lb = np.array([0.0, -1.0, -2.0, -np.inf, 0.0])
    ub = np.array([1.0, 2.0, 3.0, np.inf, 4.0])
    args = dict(A=sp.csr_matrix((0, 5)), c=c, con_lb=np.zeros(0), con_ub=np.zeros(0), var_lb=lb, var_ub=ub)
    o, g = both(oracle_lp, gpu_lp, args, eps_optimal_relative=1e-10, eps_feasible_relative=1e-10)
    want = np.where(c > 0, lb, np.where(c < 0, ub, 0.0))
    assert o.status == g.status == "OPTIMAL" and o.iterations == g.iterations
    assert rel_err(g.x, o.x) <= ITER_RTOL
    np.testing.assert_allclose(g.x, want, atol=1e-12)
    assert g.primal_objective == pytest.approx(float(c @ want), abs=1e-12)
    assert g.y.shape == (0,)


def test_rows_without_nonzeros(oracle_lp, gpu_lp, monkeypatch):
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    args = dict(A=sp.csr_matrix((3, 4)), c=np.array([1.0, 2.0, 0.0, 1.0]),
                con_lb=np.array([-1.0, 0.0, -np.inf]), con_ub=np.array([1.0, 0.0, 2.0]),
                var_lb=np.zeros(4), var_ub=np.ones(4))
    o, g = both(oracle_lp, gpu_lp, args, eps_optimal_relative=1e-10, eps_feasible_relative=1e-10)
    assert o.status == g.status == "OPTIMAL" and o.iterations == g.iterations
    assert np.abs(g.x - o.x).max() <= 1e-14 and np.abs(g.y - o.y).max() <= 1e-14
    np.testing.assert_allclose(g.x, np.zeros(4), atol=1e-12)
    assert g.primal_objective == pytest.approx(0.0, abs=1e-12)


def test_single_variable(oracle_lp, gpu_lp, monkeypatch):
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    # min -x s.t. 2 x <= 3, x >= 0: x = 1.5, y = -0.5
    args = dict(A=sp.csr_matrix(np.array([[2.0]])), c=np.array([-1.0]), con_lb=np.array([-np.inf]),
                con_ub=np.array([3.0]), var_lb=np.zeros(1), var_ub=np.array([np.inf]))
    o, g = both(oracle_lp, gpu_lp, args, eps_optimal_relative=1e-12, eps_feasible_relative=1e-12)
    assert o.status == g.status == "OPTIMAL" and o.iterations == g.iterations
    assert g.x[0] == pytest.approx(1.5, abs=1e-10) and g.y[0] == pytest.approx(-0.5, abs=1e-10)
    assert abs(g.x[0] - o.x[0]) <= 1e-14 and abs(g.y[0] - o.y[0]) <= 1e-14


def test_zero_objective(oracle_lp, gpu_lp, monkeypatch):
    from pdal_amd.generators import planted_lp
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    inst = planted_lp(100, 200, 6, seed=9)
    args = dict(inst.lp_args(), c=np.zeros(inst.n))
    o, g = both(oracle_lp, gpu_lp, args, eps_optimal_relative=1e-6, eps_feasible_relative=1e-6,
                termination_evaluation_frequency=64)
    assert o.status == g.status and o.iterations == g.iterations
    assert rel_err(g.x, o.x) <= ITER_RTOL
    assert g.primal_objective == 0.0 and o.primal_objective == 0.0
