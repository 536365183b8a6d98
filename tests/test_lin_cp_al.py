"""RHR-Lin-CP-AL, the linearized CP-AL QP solver (QP/LIN_CP_AL): one projected
gradient step per iteration in place of FA_CP's inner solve, a fixed penalty
sigma in [sigma_base, 100 sigma_base], equality constraints A x = b (b = the
constraint lower bound).  librhr_lin_cp_al.so / lin_cp_al.h.

Pins of the oracle (CPU), independent of its code: the solves reach the
planted optimum of an equality LP and of an equality-constrained convex QP
(explicit Q and the low-rank form), in both the AL and the non-augmented
branch, and every result satisfies A x = b and the bounds to the tolerance.
Kept from the reference: the Q norm is only estimated for sparse and diagonal
Q (LIN_CP_AL solver_state.cu:443-474; for the low-rank forms and LPs the field
is left unset there and is 0 here), so the non-augmented branch steps a
low-rank Q without its curvature and does not converge on it -- the test
records that it stops at the iteration limit.  The GPU path is held to the
oracle: iterates within 1e-10 of it through restarts, the same iteration count
and objective on full solves.
"""
from __future__ import annotations

import numpy as np
import pytest

from pdal_amd.generators import planted_lp, planted_qp
from pdal_amd.qp import LinCpAl

ROOT_ORACLE = "oracle/_build/liblin_cp_al_oracle.so"
TOL = dict(eps_optimal_relative=1e-6, eps_feasible_relative=1e-6, iteration_limit=300000)


@pytest.fixture(scope="module")
def oracle_lin():
    from pathlib import Path
    return LinCpAl(Path(__file__).resolve().parent.parent / ROOT_ORACLE)


def lp_case():
    lp = planted_lp(80, 160, 6, seed=3)
    return lp, dict(c=lp.c, A=lp.A, con_lb=lp.b, con_ub=lp.b, var_lb=np.zeros(lp.n), var_ub=np.full(lp.n, np.inf))


def qp_case(form):
    qp = planted_qp(120, 60, 30, seed=5)
    return qp, qp.qp_args(form, equality=True)


def solve(lib, args, **params):
    lib.reset_rng()
    return lib.solve(**args, params=lib.default_params(**params))


def feasible(args, x, tol):
    ax = args["A"] @ x
    b = args["con_lb"]
    assert np.linalg.norm(ax - b, np.inf) <= tol * (1 + np.linalg.norm(b, np.inf))
    assert np.all(x >= args["var_lb"] - 1e-12) and np.all(x <= args["var_ub"] + 1e-12)


@pytest.mark.parametrize("al", [True, False])
def test_oracle_lp_reaches_planted_optimum(oracle_lin, al):
    lp, args = lp_case()
    r = solve(oracle_lin, args, use_al_qp=al, **TOL)
    assert r.status == "OPTIMAL" and r.inner_iterations == 0
    assert abs(r.primal_objective - lp.optimal_value) <= 1e-5 * (1 + abs(lp.optimal_value))
    assert np.linalg.norm(r.x - lp.x_star) <= 1e-3 * np.linalg.norm(lp.x_star)
    feasible(args, r.x, 1e-5)


@pytest.mark.parametrize("form,al", [("explicit", True), ("explicit", False), ("lowrank", True)])
def test_oracle_qp_reaches_planted_optimum(oracle_lin, form, al):
    qp, args = qp_case(form)
    r = solve(oracle_lin, args, use_al_qp=al, **TOL)
    assert r.status == "OPTIMAL"
    assert abs(r.primal_objective - qp.optimal_value) <= 1e-6 * (1 + abs(qp.optimal_value))
    assert np.linalg.norm(r.x - qp.x_star) <= 1e-3 * np.linalg.norm(qp.x_star)
    feasible(args, r.x, 1e-5)


def test_oracle_lowrank_without_al_steps_blind(oracle_lin):
    """No curvature estimate for a low-rank Q: the non-augmented step ignores it."""
    qp, args = qp_case("lowrank")
    r = solve(oracle_lin, args, use_al_qp=False, **{**TOL, "iteration_limit": 20000})
    assert r.status == "ITERATION_LIMIT"


def test_oracle_defaults_and_sigma_init(oracle_lin):
    """Defaults of LIN_CP_AL utils.cu:466-497; a sigma_init inside the safe
    interval changes the trajectory, one above it is clipped to 100 sigma_base
    (two larger requests give the same iterates)."""
    p = oracle_lin.default_params()
    assert p.reflection_coefficient == 0.0 and p.use_al_qp and p.sigma_init == 0.0
    assert p.termination_evaluation_frequency == 200 and p.optimality_norm == 1
    assert p.restart_params.sufficient_reduction_for_restart == 0.2
    lp, args = lp_case()
    base = solve(oracle_lin, args, iteration_limit=300)
    big1 = solve(oracle_lin, args, iteration_limit=300, sigma_init=1e6)
    big2 = solve(oracle_lin, args, iteration_limit=300, sigma_init=1e9)
    assert np.array_equal(big1.x, big2.x)
    assert not np.allclose(base.x, big1.x)
