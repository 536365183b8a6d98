"""Feasibility polishing of FA_CP (QP/FA_CP/src/fa_cp_core_op.cu:1242-1735,
utils.cu:1360-1575): a primal phase (c = 0, Q kept) and a dual phase (finite
bounds = 0) after the main solve, each iterating until its relative residual
reaches eps_feas_polish_relative.

Pins of the oracle (CPU), computed in numpy on the unscaled problem, on an
LP-type instance (x >= 0, Q = 0) after a 1e-3 main solve that did not meet
them: the returned x's primal residual ||Ax - P_[l,u](Ax)|| / (1 + ||b||), its
bounds, and the returned y's dual residual ||min(c - A'y, 0)|| / (1 + ||c||).

With Q != 0 the reference's phases cannot succeed, and this build keeps that:
the dual phase's residual is c - A'y - slack without the Q x term
(utils.cu:1472-1489) and the slack is written only by the LP primal kernel and
the main residual (fa_cp_core_op.cu:110-120, utils.cu:741-768), so it stays
at about ||Q x||; in the primal phase a restart resets the Halpern counter of
both sides while only the primal anchor moves, pulling y back to its zero
anchor.  Both phases run to the iteration limit and the main solution is
returned unchanged -- which the explicit-Q test checks.  The GPU path is held
to the oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

from pdal_amd.generators import planted_lp, planted_qp

POLISH = dict(eps_optimal_relative=1e-3, eps_feasible_relative=1e-3, feasibility_polishing=True,
              eps_feas_polish_relative=1e-8, termination_evaluation_frequency=64, iteration_limit=200000,
              optimality_norm="l2")


def bound_norm(l, u):
    """constraint_bound_norm of the working problem (L2): finite lower bounds
    that differ from the upper one, and finite upper bounds."""
    lo = np.where(np.isfinite(l) & (l != u), l, 0.0)
    hi = np.where(np.isfinite(u), u, 0.0)
    return float(np.sqrt(np.sum(lo * lo) + np.sum(hi * hi)))


def primal_rel(args, x):
    ax = args["A"] @ x
    res = ax - np.clip(ax, args["con_lb"], args["con_ub"])
    return float(np.linalg.norm(res) / (1.0 + bound_norm(args["con_lb"], args["con_ub"])))


def qp_case():
    inst = planted_qp(200, 100, 40, seed=15)
    return inst.qp_args("explicit")


def lp_case():
    lp = planted_lp(120, 240, 6, seed=16)
    return dict(c=lp.c, A=lp.A, con_lb=lp.b, con_ub=lp.b, var_lb=np.zeros(lp.n), var_ub=np.full(lp.n, np.inf))


def solve(lib, args, **params):
    lib.reset_rng()
    return lib.solve(**args, params=lib.default_params(**params))


def test_oracle_qp_polishing_meets_tolerance(oracle_qp, capfd):
    args = lp_case()
    plain = solve(oracle_qp, args, **{**POLISH, "feasibility_polishing": False})
    r = solve(oracle_qp, args, **POLISH)
    assert "Feasibility Polishing Summary" in capfd.readouterr().out
    assert r.status == plain.status == "OPTIMAL" and r.iterations == plain.iterations
    assert r.feasibility_iteration > 0 and r.feasibility_polishing_time > 0.0
    assert primal_rel(args, plain.x) > 1e-7  # the loose solve needed polishing
    assert primal_rel(args, r.x) <= 1e-8 * 1.0001
    assert np.all(r.x >= args["var_lb"]) and np.all(r.x <= args["var_ub"])
    rc = args["c"] - args["A"].T @ r.y
    assert np.linalg.norm(np.minimum(rc, 0.0)) / (1.0 + np.linalg.norm(args["c"])) <= 1e-8 * 1.0001


LIMIT = 4000


def test_oracle_qp_polishing_with_q_keeps_the_solution(oracle_qp):
    params = {**POLISH, "iteration_limit": LIMIT}
    args = qp_case()
    plain = solve(oracle_qp, args, **{**params, "feasibility_polishing": False})
    r = solve(oracle_qp, args, **params)
    assert r.status == plain.status == "OPTIMAL" and r.iterations == plain.iterations
    # each phase stops at the print point total == LIMIT, runs the loop body once
    # more (total = LIMIT + 1) and is counted as total - 1
    assert r.feasibility_iteration == 2 * LIMIT
    assert np.array_equal(r.x, plain.x) and np.array_equal(r.y, plain.y)
    assert r.relative_primal_residual == plain.relative_primal_residual


def test_oracle_qp_polishing_skipped_when_feasible(oracle_qp, capfd):
    r = solve(oracle_qp, lp_case(), **{**POLISH, "eps_optimal_relative": 1e-10, "eps_feasible_relative": 1e-10,
                                       "eps_feas_polish_relative": 1e-6})
    assert "Skipping feasibility polishing" in capfd.readouterr().out
    assert r.feasibility_iteration == 0 and r.feasibility_polishing_time == 0.0


@pytest.mark.gpu
def test_gpu_qp_polishing_lp_case_matches_oracle(oracle_qp, gpu_qp):
    args = lp_case()
    print("[case] FA_CP polishing, LP-type instance", flush=True)
    o = solve(oracle_qp, args, **POLISH)
    g = solve(gpu_qp, args, **POLISH)
    assert g.status == o.status == "OPTIMAL" and g.iterations == o.iterations
    assert g.feasibility_iteration == o.feasibility_iteration > 0
    assert np.linalg.norm(g.x - o.x) <= 1e-7 * np.linalg.norm(o.x)
    assert np.linalg.norm(g.y - o.y) <= 1e-7 * np.linalg.norm(o.y)
    assert primal_rel(args, g.x) <= 1e-8 * 1.0001


@pytest.mark.gpu
def test_gpu_qp_polishing_with_q(oracle_qp, gpu_qp):
    """Explicit Q (the projected-BB inner solve): both phases run to the limit
    on the GPU as in the oracle, and the main solution comes back unchanged;
    after the main solve's restarts the BB path is rounding-sensitive, so the
    two main solutions are compared at their own tolerance."""
    params = {**POLISH, "iteration_limit": LIMIT}
    args = qp_case()
    print("[case] FA_CP polishing, explicit-Q QP", flush=True)
    o = solve(oracle_qp, args, **params)
    g = solve(gpu_qp, args, **params)
    plain = solve(gpu_qp, args, **{**params, "feasibility_polishing": False})
    assert g.status == o.status == "OPTIMAL"
    assert g.feasibility_iteration == o.feasibility_iteration == 2 * LIMIT
    assert np.array_equal(g.x, plain.x) and np.array_equal(g.y, plain.y)
    # two 1e-3 solutions of a rounding-sensitive trajectory (measured: 2e-5 apart)
    assert np.linalg.norm(g.x - o.x) <= 1e-3 * np.linalg.norm(o.x)
    assert abs(g.primal_objective - o.primal_objective) <= 1e-3 * (1 + abs(o.primal_objective))
