"""FA_CP's restart-candidate selection (choose_restart_candidate,
QP/FA_CP/src/fa_cp_core_op.cu:1091-1169; restart to the current iterate,
:908-1020): off by default (FA_CP_RESTART_SELECT unset or 0).  Mode 1 restarts
to the current iterate when its score beats the T output's fixed-point error
by the margin, mode 2 always.

CPU: with selection on the oracle still reaches the planted optimum, and the
trajectory changes (so the switch is not a no-op); mode 2 with a margin has
the margin ignored.  GPU: on an LP solved through FA_CP (no inner solve, so
the device tracks the oracle through restarts) the device stops in the
oracle's block with its status and objective.
"""
from __future__ import annotations

import numpy as np
import pytest

from pdal_amd.generators import planted_lp, planted_qp


def _lp():
    lp = planted_lp(150, 300, 6, seed=8)
    return lp, dict(c=lp.c, A=lp.A, con_lb=lp.b, con_ub=lp.b, var_lb=np.zeros(lp.n), var_ub=np.full(lp.n, np.inf))


def _solve(lib, args, **kw):
    lib.reset_rng()
    p = lib.default_params(eps_optimal_relative=1e-7, eps_feasible_relative=1e-7, iteration_limit=200000,
                           termination_evaluation_frequency=64, **kw)
    return lib.solve(**args, params=p)


@pytest.mark.parametrize("mode", ["1", "2"])
def test_selection_reaches_the_planted_optimum(oracle_qp, monkeypatch, mode):
    q = planted_qp(200, 100, 40, seed=3)
    base = _solve(oracle_qp, q.qp_args("explicit"))
    monkeypatch.setenv("FA_CP_RESTART_SELECT", mode)
    r = _solve(oracle_qp, q.qp_args("explicit"))
    assert r.status == "OPTIMAL"
    assert abs(r.primal_objective - q.optimal_value) <= 1e-6 * abs(q.optimal_value)
    assert r.iterations != base.iterations or not np.array_equal(r.x, base.x)


def test_legacy_variable_and_margin(oracle_qp, monkeypatch):
    lp, args = _lp()
    monkeypatch.setenv("FA_CP_RESTART_SELECT", "1")
    a = _solve(oracle_qp, args)
    monkeypatch.delenv("FA_CP_RESTART_SELECT")
    monkeypatch.setenv("FA_CP_RESTART_SELECT_LEGACY", "1")
    b = _solve(oracle_qp, args)
    assert a.iterations == b.iterations and np.array_equal(a.x, b.x)
    # a margin of 0 never prefers the current iterate in mode 1: the default trajectory
    monkeypatch.setenv("FA_CP_RESTART_SELECT_MARGIN", "0")
    c = _solve(oracle_qp, args)
    monkeypatch.delenv("FA_CP_RESTART_SELECT_LEGACY")
    monkeypatch.delenv("FA_CP_RESTART_SELECT_MARGIN")
    d = _solve(oracle_qp, args)
    assert c.iterations == d.iterations and np.array_equal(c.x, d.x)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["1", "2"])
def test_device_selection_matches_oracle(oracle_qp, gpu_qp, monkeypatch, mode):
    lp, args = _lp()
    monkeypatch.setenv("FA_CP_RESTART_SELECT", mode)
    o = _solve(oracle_qp, args)
    g = _solve(gpu_qp, args)
    assert o.status == g.status == "OPTIMAL" and o.iterations == g.iterations
    assert abs(g.primal_objective - o.primal_objective) <= 1e-9 * abs(o.primal_objective)
    assert abs(g.primal_objective - lp.optimal_value) <= 1e-6 * (1 + abs(lp.optimal_value))
