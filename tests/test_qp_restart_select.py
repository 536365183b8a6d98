"""FA_CP's restart-candidate selection (choose_restart_candidate,
QP/FA_CP/src/fa_cp_core_op.cu:1091-1169; restart to the current iterate,
:908-1020): off by default (FA_CP_RESTART_SELECT unset or 0).  Mode 1 restarts
to the current iterate when its score beats the T output's fixed-point error
by the margin, mode 2 always.

CPU: with selection on, the oracle still reaches the planted optimum of a QP
and the trajectory changes (so the switch is not a no-op); the legacy
variable name works; a zero margin in mode 1 gives the default trajectory.
GPU: on an LP solved through FA_CP (no inner solve, so the device tracks the
oracle through restarts) the device's iterates after 3000 iterations are the
oracle's.
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
    kw.setdefault("iteration_limit", 200000)
    p = lib.default_params(eps_optimal_relative=1e-7, eps_feasible_relative=1e-7, termination_evaluation_frequency=64,
                           **kw)
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


def test_selection_can_stall_an_lp(oracle_qp, monkeypatch):
    """On this LP (no quadratic term) restarting to the current iterate keeps
    the solve from converging (rel KKT ~1e-3 after 2e5 iterations), where the
    default restart converges in about 2.3e4: the option is off by default
    for good reason, and the oracle takes it as it stands."""
    lp, args = _lp()
    base = _solve(oracle_qp, args)
    monkeypatch.setenv("FA_CP_RESTART_SELECT", "1")
    r = _solve(oracle_qp, args, iteration_limit=60000)
    assert base.status == "OPTIMAL" and base.iterations < 60000
    assert r.status == "ITERATION_LIMIT" and r.kkt > 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["1", "2"])
def test_device_selection_matches_oracle(oracle_qp, gpu_qp, monkeypatch, mode):
    """The LP has no inner solve, so the device tracks the oracle through the
    restarts: the iterates after 3000 iterations agree, and differ from the
    default restart's."""
    lp, args = _lp()
    monkeypatch.setenv("FA_CP_RESTART_SELECT", mode)
    o = _solve(oracle_qp, args, iteration_limit=3000)
    g = _solve(gpu_qp, args, iteration_limit=3000)
    assert o.status == g.status == "ITERATION_LIMIT" and o.iterations == g.iterations == 3000
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)  # noqa: E731
    assert rel(g.x, o.x) <= 1e-9 and rel(g.y, o.y) <= 1e-9
    monkeypatch.setenv("FA_CP_RESTART_SELECT", "0")
    d = _solve(oracle_qp, args, iteration_limit=3000)
    assert rel(d.x, o.x) > 1e-6
