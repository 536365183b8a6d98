"""Feasibility polishing of CP_AL (LP/CP_AL/src/solver.cu:3912-4456,
utils.cu:1194-1440): after the main solve, a primal phase (c = 0) and a dual
phase (finite bounds = 0) run the same restarted Halpern iteration until
their relative residual reaches eps_feas_polish_relative.

Pins of the oracle (CPU), independent of its code: the returned x's primal
residual ||Ax - P_[l,u](Ax)|| / (1 + ||b||) and the returned y's dual residual
(the reduced cost's distance to the cone the bounds allow, which the solver's
slack can only exceed) are computed in numpy from the unscaled problem and
must meet the polish tolerance; x stays in its bounds; an already feasible
solve skips polishing.  The GPU path is held to the oracle.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from pdal_amd.generators import planted_lp

ROOT = Path(__file__).resolve().parent.parent


def primal_rel(inst, x):
    ax = inst.A @ x
    res = ax - np.clip(ax, inst.b, inst.b)  # equality rows: l = u = b
    return np.linalg.norm(res) / (1.0 + np.linalg.norm(inst.b))


def dual_rel(inst, y):
    rc = inst.c - inst.A.T @ y
    res = np.minimum(rc, 0.0)  # x >= 0: the slack must be nonnegative
    return np.linalg.norm(res) / (1.0 + np.linalg.norm(inst.c))


def solve(lib, inst, **params):
    lib.reset_rng()
    return lib.solve(**inst.lp_args(), params=lib.default_params(**params))


LOOSE = dict(eps_optimal_relative=1e-3, eps_feasible_relative=1e-3, feasibility_polishing=True,
             eps_feas_polish_relative=1e-8, termination_evaluation_frequency=64, iteration_limit=400000)


@pytest.mark.parametrize("lam", ["0", "0.05"])
def test_oracle_polishing_meets_tolerance(oracle_lp, monkeypatch, capfd, lam):
    monkeypatch.setenv("CP_AL_LAMBDA", lam)
    inst = planted_lp(150, 300, 6, seed=41)
    plain = solve(oracle_lp, inst, **{**LOOSE, "feasibility_polishing": False})
    r = solve(oracle_lp, inst, **LOOSE)
    out = capfd.readouterr().out
    assert "Feasibility Polishing Summary" in out
    assert r.status == plain.status == "OPTIMAL"  # the main solve's reason is kept
    assert r.iterations == plain.iterations
    assert r.extra["feasibility_iteration"] > 0 and r.extra["feasibility_polishing_time"] > 0.0
    assert primal_rel(inst, plain.x) > 1e-6  # the loose solve needed polishing
    assert primal_rel(inst, r.x) <= 1e-8 * 1.0001
    assert dual_rel(inst, r.y) <= 1e-8 * 1.0001
    assert np.all(r.x >= 0.0)
    assert r.relative_primal_residual <= 1e-8 and r.relative_dual_residual <= 1e-8


def test_oracle_polishing_skipped_when_feasible(oracle_lp, monkeypatch, capfd):
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    inst = planted_lp(80, 160, 6, seed=42)
    r = solve(oracle_lp, inst, **{**LOOSE, "eps_optimal_relative": 1e-10, "eps_feasible_relative": 1e-10,
                                   "eps_feas_polish_relative": 1e-6})
    assert "Skipping feasibility polishing" in capfd.readouterr().out
    assert r.extra["feasibility_iteration"] == 0 and r.extra["feasibility_polishing_time"] == 0.0


def test_oracle_polished_objective_drops_the_constant(oracle_lp, monkeypatch):
    """Kept from the reference: the primal phase solves with objective constant
    0, and its objective replaces the main one (solver.cu:4144, 3943-3945)."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    inst = planted_lp(120, 240, 6, seed=43)
    args = dict(inst.lp_args(), obj_const=5.0)
    oracle_lp.reset_rng()
    r = oracle_lp.solve(**args, params=oracle_lp.default_params(**LOOSE))
    assert abs(r.primal_objective - float(inst.c @ r.x)) <= 1e-9 * (1 + abs(r.primal_objective))
    # the dual phase keeps the constant; its objective is that of a feasible, not optimal, pair
    d = r.dual_objective
    assert abs(d - (inst.optimal_value + 5.0)) <= 1e-2 * (1 + abs(inst.optimal_value)) < abs(d - inst.optimal_value)


@pytest.mark.gpu
@pytest.mark.parametrize("lam", ["0", "0.05"])
def test_gpu_polishing_matches_oracle(oracle_lp, gpu_lp, monkeypatch, lam):
    monkeypatch.setenv("CP_AL_LAMBDA", lam)
    inst = planted_lp(150, 300, 6, seed=41)
    print(f"[case] polishing lam={lam}", flush=True)
    o = solve(oracle_lp, inst, **LOOSE)
    g = solve(gpu_lp, inst, **LOOSE)
    assert g.status == o.status == "OPTIMAL" and g.iterations == o.iterations
    assert g.extra["feasibility_iteration"] == o.extra["feasibility_iteration"] > 0
    assert np.linalg.norm(g.x - o.x) <= 1e-7 * np.linalg.norm(o.x)
    assert np.linalg.norm(g.y - o.y) <= 1e-7 * np.linalg.norm(o.y)
    assert primal_rel(inst, g.x) <= 1e-8 * 1.0001 and dual_rel(inst, g.y) <= 1e-8 * 1.0001
    for f in ("primal_objective", "dual_objective"):
        assert abs(getattr(g, f) - getattr(o, f)) <= 1e-9 * (1 + abs(getattr(o, f)))


@pytest.mark.gpu
def test_cli_feasibility_polishing_flag(tmp_path):
    """`cp_al -f` writes the polishing lines of the summary (cli.c:578-583)."""
    from test_lp_cli import write_mps
    inst = planted_lp(100, 200, 6, seed=44)
    mps = tmp_path / "p.mps"
    write_mps(mps, inst.A, inst.c, inst.b, inst.b, np.zeros(inst.n), np.full(inst.n, np.inf))
    out = tmp_path / "out"
    out.mkdir()
    exe = ROOT / "pdal_amd" / "lib" / "cp_al"
    r = subprocess.run([str(exe), "-f", "--eps_opt", "1e-3", "--eps_feas", "1e-3", "--eps_feas_polish", "1e-8",
                        "--no_presolve", str(mps), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    text = (out / "p_summary.txt").read_text()
    assert "Feasibility Polishing Time (sec):" in text
    assert "Feasibility Polishing Iteration Count:" in text
