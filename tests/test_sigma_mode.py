"""Sigma mode of CP_AL (LP/CP_AL/src/solver.cu:3141-3164 parameters, :3106-3126
step schedule, :3664-3671 fixed-point error, :2857-2966 PID on sigma at
restarts), switched on by CP_AL_SIGMA_MODE=1 or CP_AL_SIGMA, the mode of the
reference's random-LP protocol (experiments/random_lp/run_random_lp.sbatch:81-97).

Pins of the oracle (CPU):
  * cross-mode identity: with lambda_param = (2 / step_margin) lambda_est and
    sigma = 1/sqrt(lambda_param), sigma mode's steps (tau = 1/(sigma
    lambda_param), dual step = penalty = sigma) are alpha mode's at alpha = 0.5
    with primal weight 1 -- two different branches of the step schedule give
    the same first evaluation block;
  * zero PID gains: adaptive sigma with k_p = k_i = k_d = 0 never moves sigma,
    so it equals the fixed-sigma run through several restarts;
  * the fixed point: sigma-mode solves reach the planted optimum.
The GPU path is held to the oracle: iterates within 1e-10 through restarts,
the protocol's solves with the same iteration count.
"""
from __future__ import annotations

import numpy as np
import pytest

from pdal_amd.generators import config0, planted_lp

ITER_RTOL = 1e-10
MARGIN = 0.998 * 0.998

PROTOCOL = {  # run_random_lp.sbatch:81-97 ("rhr_cpal")
    "CP_AL_SIGMA_MODE": "1", "CP_AL_SIGMA_ADAPTIVE": "1", "CP_AL_SIGMA_INIT_FACTOR": "0.5",
    "CP_AL_SIGMA_MIN_FACTOR": "1e-4", "CP_AL_SIGMA_MAX_FACTOR": "100", "CP_AL_SIGMA_BALANCE_GAIN": "0",
    "CP_AL_SIGMA_KP": "0.5", "CP_AL_SIGMA_KI": "0", "CP_AL_SIGMA_KD": "0", "CP_AL_SIGMA_UPDATE_MIN": "1e-6",
    "CP_AL_SIGMA_UPDATE_MAX": "1e6", "CP_AL_LAMBDA_FACTOR": "2.01", "CP_AL_LAMBDA": "0",
    "CP_AL_RESTART_SUFFICIENT": "0.2", "CP_AL_RESTART_NECESSARY": "0.8", "CP_AL_RESTART_ARTIFICIAL": "0.36",
}


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def solve(lib, inst, **params):
    lib.reset_rng()
    return lib.solve(**inst.lp_args(), params=lib.default_params(**params))


def test_sigma_schedule_equals_alpha_half(oracle_lp, monkeypatch):
    """One block (the restart at its end is unconditional): the sigma-mode
    schedule with lambda_param = 2 lambda / margin and sigma = 1/sqrt(lambda_param)
    gives alpha mode's iterates at alpha = 0.5."""
    inst = config0(seed=5)
    p = dict(iteration_limit=64, termination_evaluation_frequency=64, eps_optimal_relative=1e-14,
             eps_feasible_relative=1e-14)
    monkeypatch.setenv("CP_AL_LAMBDA", "0.5")
    a = solve(oracle_lp, inst, **p)
    monkeypatch.setenv("CP_AL_SIGMA_MODE", "1")
    monkeypatch.setenv("CP_AL_SIGMA_ADAPTIVE", "0")
    monkeypatch.setenv("CP_AL_LAMBDA_FACTOR", repr(2.0 / MARGIN))
    monkeypatch.setenv("CP_AL_SIGMA_INIT_FACTOR", "1")
    s = solve(oracle_lp, inst, **p)
    assert a.iterations == s.iterations == 64
    assert rel_err(s.x, a.x) <= 1e-12 and rel_err(s.y, a.y) <= 1e-12
    # and the sigma branch is really taken: a different init factor moves the iterates
    monkeypatch.setenv("CP_AL_SIGMA_INIT_FACTOR", "2")
    s2 = solve(oracle_lp, inst, **p)
    assert rel_err(s2.x, a.x) > 1e-6


def test_zero_gain_adaptive_sigma_is_fixed_sigma(oracle_lp, monkeypatch):
    inst = config0(seed=6)
    p = dict(iteration_limit=1600, termination_evaluation_frequency=64, eps_optimal_relative=1e-14,
             eps_feasible_relative=1e-14)
    monkeypatch.setenv("CP_AL_SIGMA_MODE", "1")
    monkeypatch.setenv("CP_AL_SIGMA_ADAPTIVE", "0")
    fixed = solve(oracle_lp, inst, **p)
    monkeypatch.setenv("CP_AL_SIGMA_ADAPTIVE", "1")
    for k in ("CP_AL_SIGMA_KP", "CP_AL_SIGMA_KI", "CP_AL_SIGMA_KD", "CP_AL_SIGMA_BALANCE_GAIN"):
        monkeypatch.setenv(k, "0")
    zero = solve(oracle_lp, inst, **p)
    assert zero.iterations == fixed.iterations
    assert rel_err(zero.x, fixed.x) <= 1e-14 and rel_err(zero.y, fixed.y) <= 1e-14
    monkeypatch.delenv("CP_AL_SIGMA_KP")
    moved = solve(oracle_lp, inst, **p)  # the default gain k_p = 0.99 moves sigma at restarts
    assert rel_err(moved.x, fixed.x) > 1e-9


@pytest.mark.parametrize("env", ["default", "protocol", "fixed"])
def test_oracle_sigma_mode_reaches_planted_optimum(oracle_lp, monkeypatch, env):
    inst = planted_lp(120, 240, 6, seed=31)
    if env == "protocol":
        for k, v in PROTOCOL.items():
            monkeypatch.setenv(k, v)
    elif env == "fixed":
        monkeypatch.setenv("CP_AL_SIGMA", "0.3")
    else:
        monkeypatch.setenv("CP_AL_SIGMA_MODE", "1")
    r = solve(oracle_lp, inst, eps_optimal_relative=1e-8, eps_feasible_relative=1e-8, iteration_limit=400000,
              termination_evaluation_frequency=64)
    assert r.status == "OPTIMAL"
    assert abs(r.primal_objective - inst.optimal_value) <= 1e-6 * (1 + abs(inst.optimal_value))
    assert rel_err(r.x, inst.x_star) <= 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("iters", [64, 320, 1024])
@pytest.mark.parametrize("env", ["default", "protocol"])
def test_gpu_sigma_iterates_track_oracle(oracle_lp, gpu_lp, monkeypatch, iters, env):
    if env == "protocol":
        for k, v in PROTOCOL.items():
            monkeypatch.setenv(k, v)
    else:
        monkeypatch.setenv("CP_AL_SIGMA_MODE", "1")
    inst = config0(seed=7)
    p = dict(iteration_limit=iters, termination_evaluation_frequency=64, eps_optimal_relative=1e-14,
             eps_feasible_relative=1e-14)
    print(f"[case] sigma mode env={env} iters={iters}", flush=True)
    o = solve(oracle_lp, inst, **p)
    g = solve(gpu_lp, inst, **p)
    assert o.iterations == g.iterations == iters
    assert rel_err(g.x, o.x) <= ITER_RTOL and rel_err(g.y, o.y) <= ITER_RTOL


@pytest.mark.gpu
@pytest.mark.parametrize("env", ["protocol", "fixed"])
def test_gpu_sigma_solves_match_oracle(oracle_lp, gpu_lp, monkeypatch, env):
    if env == "protocol":
        for k, v in PROTOCOL.items():
            monkeypatch.setenv(k, v)
    else:
        monkeypatch.setenv("CP_AL_SIGMA", "0.3")
    inst = config0(seed=8)
    p = dict(eps_optimal_relative=1e-6, eps_feasible_relative=1e-6, iteration_limit=400000,
             termination_evaluation_frequency=64)
    print(f"[case] sigma-mode solve env={env}", flush=True)
    o = solve(oracle_lp, inst, **p)
    g = solve(gpu_lp, inst, **p)
    assert o.status == g.status == "OPTIMAL"
    assert g.iterations == o.iterations
    assert abs(g.primal_objective - o.primal_objective) <= 1e-6 * abs(o.primal_objective)
    assert abs(g.primal_objective - inst.optimal_value) <= 1e-5 * (1 + abs(inst.optimal_value))
