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


# ---------------------------------------------------------------------------
# The sigma PID at the first restart, through the closed-form iteration map of
# an equality LP with free variables (every step affine).


def _orthogonal_rows(m, n, s, seed):
    Q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((n, n)))
    return np.sqrt(s) * Q[:m, :]  # A A^T = s I: the power iteration is exact at its first step


def _sigma_T(A, b, c, tau, sig):
    """One sigma-mode step as z -> M z + d on z = (x, y): the box-AL predictor
    u = y - sig (A x - b) (equality rows, penalty = sig), x+ = x - tau (c - A^T u)
    (free variables), y+ = y - sig (A (2 x+ - x) - b) (solver.cu:2459-2479,
    2340-2457 with l = u = b)."""
    m, n = A.shape
    I_n, I_m = np.eye(n), np.eye(m)
    Mx = np.hstack([I_n - tau * sig * A.T @ A, tau * A.T])
    dx = -tau * c + tau * sig * A.T @ b
    Mxr = 2 * Mx - np.hstack([I_n, np.zeros((n, m))])  # 2 x+ - x
    dxr = 2 * dx
    My = np.hstack([np.zeros((m, n)), I_m]) - sig * A @ Mxr
    dy = -sig * A @ dxr + sig * b
    return np.vstack([Mx, My]), np.concatenate([dx, dy])


def _halpern_T_output(M, d, z0, K):
    """T(z_{K-1}) with z_{K-1} = z* + (1/K) sum_{j<K} (2M - I)^j (z0 - z*)."""
    N = M.shape[0]
    zs, *_ = np.linalg.lstsq(np.eye(N) - M, d, rcond=None)
    assert np.linalg.norm((np.eye(N) - M) @ zs - d) < 1e-10
    R = 2 * M - np.eye(N)
    v, acc = z0 - zs, np.zeros(N)
    for _ in range(K):
        acc += v
        v = R @ v
    return M @ (zs + acc / K) + d


@pytest.mark.parametrize("case", ["ratio_free_balance_clipped", "balance_free_wide_clip", "ratio_clipped"])
def test_sigma_pid_first_restart_closed_form(oracle_lp, monkeypatch, case):
    """Block 1 restarts unconditionally (total == K, utils.cu:336-339) to its
    T-output; the sigma PID (solver.cu:2857-2966) computes, from the restart
    displacement (d_x, d_y), ||A d_x||, lambda_param and the residual ratio,
    err = ln||d_y|| - ln(lambda_param ||d_x||^2 - ||A d_x||^2)/2 - ln sigma
    + 0.15 clip(ln(r_d/r_p)/2, -2, 2) and sigma <- sigma clip(exp((k_p + k_i) err),
    lo, hi) (lo, hi = 0.6, 1.7 unless set); block 2 is again a Halpern
    polynomial from the new anchor with the new steps.  Every quantity here
    comes from the closed-form map, not from the oracle, except the two
    residuals of block 1.  The three cases put the PID step and the balance
    term on each side of their clips."""
    import scipy.sparse as sp
    monkeypatch.setenv("CP_AL_SIGMA_MODE", "1")
    m, n, s, K, seed, f = {"ratio_free_balance_clipped": (9, 15, 1.3, 21, 8, 2.0),
                           "balance_free_wide_clip": (12, 30, 0.7, 9, 3, 3.0),
                           "ratio_clipped": (9, 15, 1.3, 21, 8, 6.0)}[case]
    lo, hi = (1e-6, 1e6) if case == "balance_free_wide_clip" else (0.6, 1.7)
    monkeypatch.setenv("CP_AL_SIGMA_INIT_FACTOR", repr(f))
    if case == "balance_free_wide_clip":
        monkeypatch.setenv("CP_AL_SIGMA_UPDATE_MIN", repr(lo))
        monkeypatch.setenv("CP_AL_SIGMA_UPDATE_MAX", repr(hi))
    A = _orthogonal_rows(m, n, s, seed=seed)
    rng = np.random.default_rng(seed + 1)
    b = rng.standard_normal(m)
    c = A.T @ rng.standard_normal(m)
    lam_param = 2.01 * s
    sigma = f / np.sqrt(lam_param)  # init factor x sigma_base (solver.cu:3141-3164)
    common = dict(l_inf_ruiz_iterations=0, has_pock_chambolle_alpha=False, bound_objective_rescaling=False,
                  termination_evaluation_frequency=K, eps_optimal_relative=1e-15, eps_feasible_relative=1e-15)
    oracle_lp.reset_rng()
    r1 = oracle_lp.solve(A=sp.csr_matrix(A), c=c, con_lb=b, con_ub=b,
                         params=oracle_lp.default_params(iteration_limit=K, **common))
    M1, d1 = _sigma_T(A, b, c, 1.0 / (sigma * lam_param), sigma)
    z1 = _halpern_T_output(M1, d1, np.zeros(n + m), K)
    np.testing.assert_allclose(np.concatenate([r1.x, r1.y]), z1, atol=1e-10 * np.linalg.norm(z1))
    dxn, dyn = np.linalg.norm(z1[:n]), np.linalg.norm(z1[n:])
    axn = np.linalg.norm(A @ z1[:n])
    ratio_infeas = r1.relative_dual_residual / r1.relative_primal_residual
    assert 1e-8 < ratio_infeas < 1e8
    balance = 0.5 * np.log(ratio_infeas)
    err = np.log(dyn) - 0.5 * np.log(lam_param * dxn ** 2 - axn ** 2) - np.log(sigma)
    err += 0.15 * np.clip(balance, -2.0, 2.0)
    step = np.exp(0.99 * err + 0.01 * err)
    if case == "ratio_free_balance_clipped":
        assert balance < -2.0 and lo < step < hi
    elif case == "balance_free_wide_clip":
        assert -2.0 < balance < -1.0 and not 0.6 < step < 1.7
    else:
        assert step < lo
    sigma2 = np.clip(sigma * np.clip(step, lo, hi), 1e-4 / np.sqrt(lam_param), 100.0 / np.sqrt(lam_param))
    oracle_lp.reset_rng()
    r2 = oracle_lp.solve(A=sp.csr_matrix(A), c=c, con_lb=b, con_ub=b,
                         params=oracle_lp.default_params(iteration_limit=2 * K, **common))
    M2, d2 = _sigma_T(A, b, c, 1.0 / (sigma2 * lam_param), sigma2)
    z2 = _halpern_T_output(M2, d2, z1, K)
    np.testing.assert_allclose(np.concatenate([r2.x, r2.y]), z2, atol=1e-9 * np.linalg.norm(z2))
    # a sigma off by 1 % lands ten times the tolerance away: the check resolves the update
    z2_off = _halpern_T_output(*_sigma_T(A, b, c, 1.0 / (1.01 * sigma2 * lam_param), 1.01 * sigma2), z1, K)
    assert np.linalg.norm(z2_off - z2) > 1e-8 * np.linalg.norm(z2)
