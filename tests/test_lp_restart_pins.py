"""Closed-form and log-consistency pins of two CP_AL controllers on the oracle (CPU).

Restart-candidate selection (LP/CP_AL/src/solver.cu:2663-2819, restart to the
chosen point :3019-3076).  On an equality LP with free variables and AA^T = sI
(scaling off) the iteration is affine, so after one block of K iterations the
three candidates are known in closed form: the T output x_T = T(z_{K-1}), the
current (reflected Halpern) iterate z_K and the Halpern point
(K x_T + z_0)/(K + 1), with z_i = z* + (1/(i+1)) sum_{j<=i} R^j (z_0 - z*).
Each candidate's score is the fixed-point metric of (x_T - candidate):
sqrt(pw |dx|^2 + |dy|^2 / pw + 2 eta <dx, A^T dy>).  The printed scores must
be those, the chosen target the smallest, and the second block must be the
Halpern polynomial started from the chosen point with the PID primal weight of
(chosen - 0).

Positive-alpha trial's acceptance gate (solver.cu:777-861): the decision line
must follow from the two zero-alpha epochs' printed rates and the trial's
printed start metrics (prediction = start x exp(-mean rate x K); accept iff fp
<= 0.90, KKT <= 0.95 and each of p, d, g <= 1.05 times its prediction).
"""
from __future__ import annotations

import ctypes
import re

import numpy as np
import pytest
import scipy.sparse as sp

from pdal_amd.generators import planted_lp

LIBC = ctypes.CDLL(None)


def _orthogonal_rows(m, n, s, seed):
    Q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((n, n)))
    return np.sqrt(s) * Q[:m, :]


def _affine(A, b, c, tau, sig):
    """z -> M z + d: one primal-dual step of the free-variable equality LP (T)."""
    m, n = A.shape
    M = np.block([[np.eye(n), tau * A.T], [-sig * A, np.eye(m) - 2 * sig * tau * A @ A.T]])
    d = np.concatenate([-tau * c, 2 * sig * tau * A @ c + sig * b])
    return M, d


def _block(M, d, z0, K):
    """(z_{K-1}, T(z_{K-1}), z_K) of K reflected-Halpern steps from z0 (reflection 1)."""
    N = M.shape[0]
    R = 2 * M - np.eye(N)
    zs = np.linalg.lstsq(np.eye(N) - M, d, rcond=None)[0]
    v = z0 - zs
    terms = [v]
    for _ in range(K):
        terms.append(R @ terms[-1])
    z_km1 = zs + sum(terms[:K]) / K
    z_k = zs + sum(terms[:K + 1]) / (K + 1)
    return z_km1, M @ z_km1 + d, z_k


def _metric(A, pw, eta, dz):
    n = A.shape[1]
    dx, dy = dz[:n], dz[n:]
    return np.sqrt(pw * dx @ dx + dy @ dy / pw + 2 * eta * dx @ (A.T @ dy))


@pytest.mark.parametrize("mode", ["2", "3"])
def test_restart_candidates_in_closed_form(oracle_lp, monkeypatch, capfd, mode):
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    monkeypatch.setenv("CP_AL_RESTART_SELECT", mode)
    monkeypatch.setenv("CP_AL_RESTART_SELECT_TRACE", "1")
    m, n, s, K = 10, 16, 1.7, 25
    A = _orthogonal_rows(m, n, s, seed=3)
    rng = np.random.default_rng(4)
    b = rng.standard_normal(m)
    c = A.T @ rng.standard_normal(m)
    eta = np.sqrt(0.998 ** 2 / s)
    pw = (np.linalg.norm(c) + 1) / (np.linalg.norm(b) + 1)
    common = dict(l_inf_ruiz_iterations=0, has_pock_chambolle_alpha=False, bound_objective_rescaling=False,
                  termination_evaluation_frequency=K, eps_optimal_relative=1e-15, eps_feasible_relative=1e-15)
    args = dict(A=sp.csr_matrix(A), c=c, con_lb=b, con_ub=b)
    M, d = _affine(A, b, c, eta / pw, eta * pw)
    z0 = np.zeros(n + m)
    z_km1, zT, zK = _block(M, d, z0, K)
    w = K / (K + 1)
    zH = w * zT + (1 - w) * z0
    want = {"current": _metric(A, pw, eta, zT - z_km1), "reflected_halpern": _metric(A, pw, eta, zT - zK),
            "halpern": _metric(A, pw, eta, zT - zH)}
    oracle_lp.reset_rng()
    r1 = oracle_lp.solve(**args, params=oracle_lp.default_params(iteration_limit=K, **common))
    capfd.readouterr()
    oracle_lp.reset_rng()
    r2 = oracle_lp.solve(**args, params=oracle_lp.default_params(iteration_limit=2 * K, **common))
    LIBC.fflush(None)
    log = capfd.readouterr().err
    sel = re.findall(r"\[restart-select\] iter=(\d+) target=(\w+) score_current=(\S+) score_reflected_halpern=(\S+) "
                     r"score_halpern=(\S+) candidate=(\S+)", log)
    assert len(sel) == 1 and sel[0][0] == str(K)
    target = sel[0][1]
    got = {"current": float(sel[0][2]), "reflected_halpern": float(sel[0][3]), "halpern": float(sel[0][4])}
    assert got["current"] == pytest.approx(want["current"], rel=1e-9)
    assert got["reflected_halpern"] == pytest.approx(want["reflected_halpern"], rel=1e-9)
    if mode == "3":
        assert got["halpern"] == pytest.approx(want["halpern"], rel=1e-9)
        assert target == min(want, key=want.get)  # margin 1: the smallest score wins
    else:
        assert np.isinf(got["halpern"]) and target == "reflected_halpern"  # mode 2 always takes the iterate
    assert float(sel[0][5]) == pytest.approx(want[target], rel=1e-9)
    chosen = {"current": zT, "reflected_halpern": zK, "halpern": zH}[target]
    # the second block: Halpern from the chosen point, primal weight from the PID on (chosen - 0)
    ratio = r1.relative_dual_residual / max(r1.relative_primal_residual, 1e-300)
    assert 1e-8 < ratio < 1e8
    err = np.log(np.linalg.norm(chosen[n:])) - np.log(np.linalg.norm(chosen[:n])) - np.log(pw)
    pw2 = pw * np.exp(0.99 * err + 0.01 * err)
    M2, d2 = _affine(A, b, c, eta / pw2, eta * pw2)
    _, zT2, _ = _block(M2, d2, chosen, K)
    np.testing.assert_allclose(np.concatenate([r2.x, r2.y]), zT2, atol=1e-9 * np.linalg.norm(zT2))


def _fields(log, tag):
    return [dict(re.findall(r"(\w+)=([^\s()]+)", line)) for line in log.splitlines() if line.startswith(f"[{tag}]")]


def _rates(line):
    m = re.search(r"rates\(fp=(\S+) p=(\S+) d=(\S+) g=(\S+) max=(\S+)\)", line)
    return [float(v) for v in m.groups()]


@pytest.mark.parametrize("force_reject", [False, True])
def test_trial_gate_follows_from_its_inputs(oracle_lp, monkeypatch, capfd, force_reject):
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    monkeypatch.setenv("CP_AL_ALPHA_ROLLBACK_TRIAL", "1")
    monkeypatch.setenv("CP_AL_ALPHA_ROLLBACK_TRACE", "1")
    monkeypatch.setenv("PDAL_ALPHA_TRIAL_MIN_SEC", "0")
    if force_reject:
        monkeypatch.setenv("CP_AL_ALPHA_ROLLBACK_FORCE_REJECT", "1")
    inst = planted_lp(300, 600, 8, seed=5)
    p = oracle_lp.default_params(termination_evaluation_frequency=200, iteration_limit=40000,
                                 eps_optimal_relative=1e-10, eps_feasible_relative=1e-10)
    oracle_lp.reset_rng()
    capfd.readouterr()
    oracle_lp.solve(**inst.lp_args(), params=p)
    LIBC.fflush(None)
    log = capfd.readouterr().err
    lines = log.splitlines()
    start_i = next(i for i, l in enumerate(lines) if l.startswith("[alpha-trial-start]"))
    epochs = [_rates(l) for l in lines[:start_i] if l.startswith("[alpha-zero-epoch]")]
    assert len(epochs) >= 2
    h0, h1 = epochs[-2], epochs[-1]  # the history the trial is judged against: the last two zero-alpha epochs
    st = _fields(log, "alpha-trial-start")[0]
    dec = _fields(log, "alpha-trial-decision")[0]
    k = int(dec["K"])
    assert k == int(dec["trajectory_iter"]) - int(st["trajectory_iter"])
    mean = [(a + b) / 2 for a, b in zip(h0, h1)]

    def pred(rate):
        return np.exp(np.clip(-rate * k, -700, 700))
    t_p, t_d, t_g = float(st["primal"]), float(st["dual"]), float(st["gap"])
    want = {"fp_ratio_pred": pred(mean[0]), "p_pred": t_p * pred(mean[1]), "d_pred": t_d * pred(mean[2]),
            "g_pred": t_g * pred(mean[3]), "kkt_pred": max(t_p, t_d, t_g) * pred(mean[4])}
    for key, v in want.items():
        assert float(dec[key]) == pytest.approx(v, rel=1e-9), key
    obs = {key: float(dec[key]) for key in ("fp_ratio_obs", "kkt_obs", "p_obs", "d_obs", "g_obs")}
    assert obs["kkt_obs"] == max(obs["p_obs"], obs["d_obs"], obs["g_obs"])
    gate = (obs["fp_ratio_obs"] <= 0.90 * want["fp_ratio_pred"] and obs["kkt_obs"] <= 0.95 * want["kkt_pred"]
            and obs["p_obs"] <= 1.05 * want["p_pred"] and obs["d_obs"] <= 1.05 * want["d_pred"]
            and obs["g_obs"] <= 1.05 * want["g_pred"])
    if force_reject:
        assert dec["accept"] == "0" and dec["reason"] == "forced-reject"
    else:
        assert dec["accept"] == str(int(gate))
        assert dec["reason"] == ("accepted" if gate else "acceptance-gate")


def test_dense_trace_in_closed_form(oracle_lp, monkeypatch, capfd):
    """CP_AL_LOG_FREQ with verbose (solver.cu:321-358): inside a block the trace
    prints the fixed-point metric of what the buffers hold at that iteration:
    x_R of the iteration itself, x_T, y_T and y_R of the block's first (major)
    iteration.  The checkpoint line at the evaluation is the block's fixed-point
    error and the progress line carries the evaluation's residuals."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    monkeypatch.setenv("CP_AL_LOG_FREQ", "5")
    monkeypatch.setenv("CP_AL_PROGRESS_LOG_FREQ", "25")
    m, n, s, K = 10, 16, 1.7, 25
    A = _orthogonal_rows(m, n, s, seed=3)
    rng = np.random.default_rng(4)
    b = rng.standard_normal(m)
    c = A.T @ rng.standard_normal(m)
    eta = np.sqrt(0.998 ** 2 / s)
    pw = (np.linalg.norm(c) + 1) / (np.linalg.norm(b) + 1)
    M, d = _affine(A, b, c, eta / pw, eta * pw)
    z0 = np.zeros(n + m)
    T0 = M @ z0 + d                         # the major iteration 1: x_T, y_T
    yR1 = 2 * T0[n:] - z0[n:]
    oracle_lp.reset_rng()
    capfd.readouterr()
    r = oracle_lp.solve(A=sp.csr_matrix(A), c=c, con_lb=b, con_ub=b, params=oracle_lp.default_params(
        verbose=True, l_inf_ruiz_iterations=0, has_pock_chambolle_alpha=False, bound_objective_rescaling=False,
        termination_evaluation_frequency=K, iteration_limit=K, eps_optimal_relative=1e-15,
        eps_feasible_relative=1e-15))
    LIBC.fflush(None)
    log = capfd.readouterr().out
    trace = [(int(a), float(b_)) for a, b_ in re.findall(r"^trace iter (\d+) fp (\S+)$", log, re.M)]
    assert [t for t, _ in trace] == [5, 10, 15, 20]  # every 5th iteration but not the evaluation's
    for t, fp in trace:
        z_prev, Tt, _ = _block(M, d, z0, t)  # z_{t-1} and T(z_{t-1})
        xR = 2 * Tt[:n] - z_prev[:n]
        want = _metric(A, pw, eta, np.concatenate([xR - T0[:n], yR1 - T0[n:]]))
        assert fp == pytest.approx(want, rel=1e-9), t
    z_km1, zT, _ = _block(M, d, z0, K)
    cp = re.findall(r"^checkpoint iter (\d+) fp (\S+)$", log, re.M)
    assert len(cp) == 1 and cp[0][0] == str(K)
    assert float(cp[0][1]) == pytest.approx(_metric(A, pw, eta, zT - z_km1), rel=1e-9)
    pg = re.findall(r"^progress iter (\d+) time \S+ rpr (\S+) rdr (\S+) rgap (\S+) fp \S+ pw (\S+)", log, re.M)
    assert len(pg) == 1 and pg[0][0] == str(K)
    assert float(pg[0][1]) == pytest.approx(r.relative_primal_residual, rel=1e-11)
    assert float(pg[0][2]) == pytest.approx(r.relative_dual_residual, rel=1e-11)
    assert float(pg[0][3]) == pytest.approx(r.relative_objective_gap, rel=1e-11)
    assert float(pg[0][4]) == pytest.approx(pw, rel=1e-11)
