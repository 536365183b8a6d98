"""Pins the CPU oracle (oracle/cp_al_oracle.cpp) without the reference running.

The reference's tests hold no numeric golden vectors for this path, so each
oracle function is pinned by something independent of its own code:
  * known answers from the reference documentation (tests/golden/kat_lp.json);
  * planted optima of the generators (KKT by construction);
  * scipy's HiGHS on small LPs with general bounds;
  * numpy's SVD for the power iteration;
  * closed forms of Ruiz, Pock-Chambolle, bound-objective and Curtis-Reid scaling;
  * the Halpern polynomial  z_K - z* = q_K(R)(z_0 - z*),  q_K(t) = sum_{j<=K} t^j/(K+1),
    for the affine reflected map R of an equality-constrained free-variable LP,
    which pins step sizes, reflection, Halpern weights, the restart anchor and
    the PID primal-weight update.
"""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.optimize import linprog

from pdal_amd import lp as L
from pdal_amd.generators import planted_lp

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_kat_lp_from_docs(oracle_lp):
    kat = json.loads((GOLDEN / "kat_lp.json").read_text())
    for case in kat["cases"]:
        A = np.array(case["A"], dtype=float)
        conv = lambda v: np.array([float(t) for t in v])  # noqa: E731  ("inf" strings)
        oracle_lp.reset_rng()
        p = oracle_lp.default_params(eps_optimal_relative=1e-10, eps_feasible_relative=1e-10)
        r = oracle_lp.solve(A=sp.csr_matrix(A), c=conv(case["c"]), con_lb=conv(case["l"]), con_ub=conv(case["u"]),
                            params=p)
        assert r.status == "OPTIMAL", case["source"]
        np.testing.assert_allclose(r.x, case["x"], atol=1e-7)
        np.testing.assert_allclose(r.y, case["y"], atol=1e-7)
        assert abs(r.primal_objective - case["objective"]) < 1e-7


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_planted_optimum_recovered(oracle_lp, monkeypatch, seed):
    monkeypatch.setenv("CP_AL_LAMBDA", "0" if seed % 2 == 0 else "0.05")
    inst = planted_lp(150, 300, 5, seed=seed)
    oracle_lp.reset_rng()
    p = oracle_lp.default_params(eps_optimal_relative=1e-9, eps_feasible_relative=1e-9)
    r = oracle_lp.solve(**inst.lp_args(), params=p)
    assert r.status == "OPTIMAL"
    assert abs(r.primal_objective - inst.optimal_value) <= 1e-7 * abs(inst.optimal_value)
    # KKT of the returned point, measured independently
    assert np.linalg.norm(inst.A @ r.x - inst.b) <= 1e-6 * (1 + np.linalg.norm(inst.b))
    assert r.x.min() >= -1e-12


def _general_lp(seed):
    rng = np.random.default_rng(seed)
    m, n = 30, 45
    A = sp.random(m, n, density=0.2, random_state=seed, format="csr")
    x0 = rng.uniform(-1, 1, n)
    ax = A @ x0
    l = ax - rng.uniform(0.1, 1, m)
    u = ax + rng.uniform(0.1, 1, m)
    l[::3] = -np.inf
    u[1::4] = np.inf
    lb = x0 - rng.uniform(0.5, 2, n)
    ub = x0 + rng.uniform(0.5, 2, n)
    ub[::5] = np.inf
    c = rng.standard_normal(n)
    return A, c, l, u, lb, ub


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_matches_highs(oracle_lp, seed):
    A, c, l, u, lb, ub = _general_lp(seed)
    rows_ub = np.isfinite(u)
    rows_lb = np.isfinite(l)
    A_ub = sp.vstack([A[rows_ub], -A[rows_lb]]).toarray()
    b_ub = np.concatenate([u[rows_ub], -l[rows_lb]])
    ref = linprog(c, A_ub=A_ub, b_ub=b_ub, bounds=list(zip(lb, np.where(np.isinf(ub), None, ub))), method="highs")
    assert ref.status == 0
    oracle_lp.reset_rng()
    p = oracle_lp.default_params(eps_optimal_relative=1e-10, eps_feasible_relative=1e-10)
    r = oracle_lp.solve(A=A, c=c, con_lb=l, con_ub=u, var_lb=lb, var_ub=ub, params=p)
    assert r.status == "OPTIMAL"
    assert abs(r.primal_objective - ref.fun) <= 1e-7 * (1 + abs(ref.fun))


def _hook(oracle_lp, name, restype, argtypes):
    f = getattr(oracle_lp.lib, name)
    f.restype, f.argtypes = restype, argtypes
    return f


def test_power_iteration_matches_svd(oracle_lp):
    ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
    f = _hook(oracle_lp, "cp_al_oracle_sigma_max", C.c_double, [C.c_int, C.c_int, ip, ip, dp, C.c_int, C.c_double])
    for seed in range(3):
        A = sp.random(40, 70, density=0.15, random_state=seed, format="csr")
        A.data = np.random.default_rng(seed).standard_normal(A.nnz)
        rp, ci, v = (np.ascontiguousarray(A.indptr, np.int32), np.ascontiguousarray(A.indices, np.int32),
                     np.ascontiguousarray(A.data))
        sig = f(40, 70, rp.ctypes.data_as(ip), ci.ctypes.data_as(ip), v.ctypes.data_as(dp), 100000, 1e-10)
        assert abs(sig - np.linalg.svd(A.toarray(), compute_uv=False)[0]) <= 1e-8
        # the reference tolerance (1e-4 on ||A A^T v - s^2 v||) still bounds the error
        sig4 = f(40, 70, rp.ctypes.data_as(ip), ci.ctypes.data_as(ip), v.ctypes.data_as(dp), 5000, 1e-4)
        assert abs(sig4 ** 2 - sig ** 2) <= 1e-4 * 10


def _rescale(oracle_lp, A, c, l, u, lb, ub, **params):
    f = _hook(oracle_lp, "cp_al_oracle_rescale", None,
              [C.POINTER(L.LpProblem), C.POINTER(L.CpAlParameters)] + [C.POINTER(C.c_double)] * 4)
    prob = oracle_lp.make_problem(A, c, l, u, lb, ub)
    p = oracle_lp.default_params(**params)
    m, n = A.shape
    vals, E, D, sc = np.zeros(A.nnz), np.zeros(m), np.zeros(n), np.zeros(2)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    f(prob, C.byref(p), dp(vals), dp(E), dp(D), dp(sc))
    oracle_lp.free_problem(prob)
    S = sp.csr_matrix((vals, A.indices, A.indptr), shape=A.shape)
    return S, E, D, sc


def _scaling_case():
    A, c, l, u, lb, ub = _general_lp(9)
    A = sp.csr_matrix(A)
    A.data = A.data * np.random.default_rng(0).lognormal(0, 2, A.nnz)
    return A, c, l, u, lb, ub


def test_ruiz_fixed_point(oracle_lp):
    A, c, l, u, lb, ub = _scaling_case()
    S, E, D, _ = _rescale(oracle_lp, A, c, l, u, lb, ub, l_inf_ruiz_iterations=60, has_pock_chambolle_alpha=False,
                          bound_objective_rescaling=False)
    Sd = np.abs(S.toarray())
    nz_r, nz_c = Sd.max(1) > 0, Sd.max(0) > 0
    np.testing.assert_allclose(Sd.max(1)[nz_r], 1.0, atol=1e-6)
    np.testing.assert_allclose(Sd.max(0)[nz_c], 1.0, atol=1e-6)
    # S = diag(1/E) A diag(1/D)
    np.testing.assert_allclose(S.toarray(), A.toarray() / E[:, None] / D[None, :], rtol=1e-12, atol=0)


def test_single_ruiz_step_closed_form(oracle_lp):
    A, c, l, u, lb, ub = _scaling_case()
    S, E, D, _ = _rescale(oracle_lp, A, c, l, u, lb, ub, l_inf_ruiz_iterations=1, has_pock_chambolle_alpha=False,
                          bound_objective_rescaling=False)
    Ad = np.abs(A.toarray())
    e = np.where(Ad.max(1) < 1e-12, 1.0, np.sqrt(Ad.max(1)))
    d = np.where(Ad.max(0) < 1e-12, 1.0, np.sqrt(Ad.max(0)))
    np.testing.assert_allclose(E, e, rtol=1e-15)
    np.testing.assert_allclose(D, d, rtol=1e-15)


def test_pock_chambolle_alpha1_closed_form(oracle_lp):
    A, c, l, u, lb, ub = _scaling_case()
    S, E, D, _ = _rescale(oracle_lp, A, c, l, u, lb, ub, l_inf_ruiz_iterations=0, has_pock_chambolle_alpha=True,
                          pock_chambolle_alpha=1.0, bound_objective_rescaling=False)
    Ad = np.abs(A.toarray())
    np.testing.assert_allclose(E, np.where(Ad.sum(1) < 1e-12, 1, np.sqrt(Ad.sum(1))), rtol=1e-13)
    np.testing.assert_allclose(D, np.where(Ad.sum(0) < 1e-12, 1, np.sqrt(Ad.sum(0))), rtol=1e-13)


def test_bound_objective_closed_form(oracle_lp):
    A, c, l, u, lb, ub = _scaling_case()
    l = l.copy()
    l[2] = u[2]  # an equality row counts once (preconditioner.cu:699-729)
    S, E, D, sc = _rescale(oracle_lp, A, c, l, u, lb, ub, l_inf_ruiz_iterations=0, has_pock_chambolle_alpha=False,
                           bound_objective_rescaling=True)
    both = np.isfinite(l) & (~np.isfinite(u) | (np.abs(l - u) > 1e-12))
    bn = np.sqrt(np.sum(np.where(both, l, 0) ** 2) + np.sum(np.where(np.isfinite(u), u, 0) ** 2))
    assert sc[0] == pytest.approx(1 / (bn + 1), rel=1e-14)
    assert sc[1] == pytest.approx(1 / (np.linalg.norm(c) + 1), rel=1e-14)
    np.testing.assert_array_equal(E, 1.0)


def test_curtis_reid_geometric_means(oracle_lp, monkeypatch):
    monkeypatch.setenv("CP_AL_CR_SCALING", "1")
    monkeypatch.setenv("CP_AL_CR_ITERS", "200")
    A, c, l, u, lb, ub = _scaling_case()
    S, E, D, _ = _rescale(oracle_lp, A, c, l, u, lb, ub, l_inf_ruiz_iterations=0, has_pock_chambolle_alpha=False,
                          bound_objective_rescaling=False)
    logs = sp.csr_matrix((np.log(np.abs(S.data)), S.indices, S.indptr), shape=S.shape)
    cnt_r = np.diff(S.indptr)
    cnt_c = np.bincount(S.indices, minlength=S.shape[1])
    # last column update makes every column's log-mean exactly 0; rows converge to 0
    np.testing.assert_allclose(np.asarray(logs.sum(0)).ravel()[cnt_c > 0] / cnt_c[cnt_c > 0], 0.0, atol=1e-12)
    np.testing.assert_allclose(np.asarray(logs.sum(1)).ravel()[cnt_r > 0] / cnt_r[cnt_r > 0], 0.0, atol=1e-6)


# ---------------------------------------------------------------------------
# Halpern polynomial on an affine reflected map


def _orthogonal_rows(m, n, s, seed):
    Q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((n, n)))
    return np.sqrt(s) * Q[:m, :]  # A A^T = s I, so the power iteration is exact at step 1


def _affine_T(A, b, c, tau, sig):
    m, n = A.shape
    M = np.block([[np.eye(n), tau * A.T], [-sig * A, np.eye(m) - 2 * sig * tau * A @ A.T]])
    d = np.concatenate([-tau * c, 2 * sig * tau * A @ c + sig * b])
    return M, d


def _halpern_T_output(M, d, z0, K, rho):
    """T(z_{K-1}) with z_{K-1} = z* + q_{K-1}(R)(z0 - z*), R = rho(2M - I) + (1-rho) I."""
    N = M.shape[0]
    Rl = rho * (2 * M - np.eye(N)) + (1 - rho) * np.eye(N)
    zs, *_ = np.linalg.lstsq(np.eye(N) - M, d, rcond=None)
    assert np.linalg.norm((np.eye(N) - M) @ zs - d) < 1e-10
    v = z0 - zs
    acc = np.zeros(N)
    for _ in range(K):  # sum_{j=0}^{K-1} R^j v, uniform weights 1/K
        acc += v
        v = Rl @ v
    zk = zs + acc / K
    return M @ zk + d


@pytest.mark.parametrize("rho", [1.0, 1.5])
def test_one_block_equals_halpern_polynomial(oracle_lp, monkeypatch, rho):
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    monkeypatch.setenv("CP_AL_REFLECTION_COEFFICIENT", str(rho))
    m, n, s, K = 12, 20, 2.5, 37
    A = _orthogonal_rows(m, n, s, seed=1)
    rng = np.random.default_rng(2)
    b = rng.standard_normal(m)
    c = A.T @ rng.standard_normal(m)  # c in range(A^T): the free-variable LP is bounded
    eta = np.sqrt(0.998 ** 2 / s)
    pw = (np.linalg.norm(c) + 1) / (np.linalg.norm(b) + 1)  # bound-objective scaling off (solver.cu:3555-3562)
    oracle_lp.reset_rng()
    p = oracle_lp.default_params(l_inf_ruiz_iterations=0, has_pock_chambolle_alpha=False,
                                 bound_objective_rescaling=False, termination_evaluation_frequency=K,
                                 iteration_limit=K, eps_optimal_relative=1e-15, eps_feasible_relative=1e-15)
    r = oracle_lp.solve(A=sp.csr_matrix(A), c=c, con_lb=b, con_ub=b, params=p)
    M, d = _affine_T(A, b, c, eta / pw, eta * pw)
    zT = _halpern_T_output(M, d, np.zeros(n + m), K, rho)
    np.testing.assert_allclose(r.x, zT[:n], rtol=0, atol=1e-10 * np.linalg.norm(zT))
    np.testing.assert_allclose(r.y, zT[n:], rtol=0, atol=1e-10 * np.linalg.norm(zT))


def test_two_blocks_restart_anchor_and_pid(oracle_lp, monkeypatch):
    """Block 1 always restarts (total == K, utils.cu:336-339) to the T-output
    with the PID primal weight; block 2 is again a Halpern polynomial from the
    new anchor with the new step sizes."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    m, n, s, K = 10, 16, 1.7, 25
    A = _orthogonal_rows(m, n, s, seed=3)
    rng = np.random.default_rng(4)
    b = rng.standard_normal(m)
    c = A.T @ rng.standard_normal(m)
    eta = np.sqrt(0.998 ** 2 / s)
    pw = (np.linalg.norm(c) + 1) / (np.linalg.norm(b) + 1)
    common = dict(l_inf_ruiz_iterations=0, has_pock_chambolle_alpha=False, bound_objective_rescaling=False,
                  termination_evaluation_frequency=K, eps_optimal_relative=1e-15, eps_feasible_relative=1e-15)
    oracle_lp.reset_rng()
    r1 = oracle_lp.solve(A=sp.csr_matrix(A), c=c, con_lb=b, con_ub=b,
                         params=oracle_lp.default_params(iteration_limit=K, **common))
    M, d = _affine_T(A, b, c, eta / pw, eta * pw)
    z1 = _halpern_T_output(M, d, np.zeros(n + m), K, 1.0)
    np.testing.assert_allclose(np.concatenate([r1.x, r1.y]), z1, atol=1e-10 * np.linalg.norm(z1))
    # PID update (solver.cu:2965-2981) with Delta = T-output - anchor (anchor 0)
    dx, dy = np.linalg.norm(z1[:n]), np.linalg.norm(z1[n:])
    ratio = r1.relative_dual_residual / max(r1.relative_primal_residual, 1e-300)
    assert 1e-8 < ratio < 1e8
    err = np.log(dy) - np.log(dx) - np.log(pw)
    pw2 = pw * np.exp(0.99 * err + 0.01 * err)
    oracle_lp.reset_rng()
    r2 = oracle_lp.solve(A=sp.csr_matrix(A), c=c, con_lb=b, con_ub=b,
                         params=oracle_lp.default_params(iteration_limit=2 * K, **common))
    M2, d2 = _affine_T(A, b, c, eta / pw2, eta * pw2)
    z2 = _halpern_T_output(M2, d2, z1, K, 1.0)
    np.testing.assert_allclose(np.concatenate([r2.x, r2.y]), z2, atol=1e-9 * np.linalg.norm(z2))


@pytest.mark.parametrize("cr_iters", ["5", "20"])
def test_curtis_reid_solve_reaches_planted_optimum(oracle_lp, monkeypatch, cr_iters):
    """Curtis-Reid scaling inside the full solve: scale -> solve -> unscale must
    return the planted optimum of the unscaled problem (x*, y*, objective),
    with coefficients spread over four decades so the scaling does work."""
    from pdal_amd.generators import planted_lp
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    monkeypatch.setenv("CP_AL_CR_SCALING", "1")
    monkeypatch.setenv("CP_AL_CR_ITERS", cr_iters)
    inst = planted_lp(150, 300, 6, seed=17)
    rng = np.random.default_rng(5)
    r_scale = 10.0 ** rng.uniform(-2, 2, inst.m)  # badly scaled rows: A' = R A, b' = R b, same x*, y' = y / R
    A = sp.diags(r_scale) @ inst.A
    b = r_scale * inst.b
    oracle_lp.reset_rng()
    p = oracle_lp.default_params(eps_optimal_relative=1e-9, eps_feasible_relative=1e-9, iteration_limit=400000)
    r = oracle_lp.solve(A=sp.csr_matrix(A), c=inst.c, con_lb=b, con_ub=b, var_lb=np.zeros(inst.n),
                        var_ub=np.full(inst.n, np.inf), params=p)
    assert r.status == "OPTIMAL"
    assert abs(r.primal_objective - inst.optimal_value) <= 1e-7 * abs(inst.optimal_value)
    np.testing.assert_allclose(A @ r.x, b, atol=1e-6 * (1 + np.abs(b).max()))
    assert r.x.min() >= 0.0
    lam = inst.c - A.T @ r.y
    assert lam.min() >= -1e-6 * (1 + np.abs(inst.c).max())
