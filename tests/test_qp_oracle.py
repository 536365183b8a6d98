"""The FA_CP CPU oracle pinned by what does not depend on it: the reference's
documented known-answer QPs, planted KKT points, closed-form optima,
brute-force active-set enumeration on tiny QPs and HiGHS on an LP.  A dropped
term, a wrong sign or a transposed operand in the oracle's iteration,
rescaling or residuals moves one of these answers."""
from __future__ import annotations

import itertools
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.optimize import linprog

from pdal_amd.generators import planted_qp

GOLDEN = Path(__file__).resolve().parent / "golden" / "kat_qp.json"


def _f(a):
    return np.array([float(v) for v in a], dtype=np.float64)


def kat_cases():
    return json.loads(GOLDEN.read_text())["cases"]


def kat_args(case):
    return dict(c=_f(case["c"]), Q=sp.csr_matrix(np.array(case["Q"])),
                R=None if case["R"] is None else sp.csr_matrix(np.array(case["R"])),
                A=sp.csr_matrix(np.array(case["A"])), con_lb=_f(case["l"]), con_ub=_f(case["u"]),
                var_lb=_f(case["lb"]), var_ub=_f(case["ub"]))


@pytest.mark.parametrize("case", kat_cases(), ids=lambda c: c["name"])
@pytest.mark.parametrize("fmt", ["dense", "csr"])
def test_known_answer_qps(oracle_qp, case, fmt):
    """QP/FA_CP/docs/C_API.md:51-120 and README.md:221-268."""
    oracle_qp.reset_rng()
    p = oracle_qp.default_params(eps_optimal_relative=1e-10, eps_feasible_relative=1e-10)
    r = oracle_qp.solve(**kat_args(case), params=p, fmt=fmt)
    assert r.status == "OPTIMAL"
    np.testing.assert_allclose(r.x, _f(case["x"]), rtol=0, atol=1e-7)
    np.testing.assert_allclose(r.y, _f(case["y"]), rtol=0, atol=1e-7)
    assert abs(r.primal_objective - case["objective"]) <= 1e-7 * abs(case["objective"])


def test_planted_qp_generator_is_a_kkt_point():
    """The planted (x*, y*, lam*) satisfies the KKT system in the solver's sign
    convention, so it is the optimum the solver must reach."""
    inst = planted_qp(300, 150, 60, seed=11)
    Q = inst.q_explicit()
    x, y, lam = inst.x_star, inst.y_star, inst.lam_star
    assert np.allclose(Q @ x + inst.c - inst.A.T @ y - lam, 0, atol=1e-10)
    ax = inst.A @ x
    assert np.all(ax >= inst.l - 1e-12) and np.all(ax <= inst.u + 1e-12)
    assert np.all((y > 0) <= np.isclose(ax, inst.l)) and np.all((y < 0) <= np.isclose(ax, inst.u))
    assert np.all((lam > 0) <= (x == inst.lb)) and np.all((lam < 0) <= (x == inst.ub))
    # convex: Q = M^T M + reg I
    assert np.linalg.eigvalsh(Q.toarray()).min() > 0


@pytest.mark.parametrize("form", ["explicit", "lowrank"])
def test_planted_qp_reaches_planted_optimum(oracle_qp, form):
    inst = planted_qp(200, 100, 40, seed=5)
    oracle_qp.reset_rng()
    p = oracle_qp.default_params(eps_optimal_relative=1e-6, eps_feasible_relative=1e-6, iteration_limit=50000)
    r = oracle_qp.solve(**inst.qp_args(form), params=p)
    assert r.status == "OPTIMAL"
    assert abs(r.primal_objective - inst.optimal_value) <= 1e-6 * abs(inst.optimal_value)
    assert np.linalg.norm(r.x - inst.x_star) <= 1e-3 * np.linalg.norm(inst.x_star)


def test_explicit_and_lowrank_forms_agree(oracle_qp):
    """Q = M^T M + reg I passed explicitly (SPARSE_Q) or as R = M plus a diagonal
    (LOW_RANK_PLUS_SPARSE_Q) is the same problem."""
    inst = planted_qp(200, 100, 40, seed=6)
    res = []
    for form in ("explicit", "lowrank"):
        oracle_qp.reset_rng()
        p = oracle_qp.default_params(eps_optimal_relative=1e-7, eps_feasible_relative=1e-7, iteration_limit=50000)
        res.append(oracle_qp.solve(**inst.qp_args(form), params=p))
    assert all(r.status == "OPTIMAL" for r in res)
    assert abs(res[0].primal_objective - res[1].primal_objective) <= 1e-6 * abs(res[0].primal_objective)


@pytest.mark.parametrize("seed", [0, 1])
def test_separable_box_qp_closed_form(oracle_qp, seed):
    """No constraints (m = 0, n from Q), diagonal Q: x = clip(-c/q, lb, ub)."""
    rng = np.random.default_rng(seed)
    n = 50
    q = rng.uniform(0.5, 3.0, n)
    c = rng.standard_normal(n) * 3
    lb, ub = -np.ones(n), np.ones(n)
    oracle_qp.reset_rng()
    p = oracle_qp.default_params(eps_optimal_relative=1e-10, eps_feasible_relative=1e-10)
    r = oracle_qp.solve(c, Q=sp.diags(q).tocsr(), var_lb=lb, var_ub=ub, params=p)
    assert r.status == "OPTIMAL"
    np.testing.assert_allclose(r.x, np.clip(-c / q, lb, ub), rtol=0, atol=1e-7)


def _active_set_optimum(H, c, A, l, u, lb, ub):
    """Brute force: every choice of active row/bound sides, solve the equality
    QP, keep the feasible candidate of least objective."""
    n, m = H.shape[0], A.shape[0]
    G = np.vstack([A, np.eye(n)])
    lo, hi = np.concatenate([l, lb]), np.concatenate([u, ub])
    best = (np.inf, None)
    for choice in itertools.product((0, 1, 2), repeat=m + n):  # 0 free, 1 at lower, 2 at upper
        rows = [i for i, s in enumerate(choice) if s]
        if any(not np.isfinite(lo[i] if choice[i] == 1 else hi[i]) for i in rows):
            continue
        if len(rows) > n:
            continue
        Ga = G[rows]
        rhs = np.array([lo[i] if choice[i] == 1 else hi[i] for i in rows])
        K = np.block([[H, Ga.T], [Ga, np.zeros((len(rows), len(rows)))]])
        try:
            sol = np.linalg.solve(K, np.concatenate([-c, rhs]))
        except np.linalg.LinAlgError:
            continue
        x = sol[:n]
        gx = G @ x
        if np.all(gx >= lo - 1e-9) and np.all(gx <= hi + 1e-9):
            f = 0.5 * x @ H @ x + c @ x
            if f < best[0] - 1e-12:
                best = (f, x)
    return best


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_tiny_qp_matches_brute_force(oracle_qp, seed):
    rng = np.random.default_rng(100 + seed)
    n, m = 3, 2
    Mh = rng.standard_normal((n, n))
    H = Mh.T @ Mh + 0.1 * np.eye(n)
    c = rng.standard_normal(n) * 2
    A = rng.standard_normal((m, n))
    l = np.array([-1.0, -np.inf])
    u = np.array([1.0, 0.5])
    lb, ub = np.full(n, -2.0), np.full(n, 2.0)
    fstar, xstar = _active_set_optimum(H, c, A, l, u, lb, ub)
    oracle_qp.reset_rng()
    p = oracle_qp.default_params(eps_optimal_relative=1e-10, eps_feasible_relative=1e-10)
    r = oracle_qp.solve(c, Q=sp.csr_matrix(H), A=sp.csr_matrix(A), con_lb=l, con_ub=u, var_lb=lb, var_ub=ub,
                        params=p)
    assert r.status == "OPTIMAL"
    assert abs(r.primal_objective - fstar) <= 1e-7 * max(1.0, abs(fstar))
    np.testing.assert_allclose(r.x, xstar, rtol=0, atol=1e-5)


def test_lp_through_fa_cp_matches_highs(oracle_qp):
    """No quadratic term (NON_Q): FA_CP is a PDHG LP solver; HiGHS pins the optimum."""
    rng = np.random.default_rng(7)
    m, n = 30, 60
    A = sp.random(m, n, density=0.2, random_state=7, format="csr") + sp.eye(m, n, format="csr")
    x0 = rng.uniform(0, 1, n)
    b = A @ x0
    c = rng.uniform(0.1, 1.0, n)
    ref = linprog(c, A_eq=A.toarray(), b_eq=b, bounds=[(0, 2)] * n, method="highs")
    assert ref.status == 0
    oracle_qp.reset_rng()
    p = oracle_qp.default_params(eps_optimal_relative=1e-9, eps_feasible_relative=1e-9, iteration_limit=200000)
    r = oracle_qp.solve(c, A=A, con_lb=b, con_ub=b, var_lb=np.zeros(n), var_ub=np.full(n, 2.0), params=p)
    assert r.status == "OPTIMAL"
    assert abs(r.primal_objective - ref.fun) <= 1e-6 * abs(ref.fun)


def test_primal_infeasible_is_detected(oracle_qp):
    """x >= 0 and x1 + x2 = -1 has no solution: the ray certificate fires
    within the first three evaluation blocks (utils.cu:1028-1132, 438-462)."""
    A = sp.csr_matrix(np.array([[1.0, 1.0]]))
    oracle_qp.reset_rng()
    p = oracle_qp.default_params(iteration_limit=2000)
    r = oracle_qp.solve(np.array([1.0, 1.0]), Q=sp.diags([1.0, 1.0]).tocsr(), A=A, con_lb=np.array([-1.0]),
                        con_ub=np.array([-1.0]), var_lb=np.zeros(2), var_ub=np.full(2, np.inf), params=p)
    assert r.status == "PRIMAL_INFEASIBLE"


# ---------------------------------------------------------------------------
# Q eigenvalue estimates (QP/FA_CP/src/utils.cu:172-315) against numpy.


def _q_eigen(Q, max_iter=50000, tol=1e-13, seed=3):
    import ctypes as C
    from conftest import ORACLE_QP
    lib = C.CDLL(str(ORACLE_QP))
    lib.fa_cp_oracle_q_eigen.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                         C.c_uint, C.c_void_p]
    Q = sp.csr_matrix(Q)
    rp, ci = Q.indptr.astype(np.int32), Q.indices.astype(np.int32)
    v = Q.data.astype(np.float64)
    out = np.zeros(2)
    lib.fa_cp_oracle_q_eigen(Q.shape[0], rp.ctypes.data, ci.ctypes.data, v.ctypes.data, max_iter, tol, seed,
                             out.ctypes.data)
    return out


def _with_spectrum(d, seed):
    U, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((len(d), len(d))))
    return (U * np.asarray(d)) @ U.T


@pytest.mark.parametrize("case", ["psd", "indefinite", "negative_dominant"])
def test_q_eigen_estimates_match_eigvalsh(case):
    """estimate_maximum_eigenvalue is a power iteration on Q (Rayleigh quotient
    of the normalised iterate): it converges to the eigenvalue of largest
    magnitude, with its sign.  estimate_minimum_eigenvalue runs the same on
    q_norm I - Q and returns q_norm minus that dominant eigenvalue: the
    smallest eigenvalue when q_norm is the largest one, and, when a negative
    eigenvalue dominates, q_norm - (q_norm - lambda_max) = lambda_max -- the
    reference's formula taken as it stands."""
    d = {"psd": [10.0, 6.0, 3.0, 2.0, 1.0, 0.7, 0.3, 0.05],
         "indefinite": [8.0, 3.0, 1.0, 0.2, -0.5, -2.0],
         "negative_dominant": [3.0, 1.0, 0.5, -6.0]}[case]
    Q = _with_spectrum(d, seed=len(d))
    Q[np.abs(Q) < 1e-14] = 0.0
    ev = np.linalg.eigvalsh(Q)
    qn, qmin = _q_eigen(Q)
    dom = ev[np.argmax(np.abs(ev))]
    assert qn == pytest.approx(dom, abs=1e-9)
    shifted = qn - ev  # eigenvalues of q_norm I - Q
    assert qmin == pytest.approx(qn - shifted[np.argmax(np.abs(shifted))], abs=1e-9)
    if case != "negative_dominant":
        assert qn == pytest.approx(ev.max(), abs=1e-9) and qmin == pytest.approx(ev.min(), abs=1e-9)
    else:
        assert qmin == pytest.approx(ev.max(), abs=1e-9)
