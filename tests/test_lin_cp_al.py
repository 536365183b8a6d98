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

import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from pdal_amd import qp as Q
from pdal_amd.generators import planted_lp, planted_qp
from pdal_amd.qp import LinCpAl

ROOT = Path(__file__).resolve().parent.parent
INCLUDE = ROOT / "include"

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


# ---------------------------------------------------------------------------
# Boundary of librhr_lin_cp_al.so (no GPU call)

@pytest.fixture(scope="module")
def gpu_lin():
    """The MI355X build; the test fails (no fallback) if it is missing."""
    return Q.load_lin_cp_al()


def _declared(header: Path) -> list:
    text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    return [n for n in re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text)
            if n not in ("if", "while", "return", "sizeof")]


def _exported(lib: Path) -> set:
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib)], check=True, capture_output=True, text=True)
    return {ln.split()[-1] for ln in out.stdout.splitlines() if " T " in ln}


def test_library_exports_every_declared_symbol(gpu_lin):
    declared = _declared(INCLUDE / "lin_cp_al.h")
    assert "solve_qp_problem" in declared and "cp_al_result_free" in declared
    exported = _exported(gpu_lin.path)
    assert not set(declared) - exported
    assert not [n for n in exported if "pdal" in n and not n.startswith("pdal_lin_cp_al_")]
    assert "fa_cp_result_free" not in exported


def test_struct_layout_matches_header(tmp_path):
    """ctypes mirror == C layout of include/lin_cp_al_types.h (the field order
    and types of QP/LIN_CP_AL/include/cp_al_types.h)."""
    structs = {"cp_al_parameters_t": Q.LinCpAlParameters, "cp_al_result_t": Q.LinCpAlResult,
               "qp_problem_t": Q.QpProblem, "matrix_desc_t": Q.MatrixDesc}
    lines = []
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    src = tmp_path / "layout.c"
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"lin_cp_al_types.h\"\nint main(void){\n"
                   + "\n".join(lines) + "\nreturn 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{INCLUDE}", str(src), "-o", str(exe)], check=True)
    got = dict(ln.rsplit(" ", 1) for ln in subprocess.run([str(exe)], check=True, capture_output=True,
                                                          text=True).stdout.splitlines())
    for cname, cls in structs.items():
        assert int(got[f"{cname} size"]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(cls, fname).offset, (cname, fname)


def test_default_parameters_match_oracle(gpu_lin, oracle_lin):
    a, b = gpu_lin.default_params(), oracle_lin.default_params()
    assert bytes(a) == bytes(b)


# ---------------------------------------------------------------------------
# GPU path against the oracle

ITER_RTOL = 1e-10


def rel_err(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


CASES = [("lp", True), ("lp", False), ("explicit", True), ("explicit", False), ("lowrank", True)]


def case_args(kind):
    return lp_case()[1] if kind == "lp" else qp_case(kind)[1]


@pytest.mark.gpu
@pytest.mark.parametrize("iters", [50, 450, 1200])
@pytest.mark.parametrize("kind,al", CASES)
def test_gpu_iterates_track_oracle(oracle_lin, gpu_lin, kind, al, iters):
    """Fixed iteration budgets through the first restarts (the evaluation and
    restart points are every 200 iterations)."""
    args = case_args(kind)
    p = dict(iteration_limit=iters, eps_optimal_relative=1e-14, eps_feasible_relative=1e-14, use_al_qp=al)
    print(f"[case] lin_cp_al {kind} al={al} iters={iters}", flush=True)
    o = solve(oracle_lin, args, **p)
    g = solve(gpu_lin, args, **p)
    assert o.iterations == g.iterations == iters
    assert rel_err(g.x, o.x) <= ITER_RTOL and rel_err(g.y, o.y) <= ITER_RTOL


@pytest.mark.gpu
@pytest.mark.parametrize("kind,al", CASES)
def test_gpu_solves_match_oracle(oracle_lin, gpu_lin, kind, al):
    args = case_args(kind)
    print(f"[case] lin_cp_al solve {kind} al={al}", flush=True)
    o = solve(oracle_lin, args, use_al_qp=al, **TOL)
    g = solve(gpu_lin, args, use_al_qp=al, **TOL)
    assert g.status == o.status == "OPTIMAL"
    assert g.iterations == o.iterations
    assert abs(g.primal_objective - o.primal_objective) <= 1e-9 * (1 + abs(o.primal_objective))
    assert rel_err(g.x, o.x) <= 1e-7
