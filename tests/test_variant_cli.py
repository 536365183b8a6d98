"""Runners of the three QP variants: rhr_lin_cp_al (QP/LIN_CP_AL/src/cli.c),
rhr_cp_al (QP/PLANTED_CP_AL/src/cli.c) and rhr_fa_cp
(QP/PLANTED_FA_CP/src/cli.c), linked against librhr_lin_cp_al.so,
librhr_cp_al.so and librhr_fa_cp.so.

CPU tests: usage and argument errors, the equality-rows check
(validate_eqbox_problem) that rejects a problem before any solve, and the
reference's exit status 0 after such errors (its main discards run_*'s
status).  GPU tests run each runner on EQTEST, QPTEST with its first row
made an equality and the second dropped: min 1.5 x1 - 2 x2 + 1/2 x'[[8,2],[2,10]]x
s.t. 2 x1 + x2 = 2, 0 <= x1 <= 20, x2 >= 0, whose optimum x = (61/80, 19/40),
objective 1399/320, multiplier 4.275 is QPTEST's (there the first row is the
active one and the second is slack), and check the summary keys, the solution
files, the closed form and the oracle.
"""
from __future__ import annotations

import subprocess
import textwrap
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from pdal_amd import qp as Q

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "pdal_amd" / "lib"
RUNNERS = {"lin": LIB / "rhr_lin_cp_al", "planted_cp": LIB / "rhr_cp_al", "planted": LIB / "rhr_fa_cp"}
ORACLES = {"lin": (Q.LinCpAl, "liblin_cp_al_oracle.so"), "planted_cp": (Q.PlantedCpAl, "libplanted_cp_al_oracle.so"),
           "planted": (Q.PlantedFaCp, "libplanted_fa_cp_oracle.so")}

EQTEST = textwrap.dedent("""\
    NAME          EQTEST
    ROWS
     N  obj
     E  r1
    COLUMNS
        c1        r1                 2.0
        c1        obj                1.5
        c2        r1                 1.0
        c2        obj               -2.0
    RHS
        rhs1      r1                 2.0
    BOUNDS
     UP bnd1      c1                20.0
    QUADOBJ
        c1        c1                 8.0
        c1        c2                 2.0
        c2        c2                10.0
    ENDATA
    """)
QPTEST_ROWS = textwrap.dedent("""\
    NAME          INEQ
    ROWS
     N  obj
     G  r1
    COLUMNS
        c1        r1                 2.0
        c1        obj                1.5
        c2        r1                 1.0
    RHS
        rhs1      r1                 2.0
    ENDATA
    """)
EQ_X = np.array([61 / 80, 19 / 40])
EQ_OBJ = 1399 / 320
EQ_Y = np.array([4.275])


def run(which, *args, timeout=120):
    return subprocess.run([str(RUNNERS[which]), *map(str, args)], capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("which", list(RUNNERS))
def test_runner_usage_and_argument_errors(which):
    assert RUNNERS[which].exists(), f"{RUNNERS[which].name} not built"
    r = run(which, "--help")
    assert r.returncode == 0 and "Usage:" in r.stderr and "This EQBOX build only accepts problems with Ax = b" in r.stderr
    assert ("--sigma_init" in r.stderr) == (which in ("lin", "planted_cp"))
    assert ("--al_lambda" in r.stderr) == (which == "lin")
    assert ("--trajectory_csv" in r.stderr) == (which in ("planted_cp", "planted"))
    assert ("--adaptive_al_sigma_rule" in r.stderr) == (which == "planted")
    assert "--sigma_update" not in r.stderr  # FA_CP's sigma rules are not flags of the variants
    r = run(which, "--opt_norm", "l1", "a", "b")
    assert r.returncode == 0 and "opt_norm must be" in r.stderr  # exit status 0, as the reference's main
    r = run(which, "only_one_arg")
    assert r.returncode == 0 and "You must specify an input file and an output directory" in r.stderr
    r = run(which, "--sigma_update", "fixed", "a", "b")
    assert r.returncode == 0 and "unrecognized option" in r.stderr


def test_planted_runner_rejects_unknown_adaptive_spellings():
    r = run("planted", "--adaptive_al_sigma_mode", "sideways", "a", "b")
    assert "adaptive_al_sigma_mode must be one of" in r.stderr
    r = run("planted", "--adaptive_al_sigma_rule", "pid", "a", "b")
    assert "adaptive_al_sigma_rule must be one of" in r.stderr


@pytest.mark.parametrize("which", list(RUNNERS))
def test_runner_rejects_inequality_rows_before_solving(which, tmp_path):
    """validate_eqbox_problem runs between read_mps_file and optimize, so no
    GPU is touched and no output file is written."""
    path = tmp_path / "INEQ.QPS"
    path.write_text(QPTEST_ROWS)
    out = tmp_path / "out"
    out.mkdir()
    r = run(which, path, out)
    assert r.returncode == 0
    assert "Error: EQBOX build requires finite equality rows; row 0 is [2, inf]." in r.stderr
    assert not list(out.iterdir())
    path.write_text(QPTEST_ROWS.replace(" G  r1", " E  r1").replace("ENDATA", "RANGES\n    rng       r1                 1.0\nENDATA"))
    r = run(which, path, out)
    assert "Error: EQBOX build requires Ax=b; row 0 has interval [2, 3]." in r.stderr
    assert not list(out.iterdir())


def _summary(path: Path) -> dict:
    return dict(line.split(": ", 1) for line in path.read_text().splitlines())


@pytest.mark.gpu
@pytest.mark.parametrize("which", list(RUNNERS))
def test_runner_solves_eqtest(which, tmp_path):
    path = tmp_path / "EQTEST.QPS"
    path.write_text(EQTEST)
    out = tmp_path / "out"
    out.mkdir()
    print(f"[case] {RUNNERS[which].name} on EQTEST", flush=True)
    r = run(which, "--eps_opt", "1e-8", "--eps_feas", "1e-8", "-v", "0", path, out)
    assert r.returncode == 0, r.stderr
    s = _summary(out / "EQTEST_summary.txt")
    expect = ["Termination Reason", "Runtime (sec)", "Iterations Count", "Inner Iterations Count",
              "Primal Objective Value", "Dual Objective Value", "Relative Primal Residual",
              "Relative Dual Residual", "Absolute Objective Gap", "Relative Objective Gap", "Rows", "Columns",
              "Nonzeros"]
    assert list(s) == expect or list(s) == [k for k in expect if k != "Inner Iterations Count"], list(s)
    assert s["Termination Reason"] == "OPTIMAL"
    assert (s["Rows"], s["Columns"], s["Nonzeros"]) == ("1", "2", "2")
    pobj = float(s["Primal Objective Value"])
    assert abs(pobj - EQ_OBJ) <= 1e-6 * EQ_OBJ
    np.testing.assert_allclose(np.loadtxt(out / "EQTEST_primal_solution.txt"), EQ_X, atol=1e-5)
    np.testing.assert_allclose(np.atleast_1d(np.loadtxt(out / "EQTEST_dual_solution.txt")), EQ_Y, atol=1e-4)
    cls, name = ORACLES[which]
    oracle = cls(ROOT / "oracle" / "_build" / name)
    oracle.reset_rng()
    o = oracle.solve(np.array([1.5, -2.0]), Q=sp.csr_matrix([[8.0, 2.0], [2.0, 10.0]]),
                     A=sp.csr_matrix([[2.0, 1.0]]), con_lb=np.array([2.0]), con_ub=np.array([2.0]),
                     var_lb=np.zeros(2), var_ub=np.array([20.0, np.inf]),
                     params=oracle.default_params(eps_optimal_relative=1e-8, eps_feasible_relative=1e-8))
    assert o.status == "OPTIMAL"
    assert abs(pobj - o.primal_objective) <= 1e-6 * abs(o.primal_objective)
    if which == "lin":  # no inner solve: the same iteration count (the others run a BB inner solve on EQTEST's Q)
        assert int(s["Iterations Count"]) == o.iterations


@pytest.mark.gpu
def test_lin_runner_polishes(tmp_path):
    """-f runs feasibility polishing after a loose solve and reports it as the
    oracle does.  The linearized solver's phases never restart (a phase's
    reference fixed-point error starts at infinity and only a restart sets it),
    so each phase is bounded by --iter_limit here."""
    path = tmp_path / "EQTEST.QPS"
    path.write_text(EQTEST)
    out = tmp_path / "out"
    out.mkdir()
    flags = dict(eps_optimal_relative=1e-3, eps_feasible_relative=1e-3, feasibility_polishing=True,
                 eps_feas_polish_relative=1e-9, iteration_limit=2000)
    print("[case] rhr_lin_cp_al -f on EQTEST", flush=True)
    r = run("lin", "--eps_opt", "1e-3", "--eps_feas", "1e-3", "-f", "--eps_feas_polish", "1e-9", "--iter_limit", "2000",
            "-v", "0", path, out)
    assert r.returncode == 0, r.stderr
    s = _summary(out / "EQTEST_summary.txt")
    cls, name = ORACLES["lin"]
    oracle = cls(ROOT / "oracle" / "_build" / name)
    oracle.reset_rng()
    o = oracle.solve(np.array([1.5, -2.0]), Q=sp.csr_matrix([[8.0, 2.0], [2.0, 10.0]]),
                     A=sp.csr_matrix([[2.0, 1.0]]), con_lb=np.array([2.0]), con_ub=np.array([2.0]),
                     var_lb=np.zeros(2), var_ub=np.array([20.0, np.inf]), params=oracle.default_params(**flags))
    assert s["Termination Reason"] == o.status == "OPTIMAL"
    assert float(s["Feasibility Polishing Time (sec)"]) > 0.0
    assert int(s["Feasibility Polishing Iteration Count"]) == o.feasibility_iteration > 0
    x = np.loadtxt(out / "EQTEST_primal_solution.txt")
    np.testing.assert_allclose(x, o.x, rtol=1e-6, atol=1e-9)
