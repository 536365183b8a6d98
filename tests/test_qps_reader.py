"""QPS reader of libfa_cp.so (include/fa_cp_mps_parser.h) and the fa_cp runner
command line (QP/FA_CP/src/mps_parser.c:371-907, QP/FA_CP/src/cli.c:120-455).

CPU tests parse through the C ABI and check every field against values
written out by hand from the file; the GPU test runs the runner on QPTEST
(the example of the QPS format description: two variables, two rows) and
checks it against the closed-form optimum and the oracle.
"""
from __future__ import annotations

import ctypes as C
import gzip
import subprocess
import textwrap
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from pdal_amd import qp as Q

ROOT = Path(__file__).resolve().parent.parent
RUNNER = ROOT / "pdal_amd" / "lib" / "fa_cp"

QPTEST = textwrap.dedent("""\
    NAME          QPTEST
    ROWS
     N  obj
     G  r1
     L  r2
    COLUMNS
        c1        r1                 2.0   r2                -1.0
        c1        obj                1.5
        c2        r1                 1.0   r2                 2.0
        c2        obj               -2.0
    RHS
        rhs1      r1                 2.0   r2                 6.0
    BOUNDS
     UP bnd1      c1                20.0
    QUADOBJ
        c1        c1                 8.0
        c1        c2                 2.0
        c2        c2                10.0
    ENDATA
    """)
# min 1.5 x1 - 2 x2 + 1/2 x'[[8,2],[2,10]]x  s.t. 2 x1 + x2 >= 2, -x1 + 2 x2 <= 6, 0 <= x1 <= 20, x2 >= 0.
# r1 is active: x2 = 2 - 2 x1, d/dx1 = 0 gives x1 = 61/80, x2 = 19/40, objective 1399/320,
# multiplier of r1: (c + Q x)_2 = 4.275.
QPTEST_X = np.array([61 / 80, 19 / 40])
QPTEST_OBJ = 1399 / 320
QPTEST_Y = np.array([4.275, 0.0])


def _reader(lib):
    f = lib.lib.read_mps_file
    f.restype = C.POINTER(Q.QpProblem)
    f.argtypes = [C.c_char_p]
    return f


def _csr(comp, rows, cols, nnz):
    if rows == 0:
        return np.zeros((0, cols))
    rp = np.ctypeslib.as_array(comp.row_ptr, shape=(rows + 1,)).copy()
    ci = np.ctypeslib.as_array(comp.col_ind, shape=(nnz,)).copy() if nnz else np.zeros(0, np.int32)
    v = np.ctypeslib.as_array(comp.val, shape=(nnz,)).copy() if nnz else np.zeros(0)
    assert rp[-1] == nnz
    return sp.csr_matrix((v, ci, rp), shape=(rows, cols))


def _read(lib, path: Path) -> dict:
    prob = _reader(lib)(str(path).encode())
    assert prob, "read_mps_file returned NULL"
    p = prob.contents
    n, m = p.num_variables, p.num_constraints
    vec = lambda ptr, k: np.ctypeslib.as_array(ptr, shape=(k,)).copy() if k else np.zeros(0)  # noqa: E731
    out = dict(n=n, m=m, c=vec(p.objective_vector, n), lb=vec(p.variable_lower_bound, n),
               ub=vec(p.variable_upper_bound, n), l=vec(p.constraint_lower_bound, m),
               u=vec(p.constraint_upper_bound, m), c0=p.objective_constant,
               A=_csr(p.constraint_matrix.contents, m, n, p.constraint_matrix_num_nonzeros),
               Q=_csr(p.objective_sparse_matrix.contents, n, n, p.objective_sparse_matrix_num_nonzeros),
               rank=p.num_rank_lowrank_obj, r_nnz=p.objective_lowrank_matrix_num_nonzeros,
               r_rp0=p.objective_lowrank_matrix.contents.row_ptr[0],
               starts=(bool(p.primal_start), bool(p.dual_start)))
    lib.free_problem(prob)
    return out


@pytest.fixture(scope="module")
def qp_lib():
    return Q.load_fa_cp()


@pytest.mark.parametrize("gz", [False, True])
def test_read_qptest(qp_lib, tmp_path, gz):
    path = tmp_path / ("QPTEST.QPS.gz" if gz else "QPTEST.QPS")
    if gz:
        with gzip.open(path, "wt") as fh:
            fh.write(QPTEST)
    else:
        path.write_text(QPTEST)
    d = _read(qp_lib, path)
    assert (d["n"], d["m"]) == (2, 2)
    np.testing.assert_array_equal(d["A"].toarray(), [[2, 1], [-1, 2]])
    np.testing.assert_array_equal(d["c"], [1.5, -2.0])
    np.testing.assert_array_equal(d["l"], [2.0, -np.inf])
    np.testing.assert_array_equal(d["u"], [np.inf, 6.0])
    np.testing.assert_array_equal(d["lb"], [0.0, 0.0])
    np.testing.assert_array_equal(d["ub"], [20.0, np.inf])
    # QUADOBJ gives the lower triangle: the off-diagonal entry is mirrored
    np.testing.assert_array_equal(d["Q"].toarray(), [[8, 2], [2, 10]])
    assert d["Q"].nnz == 4
    assert d["c0"] == 0.0 and d["rank"] == 0 and d["r_nnz"] == 0 and d["r_rp0"] == 0
    assert d["starts"] == (False, False)


def test_qmatrix_zero_and_unknown_entries_objsense(qp_lib, tmp_path):
    """QMATRIX entries are taken as given (no mirror); zero values and unknown
    columns are skipped; the objective RHS sets the constant; a MAX line with a
    second token negates c, Q and the constant, while a lone MAX line is a
    section header and changes nothing (mps_parser.c:401-441)."""
    body = textwrap.dedent("""\
        ROWS
         N  obj
         E  e1
        COLUMNS
            x         obj       1.0       e1        1.0
            y         obj       -3.0      e1        1.0
        RHS
            rhs       e1        1.0       obj       2.5
        QMATRIX
            x         x         4.0
            x         y         1.0
            y         x         1.0
            y         y         0.0
            x         nosuch    7.0
        ENDATA
        """)
    plain = tmp_path / "qm.qps"
    plain.write_text("NAME QM\n" + body)
    d = _read(qp_lib, plain)
    np.testing.assert_array_equal(d["Q"].toarray(), [[4, 1], [1, 0]])
    assert d["Q"].nnz == 3
    np.testing.assert_array_equal(d["c"], [1.0, -3.0])
    assert d["c0"] == -2.5
    lone = tmp_path / "qm_lone_max.qps"
    lone.write_text("NAME QM\nOBJSENSE\n    MAX\n" + body)
    d2 = _read(qp_lib, lone)
    np.testing.assert_array_equal(d2["c"], d["c"])
    maxed = tmp_path / "qm_max.qps"
    maxed.write_text("NAME QM\nOBJSENSE\n    MAX    OBJ\n" + body)
    d3 = _read(qp_lib, maxed)
    np.testing.assert_array_equal(d3["c"], [-1.0, 3.0])
    np.testing.assert_array_equal(d3["Q"].toarray(), [[-4, -1], [-1, 0]])
    assert d3["c0"] == 2.5


def test_lp_reader_ignores_quadratic_sections(gpu_lp, qp_lib, tmp_path):
    """libcp_al.so's reader does not know QUADOBJ: same A, c and bounds."""
    from pdal_amd import lp as L
    path = tmp_path / "QPTEST.QPS"
    path.write_text(QPTEST)
    f = gpu_lp.lib.read_mps_file
    f.restype = C.POINTER(L.LpProblem)
    f.argtypes = [C.c_char_p]
    prob = f(str(path).encode())
    assert prob
    p = prob.contents
    assert (p.num_variables, p.num_constraints, p.constraint_matrix_num_nonzeros) == (2, 2, 4)
    np.testing.assert_array_equal(np.ctypeslib.as_array(p.objective_vector, shape=(2,)), [1.5, -2.0])
    gpu_lp.free_problem(prob)
    assert _read(qp_lib, path)["Q"].nnz == 4


def test_runner_arguments():
    """Usage and argument errors (cli.c:260-425) need no GPU."""
    assert RUNNER.exists(), "fa_cp runner not built"
    r = subprocess.run([str(RUNNER), "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "Usage:" in r.stderr and "--sigma_update" in r.stderr
    r = subprocess.run([str(RUNNER), "--sigma_update", "bogus", "a", "b"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "sigma_update must be" in r.stderr
    r = subprocess.run([str(RUNNER), "--opt_norm", "l1", "a", "b"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "opt_norm must be" in r.stderr
    r = subprocess.run([str(RUNNER), "only_one_arg"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "You must specify an input file and an output directory" in r.stderr


def _summary(path: Path) -> dict:
    out = {}
    for line in path.read_text().splitlines():
        k, v = line.split(": ", 1)
        out[k] = v
    return out


@pytest.mark.gpu
def test_runner_solves_qptest(oracle_qp, tmp_path):
    """The runner on QPTEST at 1e-8: OPTIMAL, the closed-form optimum, the
    summary keys of cli.c:120-166 in order, %.10g solution files, and the
    oracle's objective on the same data."""
    path = tmp_path / "QPTEST.QPS"
    path.write_text(QPTEST)
    out = tmp_path / "out"
    out.mkdir()
    print("[case] fa_cp runner on QPTEST", flush=True)
    r = subprocess.run([str(RUNNER), "--eps_opt", "1e-8", "--eps_feas", "1e-8", "-v", "0", str(path), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    s = _summary(out / "QPTEST_summary.txt")
    keys = list(s)
    expect = ["Termination Reason", "Runtime (sec)", "Iterations Count", "Inner Iterations Count",
              "Primal Objective Value", "Dual Objective Value", "Relative Primal Residual",
              "Relative Dual Residual", "Absolute Objective Gap", "Relative Objective Gap", "Final AL Sigma",
              "Sigma Update Rule", "Sigma Update Count", "Rows", "Columns", "Nonzeros"]
    assert keys == expect or keys == [k for k in expect if k != "Inner Iterations Count"], keys
    assert s["Termination Reason"] == "OPTIMAL"
    assert s["Sigma Update Rule"] == "guarded"
    assert (s["Rows"], s["Columns"], s["Nonzeros"]) == ("2", "2", "4")
    pobj = float(s["Primal Objective Value"])
    assert abs(pobj - QPTEST_OBJ) <= 1e-6 * QPTEST_OBJ
    x = np.loadtxt(out / "QPTEST_primal_solution.txt")
    y = np.loadtxt(out / "QPTEST_dual_solution.txt")
    np.testing.assert_allclose(x, QPTEST_X, atol=1e-5)
    np.testing.assert_allclose(y, QPTEST_Y, atol=1e-4)
    # the oracle on the same data through create_qp_problem
    Qm = sp.csr_matrix([[8.0, 2.0], [2.0, 10.0]])
    A = sp.csr_matrix([[2.0, 1.0], [-1.0, 2.0]])
    oracle_qp.reset_rng()
    o = oracle_qp.solve(np.array([1.5, -2.0]), Q=Qm, A=A, con_lb=np.array([2.0, -np.inf]),
                        con_ub=np.array([np.inf, 6.0]), var_lb=np.zeros(2), var_ub=np.array([20.0, np.inf]),
                        params=oracle_qp.default_params(eps_optimal_relative=1e-8, eps_feasible_relative=1e-8))
    assert o.status == "OPTIMAL"
    assert abs(pobj - o.primal_objective) <= 1e-6 * abs(o.primal_objective)
    assert int(s["Iterations Count"]) == o.iterations
