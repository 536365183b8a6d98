"""The `cp_al` LP runner: flags, usage errors and the policy routes of
LP/CP_AL/src/cli.c.  Route checks run with CP_AL_FEATURE_AUDIT_ONLY=1, which
returns after the policy layer (cli.c:1057-1062), so no GPU is touched; the
end-to-end run with output files is a GPU test."""
from __future__ import annotations

import os
import re
import subprocess
import textwrap
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

ROOT = Path(__file__).resolve().parent.parent
CLI = ROOT / "pdal_amd" / "lib" / "cp_al"


def _run(args, env_extra=None, check=False):
    env = {k: v for k, v in os.environ.items() if not k.startswith("CP_AL_")}
    env.update(env_extra or {})
    return subprocess.run([str(CLI), *args], capture_output=True, text=True, env=env, check=check, timeout=300)


def write_mps(path: Path, A, c, l, u, lb=None, ub=None):
    """Free-format MPS writer for test instances (E/L/G rows, RANGES for l<u)."""
    A = sp.csc_matrix(A)
    m, n = A.shape
    rows, rhs, ranges = [], [], []
    for i in range(m):
        lo, hi = l[i], u[i]
        if np.isfinite(lo) and np.isfinite(hi) and lo == hi:
            rows.append(("E", f"R{i}"))
            rhs.append((f"R{i}", lo))
        elif np.isfinite(lo) and np.isfinite(hi):
            rows.append(("G", f"R{i}"))
            rhs.append((f"R{i}", lo))
            ranges.append((f"R{i}", hi - lo))
        elif np.isfinite(lo):
            rows.append(("G", f"R{i}"))
            rhs.append((f"R{i}", lo))
        else:
            rows.append(("L", f"R{i}"))
            rhs.append((f"R{i}", hi))
    out = ["NAME T", "ROWS", " N  OBJ"] + [f" {t}  {nm}" for t, nm in rows] + ["COLUMNS"]
    for j in range(n):
        if c[j] != 0:
            out.append(f"    X{j}  OBJ  {float(c[j])!r}")
        for k in range(A.indptr[j], A.indptr[j + 1]):
            out.append(f"    X{j}  R{A.indices[k]}  {float(A.data[k])!r}")
        if c[j] == 0 and A.indptr[j] == A.indptr[j + 1]:
            out.append(f"    X{j}  OBJ  0.0")
    out.append("RHS")
    out += [f"    RHS  {nm}  {float(v)!r}" for nm, v in rhs]
    if ranges:
        out.append("RANGES")
        out += [f"    RNG  {nm}  {float(v)!r}" for nm, v in ranges]
    if lb is not None:
        out.append("BOUNDS")
        for j in range(n):
            if np.isinf(lb[j]) and np.isinf(ub[j]):
                out.append(f" FR BND  X{j}")
            else:
                if np.isinf(lb[j]):
                    out.append(f" MI BND  X{j}")
                elif lb[j] != 0:
                    out.append(f" LO BND  X{j}  {float(lb[j])!r}")
                if np.isfinite(ub[j]):
                    out.append(f" UP BND  X{j}  {float(ub[j])!r}")
    out.append("ENDATA")
    path.write_text("\n".join(out) + "\n")


def test_usage_and_bad_args(tmp_path):
    r = _run(["-h"])
    assert r.returncode == 0 and "Usage:" in r.stderr and "--eval_freq" in r.stderr
    r = _run([])
    assert r.returncode == 1 and "You must specify an input file" in r.stderr
    r = _run(["--opt_norm", "l3", "a.mps", str(tmp_path)])
    assert r.returncode == 1 and "opt_norm must be" in r.stderr
    r = _run([str(tmp_path / "missing.mps"), str(tmp_path)])
    assert r.returncode == 1 and "Failed to read" in r.stderr


def _route(tmp_path, A, c, l, u, lb=None, ub=None):
    write_mps(tmp_path / "p.mps", A, c, l, u, lb, ub)
    r = _run([str(tmp_path / "p.mps"), str(tmp_path)],
             {"CP_AL_UNIFIED_POLICY": "1", "CP_AL_FEATURE_AUDIT_ONLY": "1"})
    assert r.returncode == 0, r.stderr
    route = re.search(r"\[unified-policy\] version=2 route=(\S+)", r.stderr).group(1)
    cr = int(re.search(r"cr_selected=(-?\d+)", r.stderr).group(1))
    return route, cr, r.stderr


def test_unified_policy_routes(tmp_path):
    rng = np.random.default_rng(0)
    # equality rows, dense objective -> structural-tail route; uniform columns -> CR5
    A = sp.random(30, 40, density=0.2, random_state=1, format="csr")
    route, cr, err = _route(tmp_path, A, rng.standard_normal(40), np.ones(30), np.ones(30))
    assert route == "high_equality_tail_or_base" and "[feature-matrix-stats] computed=1" in err
    # no equality rows -> zero_equality_cr0 with the CR override 0
    route, cr, _ = _route(tmp_path, A, rng.standard_normal(40), np.full(30, -np.inf), np.ones(30))
    assert route == "zero_equality_cr0" and cr == 0
    # sparse objective and highly uneven column counts -> lowobj_highcv_eval100_cr5
    B = sp.lil_matrix((400, 400))
    B[:, 0] = 1.0  # one dense column, the rest with one entry: col_nnz_cv >> 5
    for j in range(1, 400):
        B[j % 400, j] = 1.0 + j / 400
    c = np.zeros(400)
    c[3] = 1.0
    route, cr, _ = _route(tmp_path, B.tocsr(), c, np.ones(400), np.ones(400) * 2)
    assert route == "lowobj_highcv_eval100_cr5" and cr == 5


def test_feature_auto_policy_cr_thresholds(tmp_path):
    """cli.c:270-287: col_nnz_cv <= 0.19012745 -> 5, <= 0.39521255 -> 50,
    coeff_log_range <= 3.222854 -> 5, else 20."""
    write_mps(tmp_path / "u.mps", sp.random(50, 60, density=0.1, random_state=2), np.ones(60), np.ones(50),
              np.ones(50))
    # exactly 3 nonzeros per column -> col_nnz_cv = 0 -> CR5
    A = sp.lil_matrix((30, 20))
    for j in range(20):
        for k in range(3):
            A[(j + 7 * k) % 30, j] = 1.0 + 0.1 * k
    write_mps(tmp_path / "even.mps", A.tocsr(), np.ones(20), np.ones(30), np.ones(30))
    r = _run([str(tmp_path / "even.mps"), str(tmp_path)],
             {"CP_AL_FEATURE_AUTO_POLICY": "1", "CP_AL_FEATURE_AUDIT_ONLY": "1"})
    assert "cr_selected=5" in r.stderr and "col_nnz_cv=0.000000e+00" in r.stderr
    # an explicit override wins
    r = _run([str(tmp_path / "even.mps"), str(tmp_path)],
             {"CP_AL_FEATURE_AUTO_POLICY": "1", "CP_AL_FEATURE_AUDIT_ONLY": "1", "CP_AL_FEATURE_CR_ITERS": "7"})
    assert "cr_override=7 cr_selected=7" in r.stderr


@pytest.mark.gpu
def test_cli_solves_and_writes_outputs(tmp_path):
    from pdal_amd.generators import planted_lp
    inst = planted_lp(80, 160, 5, seed=3)
    write_mps(tmp_path / "planted.mps", inst.A, inst.c, inst.b, inst.b, np.zeros(inst.n), np.full(inst.n, np.inf))
    print("[case] cli planted 80x160", flush=True)
    r = _run(["--eps_opt", "1e-8", "--eps_feas", "1e-8", "--no_presolve", str(tmp_path / "planted.mps"),
              str(tmp_path)], {"PDAL_QUIET": "1"})
    assert r.returncode == 0, r.stderr
    summary = (tmp_path / "planted_summary.txt").read_text()
    assert "Termination Reason: OPTIMAL" in summary
    for key in ("Precondition time (sec):", "Runtime (sec):", "Iterations Count:", "Primal Objective Value:",
                "Dual Objective Value:", "Relative Primal Residual:", "Relative Dual Residual:",
                "Absolute Objective Gap:", "Relative Objective Gap:", "Rows: 80", "Columns: 160", "Nonzeros: 800"):
        assert key in summary
    obj = float(re.search(r"Primal Objective Value: (\S+)", summary).group(1))
    assert abs(obj - inst.optimal_value) <= 1e-6 * abs(inst.optimal_value)
    x = np.loadtxt(tmp_path / "planted_primal_solution.txt")
    y = np.loadtxt(tmp_path / "planted_dual_solution.txt")
    assert x.shape == (160,) and y.shape == (80,)
    np.testing.assert_allclose(x, inst.x_star, atol=1e-4)
