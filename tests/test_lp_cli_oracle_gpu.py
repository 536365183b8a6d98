"""The LP runner on the GPU against the same runner on the oracle.

`oracle/_build/cp_al_oracle_cli` is the runner's own source (policy layer,
MPS reader, output files; LP/CP_AL/src/cli.c:91-1092) linked against the CPU
oracle instead of libcp_al.so, so the two runs take the same routes and set
the same controller variables, and differ only in the solver.  Every route
the reference's protocols use is compared: the plain runner, the feature
policy's variants (cli.c:216-387: safe0, the sigma trigger with the dynamic
evaluation, the trigger probe, sigma mode with and without restart selection,
the trigger with selection) and the unified policy's two-phase structural tail
with the deterministic alpha pulse (cli.c:408-469, 607-832).  The tail's phases
are cut by the iteration limit, far inside their time budgets, so both runs
take the same phases.
"""
from __future__ import annotations

import os
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from pdal_amd.generators import planted_lp
from test_lp_cli import write_mps

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
GPU_CLI = ROOT / "pdal_amd" / "lib" / "cp_al"
ORACLE_CLI = ROOT / "oracle" / "_build" / "cp_al_oracle_cli"


def run_cli(exe, mps, outdir, args, env_extra):
    env = {k: v for k, v in os.environ.items() if not k.startswith("CP_AL_")}
    env.update(env_extra)
    outdir.mkdir(parents=True, exist_ok=True)
    r = subprocess.run([str(exe), *args, str(mps), str(outdir)], capture_output=True, text=True, env=env,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    summary = (outdir / (mps.stem + "_summary.txt")).read_text()
    fields = dict(re.findall(r"^([^:\n]+): (\S+)$", summary, re.M))
    x = np.loadtxt(outdir / (mps.stem + "_primal_solution.txt"))
    y = np.loadtxt(outdir / (mps.stem + "_dual_solution.txt"))
    return fields, x, y, r.stdout + r.stderr


def normalized_planted(m, n, per_col, seed):
    """A planted LP whose rows all have max |a_ij| = 1 (row_logmax_std = 0), equality rows and a dense
    objective: the unified policy's high-equality structural-tail route."""
    inst = planted_lp(m, n, per_col, seed=seed)
    A = sp.csr_matrix(inst.A)
    rmax = np.maximum(abs(A).max(axis=1).toarray().ravel(), 1e-300)
    D = sp.diags(1.0 / rmax)
    return inst, sp.csr_matrix(D @ A), inst.b / rmax


TRIGGER_NOW = {"CP_AL_FEATURE_SIGMA_TRIGGER_PRIMAL": "0", "CP_AL_FEATURE_SIGMA_TRIGGER_DUAL": "1",
               "CP_AL_FEATURE_SIGMA_TRIGGER_RATIO": "0"}
CASES = {
    "plain": ({}, ["--eps_opt", "1e-8", "--eps_feas", "1e-8"]),
    "safe0": ({"CP_AL_FEATURE_AUTO_POLICY": "1"}, ["--eps_opt", "1e-8", "--eps_feas", "1e-8"]),
    "trigger": (dict(TRIGGER_NOW, CP_AL_FEATURE_AUTO_POLICY="1", CP_AL_FEATURE_POLICY_VARIANT="trigger"),
                ["--eps_opt", "1e-8", "--eps_feas", "1e-8"]),
    # the policy's residual gates stay closed on this balanced LP; the iteration gate opens the switch
    "trigger_dynamic_eval": ({"CP_AL_FEATURE_AUTO_POLICY": "1", "CP_AL_FEATURE_POLICY_VARIANT": "trigger",
                              "CP_AL_DYNAMIC_EVAL_AFTER_ITERS": "1000"}, ["--eps_opt", "1e-8", "--eps_feas", "1e-8"]),
    "probe": (dict(TRIGGER_NOW, CP_AL_FEATURE_AUTO_POLICY="1", CP_AL_FEATURE_POLICY_VARIANT="probe"),
              ["--eps_opt", "1e-8", "--eps_feas", "1e-8"]),
    "always_select": ({"CP_AL_FEATURE_AUTO_POLICY": "1", "CP_AL_FEATURE_POLICY_VARIANT": "always_select"},
                      ["--eps_opt", "1e-8", "--eps_feas", "1e-8"]),
    "always_noselect": ({"CP_AL_FEATURE_AUTO_POLICY": "1", "CP_AL_FEATURE_POLICY_VARIANT": "always_noselect"},
                        ["--eps_opt", "1e-8", "--eps_feas", "1e-8"]),
    "trigger_select": (dict(TRIGGER_NOW, CP_AL_FEATURE_AUTO_POLICY="1",
                            CP_AL_FEATURE_POLICY_VARIANT="trigger_select"),
                       ["--eps_opt", "1e-8", "--eps_feas", "1e-8"]),
    # unified route with the tail: phase 1 stops at the iteration limit, phase 2 (warm, CR0,
    # evaluation every 200) fires the pulse past 10000 charged iterations and stops at the limit
    "unified_tail": ({"CP_AL_UNIFIED_POLICY": "1"},
                     ["--eps_opt", "1e-12", "--eps_feas", "1e-12", "--iter_limit", "12000", "--time_limit", "4000"]),
}


@pytest.mark.parametrize("name", list(CASES))
def test_runner_matches_oracle_runner(tmp_path, name):
    env, args = CASES[name]
    inst, A, b = normalized_planted(200, 400, 6, seed=11)
    mps = tmp_path / "p.mps"
    write_mps(mps, A, inst.c, b, b, np.zeros(inst.n), np.full(inst.n, np.inf))
    env = dict(env, PDAL_QUIET="1")
    print(f"[case] runner {name}", flush=True)
    fo, xo, yo, eo = run_cli(ORACLE_CLI, mps, tmp_path / "oracle", ["-v", "--no_presolve", *args], env)
    fg, xg, yg, eg = run_cli(GPU_CLI, mps, tmp_path / "gpu", ["-v", "--no_presolve", *args], env)
    assert fo["Termination Reason"] == fg["Termination Reason"]
    assert fo["Iterations Count"] == fg["Iterations Count"]
    for k in ("Primal Objective Value", "Dual Objective Value"):
        a, c = float(fo[k]), float(fg[k])
        assert abs(a - c) <= 2e-6 * max(abs(a), 1e-300), k  # printed with %e (seven digits)
    ex = np.linalg.norm(xg - xo) / max(np.linalg.norm(xo), 1e-300)
    ey = np.linalg.norm(yg - yo) / max(np.linalg.norm(yo), 1e-300)
    print(f"{name}: {fg['Termination Reason']} {fg['Iterations Count']} iterations, x {ex:.2e}, y {ey:.2e}")
    assert ex <= 1e-7 and ey <= 1e-7
    # the controllers acted the same way in both
    for tag in (r"sigma-trigger iter (\d+)", r"\[dynamic-eval\] iter=(\d+)", r"\[structural-tail-\w+\] status=(\w+)",
                r"\[alpha-direct-pulse-start\] charged_iter=(\d+)", r"\[alpha-direct-pulse-complete\] charged_iter=(\d+)",
                r"\[restart-select\] iter=(\d+) target=(\w+)"):
        assert re.findall(tag, eo) == re.findall(tag, eg), tag
    if name == "unified_tail":
        assert "route=high_equality_tail_or_base" in eg
        assert re.findall(r"\[structural-tail-final\] phases=(\d)", eg) == ["2"]
        assert re.findall(r"\[alpha-direct-pulse-start\]", eg), "the tail's pulse did not start"
    if fg["Termination Reason"] == "OPTIMAL":
        assert abs(float(fg["Primal Objective Value"]) - inst.optimal_value) <= 1e-6 * abs(inst.optimal_value)
