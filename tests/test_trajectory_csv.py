"""The trajectory CSV of FA_CP and of the planted-KKT variants
(initialize_trajectory_csv / append_trajectory_row, QP/FA_CP/src/solver.cu:35-71,
QP/PLANTED_FA_CP/src/solver.cu:35-109, QP/PLANTED_CP_AL/src/solver.cu:35-109).

One row per checkpoint (iteration 0, every evaluation, every print point),
written before a planned restart is performed, plus a last row when the
iteration limit ends the solve.  CPU: the oracle's rows follow that schedule,
the restart column agrees with the restart schedule, and the last row agrees
with the result.  GPU: the device solver writes the same rows as the oracle
(iteration and restart columns equal; residual columns close while the
trajectories agree, that is up to the first restart where an inner BB solve
runs).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from pdal_amd import qp as Q
from pdal_amd.generators import planted_qp

ROOT = Path(__file__).resolve().parent.parent
ORACLES = {"fa_cp": "libfa_cp_oracle.so", "planted_fa_cp": "libplanted_fa_cp_oracle.so",
           "planted_cp_al": "libplanted_cp_al_oracle.so"}
HEAD_FA_CP = ("iteration,time_sec,rel_primal,rel_dual,rel_gap,kkt_residual,fixed_point_residual,restart,al_sigma,"
              "sigma_update_count")
HEAD_PLANTED = ("iteration,time_sec,rel_primal,rel_dual,rel_gap,kkt_residual,fixed_point_residual,restart,sigma,"
                "primal_weight,al_sigma,lambda_est,primal_step,dual_step")


def oracle(name):
    cls = {"fa_cp": Q.FaCp, "planted_fa_cp": Q.PlantedFaCp, "planted_cp_al": Q.PlantedCpAl}[name]
    return cls(ROOT / "oracle" / "_build" / ORACLES[name])


def device(name):
    return {"fa_cp": Q.load_fa_cp, "planted_fa_cp": Q.load_planted_fa_cp, "planted_cp_al": Q.load_planted_cp_al}[name]()


def rows(path):
    lines = Path(path).read_text().splitlines()
    head = lines[0].split(",")
    return lines[0], [dict(zip(head, (float(v) for v in line.split(",")))) for line in lines[1:]]


def case(name):
    """(problem arguments, parameter overrides) of a small instance."""
    if name == "fa_cp":
        return planted_qp(120, 60, 30, seed=5).qp_args("explicit"), {}
    if name == "planted_fa_cp":
        return planted_qp(120, 60, 30, seed=5).qp_args("explicit", True), {"adaptive_al_sigma": True}
    return planted_qp(120, 60, 30, seed=5).qp_args("explicit", True), {}


def run(lib, name, path, limit=3000, K=50, **extra):
    args, kw = case(name)
    p = lib.default_params(trajectory_csv=str(path).encode(), termination_evaluation_frequency=K,
                           iteration_limit=limit, eps_optimal_relative=1e-9, eps_feasible_relative=1e-9,
                           **kw, **extra)
    lib.reset_rng()
    return lib.solve(**args, params=p)


def print_point(t):
    step, thr = 10, 1000
    while t >= thr:
        step, thr = step * 10, thr * 10
    return t % step == 0


@pytest.mark.parametrize("name", list(ORACLES))
def test_oracle_rows_follow_the_checkpoints(name, tmp_path):
    K = 50
    lib = oracle(name)
    r = run(lib, name, tmp_path / "t.csv", limit=1730, K=K)
    head, rs = rows(tmp_path / "t.csv")
    assert head == (HEAD_FA_CP if name == "fa_cp" else HEAD_PLANTED)
    its = [int(x["iteration"]) for x in rs]
    want = [t for t in range(0, r.iterations) if t == 0 or t % K == 0 or print_point(t)]
    # the limit's row after the loop, or the checkpoint that found the solve optimal
    assert r.status in ("ITERATION_LIMIT", "OPTIMAL")
    want.append(r.iterations)
    assert its == want
    assert np.isnan(rs[0]["fixed_point_residual"])
    # restarts are planned only at evaluations (or iteration 0), and FA_CP restarts at the first evaluation
    assert all(x["iteration"] % K == 0 for x in rs if x["restart"] == 1.0)
    assert next(x for x in rs if x["iteration"] == K)["restart"] == 1.0
    last = rs[-1]
    assert last["rel_primal"] == r.relative_primal_residual and last["rel_dual"] == r.relative_dual_residual
    assert last["kkt_residual"] == max(last["rel_primal"], last["rel_dual"], last["rel_gap"])
    if name == "fa_cp":
        assert last["al_sigma"] == r.final_al_sigma and last["sigma_update_count"] == r.sigma_update_count
        assert last["sigma_update_count"] == sum(x["restart"] for x in rs)  # one update per restart


def test_unwritable_trajectory_path_is_an_error(tmp_path):
    lib = oracle("fa_cp")
    with pytest.raises(RuntimeError):
        run(lib, "fa_cp", tmp_path / "missing" / "t.csv", limit=100)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(ORACLES))
def test_device_rows_equal_the_oracle(name, tmp_path):
    o, g = oracle(name), device(name)
    ro = run(o, name, tmp_path / "o.csv")
    rg = run(g, name, tmp_path / "g.csv")
    ho, xo = rows(tmp_path / "o.csv")
    hg, xg = rows(tmp_path / "g.csv")
    assert ho == hg
    assert ro.status == rg.status
    first_restart = next(x["iteration"] for x in xo if x["restart"] == 1.0)
    n = min(len(xo), len(xg))
    for a, b in zip(xo[:n], xg[:n]):
        if a["iteration"] > first_restart:
            break
        assert a["iteration"] == b["iteration"] and a["restart"] == b["restart"]
        for k in ("rel_primal", "rel_dual", "rel_gap", "primal_weight", "al_sigma", "sigma"):
            if k in a and np.isfinite(a[k]):
                assert b[k] == pytest.approx(a[k], rel=1e-8, abs=1e-14), (a["iteration"], k)
    if name == "planted_fa_cp":  # the device's sigma rule, read off its own trajectory
        from test_planted_fa_cp import _rule_output
        P = dict(trig_p=3.0, trig_g=5.0, inc=1.5, dec=0.7, min=0.005, max=0.3)  # the defaults
        rs = [k for k, x in enumerate(xg[:-1]) if x["restart"] == 1.0]
        assert rs
        for k in rs:
            assert xg[k + 1]["al_sigma"] == pytest.approx(_rule_output(xg[k], "trigger", "both", P), rel=1e-15)
