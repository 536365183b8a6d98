"""The environment-driven controllers of the CP_AL loop, on the oracle (CPU).

Reference: LP/CP_AL/src/solver.cu:1496-1528 (sigma trigger), :1228-1260 and
:1533-1584 (dynamic evaluation frequency), :1059-1145 and :3166-3536 (adaptive
lambda: the tier ladder and the race), :1590-1600 (the race's forced restart);
LP/CP_AL/src/preconditioner.cu:162-268 (Curtis-Reid with its edge-change
diagnostic and early stop).  The LP runner's feature policy switches the
trigger and the dynamic evaluation on (cli.c:319-337).

What pins each controller here (independently of the oracle's code):
  * grid spelling: the printed grid is the sorted, de-duplicated, clipped set
    of the tokens with 0 added;
  * tier ladder: with the activity floors at 0 each plateau flag is its ratio
    against the threshold, the fixed-point ratio is the quotient of the
    restarts' printed fixed-point errors, the chosen tier follows the flags,
    the hold counts down, and the new lambda is the grid value of the tier;
  * race: each epoch's score recomputed from the printed ratios and the epoch
    length (deterministic clock), the running best over non-bad epochs, and an
    exploration that walks the tiers upward one per epoch;
  * sigma trigger: it fires at the first evaluation past MIN_ITERS when the
    thresholds hold, with sigma = 6 sigma_base and sigma_base = 1/sqrt(2.01
    lambda_est); on the default thresholds it stays off on a balanced LP;
  * dynamic evaluation: with restarts off (after the unconditional first one)
    the iterates do not depend on where evaluations happen, so a run that
    switches to 50-iteration blocks equals the run that does not, bit for bit;
    the switch lands on the first evaluation the gate allows;
  * Curtis-Reid early stop: a run stopped after five passes equals a run of
    five passes; the diagnostic's edge changes equal a numpy Curtis-Reid.
The HIP path is held to the oracle on the same scenarios in
tests/test_lp_controllers_gpu.py.
"""
from __future__ import annotations

import ctypes
import math
import re

import numpy as np
import pytest

from pdal_amd.generators import planted_lp

LIBC = ctypes.CDLL(None)
TIER_ENV = {"CP_AL_LAMBDA": "0", "CP_AL_ADAPTIVE_LAMBDA": "1", "CP_AL_LAMBDA_FEAS_ACTIVE": "0",
            "CP_AL_LAMBDA_GAP_ACTIVE": "0", "CP_AL_LAMBDA_FP_ACTIVE": "0"}
RACE_ENV = {"CP_AL_LAMBDA": "0", "CP_AL_ADAPTIVE_LAMBDA": "1", "CP_AL_LAMBDA_RACE": "1",
            "CP_AL_LAMBDA_GRID": "0,0.02,0.05", "CP_AL_LAMBDA_RACE_EXPLOIT_EPOCHS": "1",
            "CP_AL_LAMBDA_FEAS_ACTIVE": "0", "CP_AL_LAMBDA_GAP_ACTIVE": "0", "CP_AL_LAMBDA_FP_ACTIVE": "0",
            "PDAL_CONTROLLER_SECONDS_PER_ITERATION": "1e-4"}
NO_RESTARTS = {"CP_AL_RESTART_SUFFICIENT": "0", "CP_AL_RESTART_NECESSARY": "0", "CP_AL_RESTART_ARTIFICIAL": "1e9"}


def run(lib, inst, capfd, env, monkeypatch, **params):
    """Solve with the given environment; return the result and the captured stdout + stderr."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    lib.reset_rng()
    capfd.readouterr()
    p = lib.default_params(**params)
    r = lib.solve(**inst.lp_args(), params=p)
    LIBC.fflush(None)
    out = capfd.readouterr()
    for k in env:
        monkeypatch.delenv(k, raising=False)
    return r, out.out + out.err


def grid_of(log):
    m = re.search(r"lambda_grid=([0-9eE.+\-/]+)", log)
    return [float(v) for v in m.group(1).split("/")]


def adapt_lines(log):
    pat = re.compile(r"lambda-adapt restart (\d+) ratios\(feas (\S+) gap (\S+) fp (\S+)\) plateau\(feas (\d) gap (\d) "
                     r"fp (\d) dom (\d) mature (\d)\) tier (\d+)->(\d+) hold (\d+) lambda (\S+)->(\S+)")
    out = []
    for m in pat.finditer(log):
        g = m.groups()
        out.append(dict(restart=int(g[0]), feas=float(g[1]), gap=float(g[2]), fp=float(g[3]),
                        pf=int(g[4]), pg=int(g[5]), pp=int(g[6]), dom=int(g[7]), mature=int(g[8]),
                        t0=int(g[9]), t1=int(g[10]), hold=int(g[11]), l0=float(g[12]), l1=float(g[13])))
    return out


def race_lines(log):
    pat = re.compile(r"lambda-race restart (\d+) score (\S+) best\(tier=(\d+) score=(\S+)\) cycle\(tier=(\d+) "
                     r"score=(\S+)\) tier (\d+)->(\d+) lambda (\S+)->(\S+) next=(\d+) exploit=(\d+) cool=(\d+) "
                     r"bad=(\d) start=(\d) mature=(\d) ratios\(feas (\S+) gap (\S+) fp (\S+)\)")
    out = []
    for m in pat.finditer(log):
        g = m.groups()
        out.append(dict(restart=int(g[0]), score=float(g[1]), best_tier=int(g[2]), best=float(g[3]),
                        cycle_tier=int(g[4]), t0=int(g[6]), t1=int(g[7]), l0=float(g[8]), l1=float(g[9]),
                        next=int(g[10]), exploit=int(g[11]), cool=int(g[12]), bad=int(g[13]), start=int(g[14]),
                        feas=float(g[16]), gap=float(g[17]), fp=float(g[18])))
    return out


def restart_fps(log):
    return [(int(a), float(b)) for a, b in re.findall(r"^restart iter (\d+) fp (\S+)$", log, re.M)]


def test_lambda_grid_spelling(oracle_lp, capfd, monkeypatch):
    inst = planted_lp(60, 120, 6, seed=1)
    cases = [
        ({"CP_AL_ADAPTIVE_LAMBDA": "1", "CP_AL_LAMBDA_GRID": "0.3, 0.1;0.1 : 2.0 x\t0.25"},
         [0.0, 0.1, 0.25, 0.3, 0.95]),
        ({"CP_AL_ADAPTIVE_LAMBDA": "1"}, [0.0, 0.1]),  # base 0, low = mid = high = 0.1
        ({"CP_AL_LAMBDA_PROFILE": "race4"}, [0.0, 0.02, 0.1, 0.2]),
        ({"CP_AL_LAMBDA_PROFILE": "safe_auto"},
         [0.0, 0.005, 0.01, 0.02, 0.035, 0.05, 0.075, 0.10, 0.15, 0.20, 0.30, 0.40]),
        ({"CP_AL_ADAPTIVE_LAMBDA": "1", "CP_AL_LAMBDA_LOW": "-1", "CP_AL_LAMBDA_MID": "0.5",
          "CP_AL_LAMBDA_HIGH": "0.5"}, [0.0, 0.5]),
    ]
    for env, want in cases:
        _, log = run(oracle_lp, inst, capfd, env, monkeypatch, verbose=True, iteration_limit=64,
                     termination_evaluation_frequency=64)
        assert grid_of(log) == pytest.approx(want, abs=5e-4), env
        assert "adaptive=1" in log
    # sigma mode switches the controller off; so does an unset controller
    _, log = run(oracle_lp, inst, capfd, {"CP_AL_ADAPTIVE_LAMBDA": "1", "CP_AL_SIGMA_MODE": "1"}, monkeypatch,
                 verbose=True, iteration_limit=64, termination_evaluation_frequency=64)
    assert "adaptive=0" in log and "lambda-adapt" not in log


def test_tier_ladder_follows_its_rule(oracle_lp, capfd, monkeypatch):
    inst = planted_lp(300, 600, 8, seed=3)
    env = dict(TIER_ENV, CP_AL_LAMBDA_GRID="0,0.02,0.05,0.1,0.2")
    r, log = run(oracle_lp, inst, capfd, env, monkeypatch, verbose=True, iteration_limit=30000,
                 termination_evaluation_frequency=100, eps_optimal_relative=1e-8, eps_feasible_relative=1e-8)
    grid = grid_of(log)
    assert grid == pytest.approx([0.0, 0.02, 0.05, 0.1, 0.2], abs=5e-4)
    lines = adapt_lines(log)
    fps = restart_fps(log)
    assert len(lines) >= 5 and len(fps) == len(lines) + 1  # the first restart only records its metrics
    assert any(l["t1"] > 0 for l in lines) and any(l["t1"] == 0 for l in lines[1:])
    low, mid, high = 1, len(grid) // 2, len(grid) - 1
    tier, hold = 0, 0
    for k, l in enumerate(lines):
        # the fixed-point ratio is the quotient of the two restarts' printed errors
        assert l["fp"] == pytest.approx(fps[k + 1][1] / fps[k][1], rel=2e-3)
        # with zero activity floors each plateau flag is its ratio against the threshold
        for flag, ratio, thr in (("pf", "feas", 0.98), ("pg", "gap", 0.95), ("pp", "fp", 0.98)):
            if abs(l[ratio] / thr - 1.0) > 2e-3:
                assert l[flag] == int(l[ratio] > thr), (k, flag)
        assert l["mature"] == 1
        want = (high if l["pg"] and l["dom"] else mid if l["pf"] or (l["pg"] and l["pp"])
                else low if l["pg"] or l["pp"] else 0)
        assert l["t0"] == tier
        if want > 0:
            tier, hold = want, 2
        elif hold > 0 and tier > 0:
            hold -= 1
        else:
            tier, hold = 0, 0
        assert (l["t1"], l["hold"]) == (tier, hold), k
        assert l["l1"] == pytest.approx(grid[tier], abs=5e-4)
    assert r.status == "OPTIMAL"


def test_race_scores_best_and_exploration(oracle_lp, capfd, monkeypatch):
    inst = planted_lp(300, 600, 8, seed=3)
    r, log = run(oracle_lp, inst, capfd, RACE_ENV, monkeypatch, verbose=True, iteration_limit=30000,
                 termination_evaluation_frequency=100, eps_optimal_relative=1e-8, eps_feasible_relative=1e-8)
    lines = race_lines(log)
    fps = restart_fps(log)
    assert len(lines) >= 6 and len(fps) == len(lines) + 1
    best, best_tier = -math.inf, 0
    for k, l in enumerate(lines):
        dt = (fps[k + 1][0] - fps[k][0]) * 1e-4
        # the score: weighted log decrease of feasibility, fixed-point error and gap per second
        s = (0.45 * -math.log(l["feas"]) + 0.35 * -math.log(l["fp"]) + 0.20 * -math.log(l["gap"])) / dt
        assert abs(l["score"] - s) <= 2e-3 * abs(s) + 3e-3 / dt, k
        assert l["fp"] == pytest.approx(fps[k + 1][1] / fps[k][1], rel=2e-3)
        if not l["bad"] and l["score"] > best:
            best, best_tier = l["score"], l["t0"]
        assert l["best_tier"] == best_tier and l["best"] == pytest.approx(best, rel=2e-3), k
    # exploration walks the tiers upward, one per epoch, and starts from tier 0
    explored = [l["t1"] for l in lines if l["t1"] > 0 and l["t0"] < l["t1"]]
    assert explored, "the race never explored a positive tier"
    for a, b in zip(lines, lines[1:]):
        if b["t1"] > b["t0"] and b["t0"] > 0:
            assert b["t1"] == b["t0"] + 1
    assert r.status == "OPTIMAL"


def test_race_forced_restart(oracle_lp, capfd, monkeypatch):
    """CP_AL_LAMBDA_FORCE_RESTART_ITERS forces a restart that many iterations after the last one."""
    inst = planted_lp(300, 600, 8, seed=3)
    env = dict(RACE_ENV, CP_AL_LAMBDA_FORCE_RESTART_ITERS="300", **NO_RESTARTS)
    _, log = run(oracle_lp, inst, capfd, env, monkeypatch, verbose=True, iteration_limit=2000,
                 termination_evaluation_frequency=100, eps_optimal_relative=0.0, eps_feasible_relative=0.0)
    forced = [int(v) for v in re.findall(r"lambda-race artificial restart iter (\d+)", log)]
    assert [it for it, _ in restart_fps(log)] == [100] + forced
    assert forced == list(range(400, 2001, 300))


def test_sigma_trigger(oracle_lp, capfd, monkeypatch):
    inst = planted_lp(300, 600, 8, seed=3)
    p = dict(verbose=True, iteration_limit=2000, termination_evaluation_frequency=100, eps_optimal_relative=0.0,
             eps_feasible_relative=0.0)
    env = {"CP_AL_LAMBDA": "0", "CP_AL_SIGMA_TRIGGER": "1", "CP_AL_SIGMA_TRIGGER_MIN_ITERS": "250",
           "CP_AL_SIGMA_TRIGGER_PRIMAL": "0", "CP_AL_SIGMA_TRIGGER_DUAL": "1", "CP_AL_SIGMA_TRIGGER_RATIO": "0"}
    r, log = run(oracle_lp, inst, capfd, env, monkeypatch, **p)
    trig = re.findall(r"sigma-trigger iter (\d+) rpr (\S+) rdr (\S+) ratio (\S+) sigma (\S+) base (\S+)", log)
    assert len(trig) == 1
    it, rp, rd, ratio, sigma, base = trig[0]
    assert int(it) == 300  # the first evaluation at or past MIN_ITERS
    lam = float(re.search(r"lambdaA=(\S+)", log).group(1))
    assert float(sigma) == pytest.approx(6.0 * float(base), rel=1e-12)
    assert float(base) == pytest.approx(1.0 / math.sqrt(2.01 * lam), rel=2e-3)
    assert float(ratio) == pytest.approx(float(rp) / float(rd), rel=1e-11)
    # the trigger restarts at once, whatever the restart test says
    assert 300 in [i for i, _ in restart_fps(log)]
    # on the default thresholds a balanced LP never triggers, and the run is the plain one
    for k in ("CP_AL_SIGMA_TRIGGER_PRIMAL", "CP_AL_SIGMA_TRIGGER_DUAL", "CP_AL_SIGMA_TRIGGER_RATIO"):
        env.pop(k)
    r1, log1 = run(oracle_lp, inst, capfd, env, monkeypatch, **p)
    r0, _ = run(oracle_lp, inst, capfd, {"CP_AL_LAMBDA": "0"}, monkeypatch, **p)
    assert "sigma-trigger" not in log1
    assert np.array_equal(r1.x, r0.x) and np.array_equal(r1.y, r0.y)
    assert not np.allclose(r.x, r0.x, rtol=1e-6)


@pytest.mark.parametrize("gate", ["iteration", "time", "residual"])
def test_dynamic_eval_switch_keeps_iterates(oracle_lp, capfd, monkeypatch, gate):
    inst = planted_lp(300, 600, 8, seed=3)
    p = dict(verbose=True, iteration_limit=1500, termination_evaluation_frequency=200, eps_optimal_relative=0.0,
             eps_feasible_relative=0.0)
    env = dict(NO_RESTARTS, CP_AL_LAMBDA="0")
    gates = {"iteration": {"CP_AL_DYNAMIC_EVAL_AFTER_ITERS": "500"},
             "time": {"CP_AL_DYNAMIC_EVAL_AFTER_SEC": "0.45", "PDAL_CONTROLLER_SECONDS_PER_ITERATION": "1e-3"},
             "residual": {"CP_AL_DYNAMIC_EVAL": "1", "CP_AL_DYNAMIC_EVAL_SMALL_RES": "1",
                          "CP_AL_DYNAMIC_EVAL_LARGE_RES": "0", "CP_AL_DYNAMIC_EVAL_MIN_ITERS": "500"}}
    r0, _ = run(oracle_lp, inst, capfd, env, monkeypatch, **p)
    r1, log = run(oracle_lp, inst, capfd, dict(env, **gates[gate]), monkeypatch, **p)
    sw = re.findall(r"\[dynamic-eval\] iter=(\d+) .*eval_frequency=(\d+)->(\d+) .*reason=(\w+)", log)
    assert sw == [("600", "200", "50", gate)]
    # evaluations every 50 iterations from there: the run stops exactly at the limit
    assert r0.iterations == 1600 and r1.iterations == 1500
    # where evaluations happen does not change the iterates (no restarts after the first)
    p2 = dict(p, iteration_limit=1600)
    r2, _ = run(oracle_lp, inst, capfd, dict(env, **gates[gate]), monkeypatch, **p2)
    assert r2.iterations == 1600
    assert np.array_equal(r2.x, r0.x) and np.array_equal(r2.y, r0.y)


def test_dynamic_eval_bootstrap_and_cr_early_stop_gate(oracle_lp, capfd, monkeypatch):
    inst = planted_lp(300, 600, 8, seed=3)
    p = dict(verbose=True, iteration_limit=1000, termination_evaluation_frequency=200, eps_optimal_relative=0.0,
             eps_feasible_relative=0.0)
    boot = {"CP_AL_LAMBDA": "0", "CP_AL_DYNAMIC_EVAL": "1", "CP_AL_DYNAMIC_EVAL_BOOTSTRAP_ITERS": "400",
            "CP_AL_DYNAMIC_EVAL_SMALL_RES": "0", "CP_AL_DYNAMIC_EVAL_LARGE_RES": "1e300",
            "CP_AL_DYNAMIC_EVAL_RATIO": "1e300"}
    _, log = run(oracle_lp, inst, capfd, boot, monkeypatch, **p)
    assert re.findall(r"^dynamic-eval iter (\d+) eval_freq (\d+)->(\d+)", log, re.M) == [("400", "50", "200")]
    assert restart_fps(log)[0][0] == 50  # the unconditional first restart closes the first (50-iteration) block
    # an early-stopped Curtis-Reid closes the iteration and time gates
    cr = {"CP_AL_LAMBDA": "0", "CP_AL_CR_SCALING": "1", "CP_AL_CR_EARLY_STOP": "1",
          "CP_AL_CR_EARLY_STOP_THRESHOLD": "100", "CP_AL_DYNAMIC_EVAL_AFTER_ITERS": "300"}
    _, log = run(oracle_lp, inst, capfd, cr, monkeypatch, **p)
    assert "[cr-early-stop] selected=20 executed=5" in log and "dynamic-eval" not in log


def numpy_cr_edge_changes(A, passes):
    """Curtis-Reid from its definition: alternately set each row's (column's) log
    scale to the mean over its nonzeros of -log|a_ij| - (the other side's log
    scale); report max |row change + column change| over the nonzeros."""
    A = A.tocsr()
    rows, cols = A.nonzero()
    la = -np.log(np.maximum(np.abs(np.asarray(A[rows, cols]).ravel()), 1e-300))
    m, n = A.shape
    r, c = np.zeros(m), np.zeros(n)
    rc, cc = np.bincount(rows, minlength=m), np.bincount(cols, minlength=n)
    out = []
    for _ in range(passes):
        r_new = np.where(rc > 0, np.bincount(rows, la - c[cols], minlength=m) / np.maximum(rc, 1), 0.0)
        c_new = np.where(cc > 0, np.bincount(cols, la - r_new[rows], minlength=n) / np.maximum(cc, 1), 0.0)
        out.append(float(np.max(np.abs((r_new - r)[rows] + (c_new - c)[cols]))))
        r, c = r_new, c_new
    return out


def test_cr_early_stop_and_diagnostic(oracle_lp, capfd, monkeypatch):
    inst = planted_lp(200, 400, 6, seed=7)
    p = dict(iteration_limit=600, termination_evaluation_frequency=100, eps_optimal_relative=0.0,
             eps_feasible_relative=0.0)
    base = {"CP_AL_LAMBDA": "0", "CP_AL_CR_SCALING": "1"}
    r_stop, log = run(oracle_lp, inst, capfd, dict(base, CP_AL_CR_EARLY_STOP="1", CP_AL_CR_DIAGNOSTIC="1",
                                                   CP_AL_CR_EARLY_STOP_THRESHOLD="100"), monkeypatch, **p)
    diag = [float(v) for v in re.findall(r"\[cr-diagnostic\] iteration=\d+ max_edge_log_change=(\S+)", log)]
    assert len(diag) == 5 and "[cr-early-stop] selected=20 executed=5" in log
    assert diag == pytest.approx(numpy_cr_edge_changes(inst.A, 5), rel=1e-9)
    r5, _ = run(oracle_lp, inst, capfd, dict(base, CP_AL_CR_ITERS="5"), monkeypatch, **p)
    assert np.array_equal(r_stop.x, r5.x) and np.array_equal(r_stop.y, r5.y)
    # a threshold below the fifth change keeps all 20 passes (the default run)
    thr = repr(diag[4] * 0.5)
    r20, log = run(oracle_lp, inst, capfd, dict(base, CP_AL_CR_EARLY_STOP="1", CP_AL_CR_EARLY_STOP_THRESHOLD=thr),
                   monkeypatch, **p)
    rdef, _ = run(oracle_lp, inst, capfd, base, monkeypatch, **p)
    assert "cr-early-stop" not in log
    assert np.array_equal(r20.x, rdef.x) and np.array_equal(r20.y, rdef.y)
    # the early stop is for the 20-pass default only
    _, log = run(oracle_lp, inst, capfd, dict(base, CP_AL_CR_ITERS="10", CP_AL_CR_EARLY_STOP="1",
                                              CP_AL_CR_EARLY_STOP_THRESHOLD="100"), monkeypatch, **p)
    assert "cr-early-stop" not in log
