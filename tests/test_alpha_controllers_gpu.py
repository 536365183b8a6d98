"""The positive-alpha controllers (solver.cu:187-1057) and the restart-candidate
selection (solver.cu:2663-2819) of the single-instance solver against the
oracle.

The deterministic pulse is what the CLI's structural tail switches on for its
second phase (cli.c:758-793: evaluation every 200 iterations, the phase-1
result as the warm anchor, CP_AL_ALPHA_DIRECT_PULSE=1 with the gate forced):
after 10000 charged iterations and two recorded zero-alpha epochs, one block
of 200 iterations runs at alpha = 0.002.  The tests drive both phases with
iteration limits instead of the CLI's time fractions (those are wall-clock
dependent) and stop at the end of the pulse block with the reference's own
test hook (CP_AL_ALPHA_DIRECT_PULSE_TEST_TRAJECTORY_LIMIT).  The controller's
log lines (where the gate opened, the epoch lengths behind it) must be the
oracle's, and so must the iterates after the pulse; a run without the pulse
must land measurably elsewhere (the pulse is not a no-op).

The rollback trial needs 2 s of elapsed time before it may start; both
implementations read PDAL_ALPHA_TRIAL_MIN_SEC to lower that floor so the trial
starts at the same deterministic point in both.
"""
from __future__ import annotations

import re

import numpy as np
import pytest

from pdal_amd.generators import planted_lp

pytestmark = pytest.mark.gpu

ITER_RTOL = 1e-7  # iterates after 10^4 iterations: the oracle's summation order is not the device's


def rel_err(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def fields(log: str, tag: str) -> list[dict]:
    """key=value pairs of every `[tag] ...` line."""
    out = []
    for line in log.splitlines():
        if line.startswith(f"[{tag}]"):
            out.append(dict(re.findall(r"(\w+)=([^\s]+)", line)))
    return out


def run(lib, inst, params, capfd, start=None):
    lib.reset_rng()
    kw = {} if start is None else dict(primal_start=start[0], dual_start=start[1])
    capfd.readouterr()
    r = lib.solve(**inst.lp_args(), params=params, **kw)
    return r, capfd.readouterr().err


@pytest.fixture
def tail(monkeypatch):
    """The tail phase's controller settings (cli.c:758-793)."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    monkeypatch.setenv("CP_AL_ALPHA_DIRECT_PULSE", "1")
    monkeypatch.setenv("CP_AL_DETERMINISTIC_PULSE_FORCE", "1")
    return monkeypatch


def test_two_phase_tail_with_pulse_matches_oracle(oracle_lp, gpu_lp, tail, capfd):
    inst = planted_lp(300, 600, 8, seed=5)
    # phase 1: the base solve (alpha 0, default evaluation frequency), cut by an iteration limit
    tail.setenv("CP_AL_ALPHA_DIRECT_PULSE", "0")
    p1 = oracle_lp.default_params(termination_evaluation_frequency=64, iteration_limit=3200,
                                  eps_optimal_relative=1e-12, eps_feasible_relative=1e-12)
    base_o, _ = run(oracle_lp, inst, p1, capfd)
    base_g, _ = run(gpu_lp, inst, p1, capfd)
    assert base_o.status == base_g.status == "ITERATION_LIMIT"
    assert rel_err(base_g.x, base_o.x) <= 1e-9 and rel_err(base_g.y, base_o.y) <= 1e-9
    # phase 2: warm anchor = the phase-1 result, evaluation every 200 iterations, the pulse forced
    tail.setenv("CP_AL_ALPHA_DIRECT_PULSE", "1")
    tail.setenv("CP_AL_ALPHA_DIRECT_PULSE_TEST_TRAJECTORY_LIMIT", "10000")
    p2 = oracle_lp.default_params(termination_evaluation_frequency=200, iteration_limit=40000,
                                  eps_optimal_relative=0.0, eps_feasible_relative=0.0)
    start = (base_o.x, base_o.y)
    ro, lo = run(oracle_lp, inst, p2, capfd, start)
    rg, lg = run(gpu_lp, inst, p2, capfd, start)
    print(f"\noracle:\n{lo}\ndevice:\n{lg}")
    for tag, keys in (("alpha-deterministic-tail-anchor", ("charged_iter", "trajectory_iter", "previous_K",
                                                           "recent_K", "pass")),
                      ("alpha-direct-pulse-start", ("charged_iter", "trajectory_iter", "pulse_iterations")),
                      ("alpha-direct-pulse-complete", ("charged_iter", "positive_iters", "term")),
                      ("alpha-direct-pulse-final", ("started", "completed", "aborted", "gate_passed",
                                                    "gate_anchor_charged_iter", "positive_charged_iters"))):
        fo, fg = fields(lo, tag), fields(lg, tag)
        assert len(fo) == len(fg) == 1, tag
        assert {k: fo[0][k] for k in keys} == {k: fg[0][k] for k in keys}, tag
    a = fields(lo, "alpha-deterministic-tail-anchor")[0]
    b = fields(lg, "alpha-deterministic-tail-anchor")[0]
    for k in ("previous_fp_rate", "recent_fp_rate"):
        assert abs(float(a[k]) - float(b[k])) <= 1e-6 * abs(float(a[k])), k
    assert int(a["charged_iter"]) >= 10000
    # the solve stops at the end of the pulse block (the test hook), at the oracle's iterates
    assert ro.status == rg.status == "ITERATION_LIMIT" and ro.iterations == rg.iterations
    assert ro.iterations == int(fields(lo, "alpha-direct-pulse-complete")[0]["charged_iter"])
    assert rel_err(rg.x, ro.x) <= ITER_RTOL and rel_err(rg.y, ro.y) <= ITER_RTOL
    # without the pulse the same number of iterations lands elsewhere
    tail.setenv("CP_AL_ALPHA_DIRECT_PULSE", "0")
    p3 = oracle_lp.default_params(termination_evaluation_frequency=200, iteration_limit=ro.iterations,
                                  eps_optimal_relative=0.0, eps_feasible_relative=0.0)
    rn, _ = run(oracle_lp, inst, p3, capfd, start)
    assert rn.iterations == ro.iterations
    assert rel_err(rn.y, ro.y) > 100 * max(rel_err(rg.y, ro.y), 1e-12)


def test_pulse_gate_needs_evaluation_every_200(oracle_lp, gpu_lp, tail, capfd):
    """At another evaluation frequency the gate is never evaluated (solver.cu:913)."""
    inst = planted_lp(200, 400, 8, seed=6)
    p = oracle_lp.default_params(termination_evaluation_frequency=64, iteration_limit=12800,
                                 eps_optimal_relative=0.0, eps_feasible_relative=0.0)
    for lib in (oracle_lp, gpu_lp):
        r, log = run(lib, inst, p, capfd)
        final = fields(log, "alpha-direct-pulse-final")
        assert len(final) == 1 and final[0]["gate_evaluated"] == "0" and final[0]["started"] == "0"
        assert not fields(log, "alpha-direct-pulse-start")


@pytest.mark.parametrize("force_reject", [False, True])
def test_rollback_trial_matches_oracle(oracle_lp, gpu_lp, monkeypatch, capfd, force_reject):
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    monkeypatch.setenv("CP_AL_ALPHA_ROLLBACK_TRIAL", "1")
    monkeypatch.setenv("PDAL_ALPHA_TRIAL_MIN_SEC", "0")
    if force_reject:
        monkeypatch.setenv("CP_AL_ALPHA_ROLLBACK_FORCE_REJECT", "1")
        monkeypatch.setenv("CP_AL_ALPHA_ROLLBACK_TEST_TRAJECTORY_LIMIT", "10000")
    inst = planted_lp(300, 600, 8, seed=5)
    p = oracle_lp.default_params(termination_evaluation_frequency=200, iteration_limit=40000,
                                 eps_optimal_relative=1e-10, eps_feasible_relative=1e-10)
    ro, lo = run(oracle_lp, inst, p, capfd)
    rg, lg = run(gpu_lp, inst, p, capfd)
    print(f"\noracle:\n{lo}\ndevice:\n{lg}")
    st_o, st_g = fields(lo, "alpha-trial-start"), fields(lg, "alpha-trial-start")
    assert len(st_o) == len(st_g) == 1
    assert {k: st_o[0][k] for k in ("charged_iter", "trajectory_iter", "max_K")} == \
        {k: st_g[0][k] for k in ("charged_iter", "trajectory_iter", "max_K")}
    do, dg = fields(lo, "alpha-trial-decision"), fields(lg, "alpha-trial-decision")
    assert [(d["reason"], d["accept"], d["charged_iter"], d["K"]) for d in do] == \
        [(d["reason"], d["accept"], d["charged_iter"], d["K"]) for d in dg]
    assert do and do[0]["accept"] == ("0" if force_reject else "1")
    fo, fg = fields(lo, "alpha-trial-final")[0], fields(lg, "alpha-trial-final")[0]
    for k in ("charged_iter", "trajectory_iter", "started", "accepted", "rejected", "positive_charged_iters"):
        assert fo[k] == fg[k], k
    if force_reject:
        # restored to the snapshot: the trajectory count went back, the charged count did not
        assert int(fo["trajectory_iter"]) < int(fo["charged_iter"])
        assert fields(lo, "alpha-trial-restore")[0]["reason"] == "forced-reject"
    assert ro.status == rg.status and ro.iterations == rg.iterations
    assert rel_err(rg.x, ro.x) <= ITER_RTOL and rel_err(rg.y, ro.y) <= ITER_RTOL


@pytest.mark.parametrize("mode", ["1", "2", "3"])
def test_restart_selection_matches_oracle(oracle_lp, gpu_lp, monkeypatch, capfd, mode):
    """CP_AL_RESTART_SELECT (the feature policy sets 3 with margin 0.9,
    cli.c:258-262): the same terminating block and status as the oracle, and a
    trajectory that differs from the default restart to the T output."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    monkeypatch.setenv("CP_AL_RESTART_SELECT", mode)
    monkeypatch.setenv("CP_AL_RESTART_SELECT_MARGIN", "0.9")
    monkeypatch.setenv("CP_AL_RESTART_SELECT_TRACE", "1")
    inst = planted_lp(300, 600, 8, seed=7)
    p = oracle_lp.default_params(termination_evaluation_frequency=64, iteration_limit=100000,
                                 eps_optimal_relative=1e-8, eps_feasible_relative=1e-8)
    ro, lo = run(oracle_lp, inst, p, capfd)
    rg, lg = run(gpu_lp, inst, p, capfd)
    # every restart's candidates, scores and choice (the scores are pinned in closed form on the oracle)
    so, sg = fields(lo, "restart-select"), fields(lg, "restart-select")
    assert so and len(so) == len(sg)
    for a, b in zip(so, sg):
        assert (a["iter"], a["target"]) == (b["iter"], b["target"])
        for k in ("score_current", "score_reflected_halpern", "score_halpern", "candidate"):
            x, y = float(a[k]), float(b[k])
            assert x == y or abs(x - y) <= 1e-6 * max(abs(x), abs(y)), (k, a, b)
    assert ro.status == rg.status == "OPTIMAL" and ro.iterations == rg.iterations
    assert rel_err(rg.x, ro.x) <= 1e-6
    assert abs(rg.primal_objective - inst.optimal_value) <= 1e-6 * (1 + abs(inst.optimal_value))
    monkeypatch.setenv("CP_AL_RESTART_SELECT", "0")
    r0, _ = run(oracle_lp, inst, p, capfd)
    assert r0.iterations != ro.iterations or rel_err(r0.x, ro.x) > 1e-9
