"""FA_CP's AL sigma update rules at restarts (update_al_sigma_at_restart,
QP/FA_CP/src/fa_cp_core_op.cu:722-906) pinned on the oracle from outside.

With every rescaling switched off the scaled problem is the problem itself,
so what each rule reads at a restart can be observed from solves cut at the
restart points: the relative residuals (the result at the cut), the
displacement of x and y from the anchor (the results at two consecutive
cuts) and the normal motion ||A dx|| / ||dx||, recomputed in numpy.  FA_CP
restarts at the first evaluation (iteration K) and, by the artificial
criterion, at 2K (the inner count K is at least 0.36 of 2K), so the sigma a
solve cut at 2K (3K) returns is the first (second) update's output.  Each
rule's documented response to those inputs is computed here and compared
with the oracle's final_al_sigma.  The clip range is opened wide and the gain
moved off its default so a wrong input (rp and rd swapped, the wrong anchor,
the wrong norm) or a hard-wired constant shows.
"""
from __future__ import annotations

import numpy as np
import pytest

from pdal_amd.generators import planted_qp

RULES = {"fixed": 0, "residual": 1, "kkt": 2, "displacement": 3, "progress": 4, "greedy": 5, "guarded": 6,
         "balanced_guarded": 7}
GAIN, CLIP_LO, CLIP_HI = 0.7, 0.05, 20.0


def _params(lib, rule, limit, sigma0=0.5, **kw):
    p = lib.default_params(l_inf_ruiz_iterations=0, has_pock_chambolle_alpha=False, bound_objective_rescaling=False,
                           sigma_update_rule=RULES[rule], al_sigma=sigma0, use_al_qp=True, sigma_update_gain=GAIN,
                           sigma_clip_low=CLIP_LO, sigma_clip_high=CLIP_HI, **kw)
    p.termination_criteria.iteration_limit = limit
    p.termination_criteria.eps_optimal_relative = 1e-14
    p.termination_criteria.eps_feasible_relative = 1e-14
    return p


@pytest.fixture(scope="module")
def qp():
    return planted_qp(200, 100, 40, seed=3)


def _cuts(lib, qp, rule, sigma0=0.5, **kw):
    """Results of the same solve cut at 0, K, 2K and 3K iterations."""
    K = lib.default_params().termination_evaluation_frequency
    out = []
    for mult in range(4):
        lib.reset_rng()
        out.append(lib.solve(**qp.qp_args("explicit"), params=_params(lib, rule, mult * K, sigma0, **kw)))
    return K, out


def _inputs(qp, prev, cur):
    """What the restart at `cur` reads: residuals, displacement from the
    anchor `prev`, normal motion."""
    dx, dy = cur.x - prev.x, cur.y - prev.y
    pd, dd = float(np.linalg.norm(dx)), float(np.linalg.norm(dy))
    normal = float(np.linalg.norm(qp.A @ dx) / pd) if pd > 1e-16 else np.inf
    return dict(rp=cur.relative_primal_residual, rd=cur.relative_dual_residual, rg=cur.relative_objective_gap,
                pd=pd, dd=dd, normal=normal, score=max(cur.relative_primal_residual, cur.relative_dual_residual,
                                                       cur.relative_objective_gap))


def _mult(cur, ratio):
    return cur * min(max(ratio ** GAIN, CLIP_LO), CLIP_HI)


def _clamp(s, lo=1e-8, hi=10.0):
    return min(max(s, lo), hi)


def test_cuts_land_on_the_restarts(oracle_qp, qp):
    K, r = _cuts(oracle_qp, qp, "residual")
    assert [x.iterations for x in r] == [0, K, 2 * K, 3 * K]
    assert [x.sigma_update_count for x in r] == [0, 0, 1, 2]


def test_fixed_rule_never_moves(oracle_qp, qp):
    _, r = _cuts(oracle_qp, qp, "fixed")
    assert [x.final_al_sigma for x in r] == [0.5] * 4


@pytest.mark.parametrize("rule", ["residual", "kkt", "displacement", "progress", "greedy"])
def test_multiplicative_rules(oracle_qp, qp, rule):
    _, r = _cuts(oracle_qp, qp, rule)
    i1, i2 = _inputs(qp, r[0], r[1]), _inputs(qp, r[1], r[2])
    s = 0.5
    expected = []
    last = None
    direction = 0
    for n, i in enumerate((i1, i2)):
        other = max(i["rd"], i["rg"])
        if rule == "residual":
            s = _mult(s, i["rp"] / i["rd"])
        elif rule == "kkt":
            s = _mult(s, i["rp"] / other)
        elif rule == "displacement":
            s = _mult(s, _clamp(i["dd"] / i["pd"]) / s)
        elif rule == "progress":
            if last is None:
                s = _mult(s, i["rp"] / other)
            else:  # residual reductions since the last update: the others' against the primal's
                s = _mult(s, (max(last["rd"], last["rg"]) / other) / (last["rp"] / i["rp"]))
        elif rule == "greedy":
            if n == 0:
                direction = 1 if i["rp"] >= other else -1
            elif i["score"] > last["score"]:
                direction = -direction
            s = _mult(s, CLIP_HI if direction > 0 else CLIP_LO)
        s = _clamp(s)
        expected.append(s)
        last = i
    got = [r[2].final_al_sigma, r[3].final_al_sigma]
    np.testing.assert_allclose(got, expected, rtol=1e-9)
    assert abs(got[0] - 0.5) > 1e-3  # the rule moved sigma


@pytest.mark.parametrize("rule", ["guarded", "balanced_guarded"])
def test_guarded_rules_when_active(oracle_qp, qp, rule):
    """With sigma above the zero threshold: the kkt-balance step while the
    primal residual is at least twice the others and the normal motion is not
    negligible, the low clip otherwise."""
    _, r = _cuts(oracle_qp, qp, rule, sigma0=1e-3)
    s = 1e-3
    for k, (prev, cur) in enumerate(((r[0], r[1]), (r[1], r[2]))):
        i = _inputs(qp, prev, cur)
        other = max(i["rd"], i["rg"])
        if i["normal"] > 1e-10 and i["rp"] >= 2.0 * other:
            s = _mult(s, i["rp"] / other)
        else:
            s = _mult(s, CLIP_LO)
        assert r[2 + k].final_al_sigma == pytest.approx(s, rel=1e-9), k
        s = r[2 + k].final_al_sigma


def test_guarded_rule_stays_off_without_a_stall(oracle_qp, qp):
    """At sigma 0 the guarded rule activates only when the score stalled
    (above 0.9 of the last update's) with a primal signal; here the score
    falls, so sigma stays 0 and AL stays off."""
    _, r = _cuts(oracle_qp, qp, "guarded", sigma0=0.0)
    i1, i2 = _inputs(qp, r[0], r[1]), _inputs(qp, r[1], r[2])
    assert i2["score"] <= 0.9 * i1["score"]
    assert [x.final_al_sigma for x in r] == [0.0] * 4
    assert [x.sigma_update_count for x in r] == [0, 0, 1, 2]


def test_rollback_returns_to_the_best_sigma(oracle_qp, qp):
    """A rollback tolerance below 1 makes every update a rollback to the best
    score's sigma, which is the sigma in force at the first update."""
    _, r = _cuts(oracle_qp, qp, "residual", sigma_rollback_tol=0.5)
    assert [x.final_al_sigma for x in r] == [0.5] * 4
    assert [x.sigma_update_count for x in r] == [0, 0, 1, 2]


def test_sigma_bounds_clamp_the_update(oracle_qp, qp):
    """sigma_max caps a growing rule (the greedy step is a clip bound raised to the gain)."""
    _, r = _cuts(oracle_qp, qp, "greedy", sigma_max=0.6)
    i1 = _inputs(qp, r[0], r[1])
    if i1["rp"] >= max(i1["rd"], i1["rg"]):
        assert r[2].final_al_sigma == 0.6
    else:
        assert r[2].final_al_sigma == pytest.approx(0.5 * CLIP_LO ** GAIN, rel=1e-12)
