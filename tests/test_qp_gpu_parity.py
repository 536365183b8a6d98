"""HIP FA_CP path vs the CPU FP64 oracle, through the fa_cp.h C ABI.

Bars:
  * iterates (x_T, y_T unscaled, reduced cost) after a fixed iteration count
    within ITER_RTOL relative of the oracle, for every quadratic-term type
    (none, diagonal, sparse, low-rank, low-rank plus sparse), with the AL
    branch off and on, and under every sigma update rule;
  * the projected-BB inner solve (non-diagonal Q, or the AL branch with a
    sparse Q) makes a trajectory sensitive to rounding once the first restart
    has tightened the inner tolerance: the FP64 oracle summing in reverse
    order drifts from the forward oracle by 1e-6..1e-1 relative over the next
    hundreds of iterations (stop tests ||d|| <= tol flip).  Those iterates are
    held to ITER_RTOL up to the first restart (the low-rank forms already
    split at the first BB solve after it, by 1e-8 between the two oracle
    builds) and, later, to the distance between the two oracle builds
    (NOISE_FACTOR times it); resumed from the device's own state after 1000
    iterations, both sides are held to ITER_RTOL again with equal inner step
    counts, over the lengths the two oracle builds stay within 3e-11 (48-200
    iterations);
  * same termination status and iteration count, objective within 1e-6
    relative of the oracle and of the planted optimum when solved;
  * the documented known answers, infeasibility detection, warm start,
    constraint-free problems and every input format.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from pdal_amd.generators import planted_lp, planted_qp

pytestmark = pytest.mark.gpu

ITER_RTOL = 1e-10
OBJ_RTOL = 1e-6
NOISE_FACTOR = 10.0
BB_KINDS = ["sparse", "lowrank", "lowrank_plus_sparse"]
GOLDEN = Path(__file__).resolve().parent / "golden" / "kat_qp.json"


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_both(oracle_qp, gpu_qp, args, start=None, fmt="csr", **params):
    out = []
    for lib in (oracle_qp, gpu_qp):
        lib.reset_rng()
        p = lib.default_params(**params)
        print(f"[case] lib={lib.path.name} params={params} fmt={fmt}", flush=True)
        kw = {} if start is None else dict(primal_start=start[0], dual_start=start[1])
        out.append(lib.solve(**args, params=p, fmt=fmt, **kw))
    return out


def assert_tracks(o, g, rtol=ITER_RTOL):
    assert o.status == g.status
    assert o.iterations == g.iterations
    assert rel_err(g.x, o.x) <= rtol
    if o.y.size:
        assert rel_err(g.y, o.y) <= rtol
    assert rel_err(g.reduced_cost, o.reduced_cost) <= rtol
    for f in ("relative_primal_residual", "relative_dual_residual", "relative_objective_gap"):
        a, b = getattr(g, f), getattr(o, f)
        assert abs(a - b) <= 1e-6 * max(abs(b), 1e-300) + 1e-14, f


def _f(a):
    return np.array([float(v) for v in a], dtype=np.float64)


def _kat_args(case):
    return dict(c=_f(case["c"]), Q=sp.csr_matrix(np.array(case["Q"])),
                R=None if case["R"] is None else sp.csr_matrix(np.array(case["R"])),
                A=sp.csr_matrix(np.array(case["A"])), con_lb=_f(case["l"]), con_ub=_f(case["u"]),
                var_lb=_f(case["lb"]), var_ub=_f(case["ub"]))


@pytest.mark.parametrize("case", json.loads(GOLDEN.read_text())["cases"], ids=lambda c: c["name"])
def test_known_answer_qps(oracle_qp, gpu_qp, case):
    o, g = run_both(oracle_qp, gpu_qp, _kat_args(case), eps_optimal_relative=1e-10, eps_feasible_relative=1e-10)
    assert g.status == "OPTIMAL"
    assert g.iterations == o.iterations
    np.testing.assert_allclose(g.x, _f(case["x"]), rtol=0, atol=1e-7)
    np.testing.assert_allclose(g.y, _f(case["y"]), rtol=0, atol=1e-7)
    assert abs(g.primal_objective - case["objective"]) <= 1e-7 * abs(case["objective"])
    assert abs(g.primal_objective - o.primal_objective) <= 1e-9 * abs(o.primal_objective)


def _qp_instance(kind: str):
    """Small instances covering each quadratic-term type."""
    inst = planted_qp(200, 100, 40, seed=5)
    if kind == "sparse":
        return inst.qp_args("explicit")
    if kind == "lowrank_plus_sparse":
        return inst.qp_args("lowrank")
    if kind == "lowrank":
        args = inst.qp_args("lowrank")
        args.pop("Q")
        return args
    if kind == "diag":
        args = inst.qp_args("explicit")
        rng = np.random.default_rng(3)
        args["Q"] = sp.diags(rng.uniform(0.5, 2.0, inst.n)).tocsr()
        return args
    if kind == "lp":
        lp = planted_lp(120, 240, 6, seed=4)
        return dict(c=lp.c, A=lp.A, con_lb=lp.b, con_ub=lp.b, var_lb=np.zeros(lp.n), var_ub=np.full(lp.n, np.inf))
    raise KeyError(kind)


@pytest.mark.parametrize("iters", [10, 200, 400, 1000])
@pytest.mark.parametrize("kind", ["lp", "diag"])
def test_iterates_track_oracle(oracle_qp, gpu_qp, kind, iters):
    """No inner solve: the whole trajectory is rounding-stable."""
    o, g = run_both(oracle_qp, gpu_qp, _qp_instance(kind), iteration_limit=iters, eps_optimal_relative=1e-14,
                    eps_feasible_relative=1e-14)
    assert o.status == "ITERATION_LIMIT" and o.iterations == iters
    assert_tracks(o, g)


@pytest.mark.parametrize("iters", [10, 100, 200])
@pytest.mark.parametrize("kind", BB_KINDS)
def test_bb_iterates_track_oracle_through_first_restart(oracle_qp, gpu_qp, kind, iters):
    o, g = run_both(oracle_qp, gpu_qp, _qp_instance(kind), iteration_limit=iters, eps_optimal_relative=1e-14,
                    eps_feasible_relative=1e-14)
    assert o.status == "ITERATION_LIMIT" and o.iterations == iters
    assert g.inner_iterations == o.inner_iterations
    assert_tracks(o, g)


@pytest.mark.parametrize("iters", [400, 1000])
@pytest.mark.parametrize("kind", BB_KINDS)
def test_bb_iterates_within_summation_noise(oracle_qp, oracle_qp_rev, gpu_qp, kind, iters):
    """After the first restart: GPU-vs-oracle distance within NOISE_FACTOR times
    the distance between two summation orders of the oracle itself."""
    args = _qp_instance(kind)
    res = []
    for lib in (oracle_qp, oracle_qp_rev, gpu_qp):
        lib.reset_rng()
        print(f"[case] lib={lib.path.name} kind={kind} iters={iters}", flush=True)
        res.append(lib.solve(**args, params=lib.default_params(iteration_limit=iters, eps_optimal_relative=1e-14,
                                                               eps_feasible_relative=1e-14)))
    o, r, g = res
    assert o.iterations == r.iterations == g.iterations == iters
    floor = max(rel_err(r.x, o.x), rel_err(r.y, o.y))
    dist = max(rel_err(g.x, o.x), rel_err(g.y, o.y))
    print(f"[noise] kind={kind} iters={iters} oracle-vs-reversed={floor:.3e} gpu-vs-oracle={dist:.3e}")
    assert dist <= NOISE_FACTOR * floor + ITER_RTOL
    assert abs(g.inner_iterations - o.inner_iterations) <= 0.1 * o.inner_iterations


RESUMED = {"sparse": 48, "lowrank": 48, "lowrank_plus_sparse": 200}


@pytest.mark.parametrize("kind", BB_KINDS)
def test_bb_resumed_from_device_state_tracks_oracle(oracle_qp, gpu_qp, kind):
    """The BB path on late iterates at 1e-10: the device's own x, y after 1000
    iterations (several restarts in) start both the oracle and the device
    (set_start_values), which then run RESUMED[kind] iterations.  Measured from
    that state, the forward and reversed-sum oracles differ by 1.1e-11 (sparse,
    48), 2.5e-11 (low rank, 48) and 1.4e-13 (low rank plus sparse, 200: 446
    inner steps, the inner tolerance tightened at the evaluations of
    iterations 64, 128 and 192), the device from the oracle by 3.1e-12,
    2.2e-11 and 3.2e-14, with equal inner step counts; ten iterations later the
    sparse and low-rank trajectories have drifted apart by 1e-10 in both
    oracle builds alike."""
    args = _qp_instance(kind)
    gpu_qp.reset_rng()
    print(f"[case] device state after 1000 iterations kind={kind}", flush=True)
    base = gpu_qp.solve(**args, params=gpu_qp.default_params(iteration_limit=1000, eps_optimal_relative=1e-14,
                                                             eps_feasible_relative=1e-14))
    assert base.iterations == 1000
    it = RESUMED[kind]
    o, g = run_both(oracle_qp, gpu_qp, args, start=(base.x, base.y), iteration_limit=it,
                    eps_optimal_relative=1e-14, eps_feasible_relative=1e-14)
    assert o.iterations == g.iterations == it
    assert g.inner_iterations == o.inner_iterations
    assert_tracks(o, g)


@pytest.mark.parametrize("iters", [200, 600])
@pytest.mark.parametrize("kind", ["lp", "diag"])
def test_al_branch_tracks_oracle(oracle_qp, gpu_qp, kind, iters):
    """al_sigma > 0 switches the AL branch on from the start: effective dual
    products inside the BB loop and the AL dual update."""
    o, g = run_both(oracle_qp, gpu_qp, _qp_instance(kind), iteration_limit=iters, al_sigma=0.1,
                    eps_optimal_relative=1e-14, eps_feasible_relative=1e-14)
    assert g.inner_iterations == o.inner_iterations
    assert_tracks(o, g)
    assert g.final_al_sigma == pytest.approx(o.final_al_sigma, rel=1e-12)
    assert g.sigma_update_count == o.sigma_update_count


def test_al_branch_sparse_q_tracks_oracle(oracle_qp, gpu_qp):
    o, g = run_both(oracle_qp, gpu_qp, _qp_instance("sparse"), iteration_limit=200, al_sigma=0.1,
                    eps_optimal_relative=1e-14, eps_feasible_relative=1e-14)
    assert g.inner_iterations == o.inner_iterations
    assert_tracks(o, g)


@pytest.mark.parametrize("rule", range(8))
def test_sigma_rules_track_oracle(oracle_qp, gpu_qp, rule):
    """Every rule acts at the first restart (iteration 200)."""
    o, g = run_both(oracle_qp, gpu_qp, _qp_instance("sparse"), iteration_limit=201, al_sigma=0.05,
                    sigma_update_rule=rule, eps_optimal_relative=1e-14, eps_feasible_relative=1e-14)
    assert g.inner_iterations == o.inner_iterations
    assert_tracks(o, g)
    assert g.final_al_sigma == pytest.approx(o.final_al_sigma, rel=1e-9, abs=1e-300)
    assert g.sigma_update_count == o.sigma_update_count == (0 if rule == 0 else 1)


@pytest.mark.parametrize("form", ["explicit", "lowrank"])
def test_solve_matches_oracle_and_planted(oracle_qp, gpu_qp, form):
    inst = planted_qp(200, 100, 40, seed=5)
    o, g = run_both(oracle_qp, gpu_qp, inst.qp_args(form), eps_optimal_relative=1e-6, eps_feasible_relative=1e-6,
                    iteration_limit=50000)
    assert o.status == g.status == "OPTIMAL"
    assert abs(g.primal_objective - o.primal_objective) <= OBJ_RTOL * abs(o.primal_objective)
    assert abs(g.primal_objective - inst.optimal_value) <= OBJ_RTOL * abs(inst.optimal_value)
    assert g.kkt < 1e-6


@pytest.mark.parametrize("norm", ["l2", "linf"])
@pytest.mark.parametrize("kind,iters", [("lp", 1000), ("diag", 1000), ("sparse", 201)])
def test_optimality_norms(oracle_qp, gpu_qp, norm, kind, iters):
    o, g = run_both(oracle_qp, gpu_qp, _qp_instance(kind), optimality_norm=norm, iteration_limit=iters,
                    eps_optimal_relative=1e-14, eps_feasible_relative=1e-14)
    assert_tracks(o, g)


def test_primal_infeasible_detected(oracle_qp, gpu_qp):
    args = dict(c=np.array([1.0, 1.0]), Q=sp.diags([1.0, 1.0]).tocsr(), A=sp.csr_matrix(np.array([[1.0, 1.0]])),
                con_lb=np.array([-1.0]), con_ub=np.array([-1.0]), var_lb=np.zeros(2), var_ub=np.full(2, np.inf))
    o, g = run_both(oracle_qp, gpu_qp, args, iteration_limit=2000)
    assert o.status == g.status == "PRIMAL_INFEASIBLE"
    assert g.iterations == o.iterations


def test_constraint_free_box_qp(gpu_qp):
    """m = 0 (no A): x = clip(-c/q, lb, ub)."""
    rng = np.random.default_rng(0)
    n = 64
    q = rng.uniform(0.5, 3.0, n)
    c = rng.standard_normal(n) * 3
    gpu_qp.reset_rng()
    p = gpu_qp.default_params(eps_optimal_relative=1e-10, eps_feasible_relative=1e-10)
    r = gpu_qp.solve(c, Q=sp.diags(q).tocsr(), var_lb=-np.ones(n), var_ub=np.ones(n), params=p)
    assert r.status == "OPTIMAL"
    np.testing.assert_allclose(r.x, np.clip(-c / q, -1, 1), rtol=0, atol=1e-7)


@pytest.mark.parametrize("kind,iters", [("sparse", 150), ("diag", 1000)])
def test_warm_start_tracks_oracle(oracle_qp, gpu_qp, kind, iters):
    """Start near the planted optimum (set_start_values, scaled into the
    solver's space as solver_state.cu:630-657 does).  The BB path is held to
    150 iterations: from this start its summation-order noise already reaches
    4e-11 by iteration 200 between the two oracle builds."""
    inst = planted_qp(200, 100, 40, seed=5)
    rng = np.random.default_rng(1)
    start = (inst.x_star + 0.01 * rng.standard_normal(inst.n), inst.y_star + 0.01 * rng.standard_normal(inst.m))
    o, g = run_both(oracle_qp, gpu_qp, _qp_instance(kind), start=start, iteration_limit=iters,
                    eps_optimal_relative=1e-14, eps_feasible_relative=1e-14)
    assert g.inner_iterations == o.inner_iterations
    assert_tracks(o, g)


@pytest.mark.parametrize("fmt", ["dense", "csc", "coo"])
def test_input_formats_give_identical_results(gpu_qp, fmt):
    args = _qp_instance("sparse")
    res = []
    for f in ("csr", fmt):
        gpu_qp.reset_rng()
        p = gpu_qp.default_params(iteration_limit=300, eps_optimal_relative=1e-14, eps_feasible_relative=1e-14)
        res.append(gpu_qp.solve(**args, params=p, fmt=f))
    np.testing.assert_array_equal(res[0].x, res[1].x)
    np.testing.assert_array_equal(res[0].y, res[1].y)


def test_zero_iteration_limit(oracle_qp, gpu_qp):
    o, g = run_both(oracle_qp, gpu_qp, _qp_instance("sparse"), iteration_limit=0)
    assert o.status == g.status == "ITERATION_LIMIT"
    assert o.iterations == g.iterations == 0
    assert rel_err(g.reduced_cost, o.reduced_cost) <= ITER_RTOL
