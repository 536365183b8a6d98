"""Shared-A engine of the batched CP_AL (pdal_batch.h): when every instance
has the same constraint matrix, the batch keeps the instances' vectors side by
side and one pass over A serves all of them.  Every instance must still follow
the single-instance algorithm: iterates within 1e-10 of the oracle solving it
alone, and the same solves as the block-diagonal engine.
"""
from __future__ import annotations

import numpy as np
import pytest

from pdal_amd.generators import config4_thetas, lp_family
from pdal_amd.lp import LpBatch

pytestmark = pytest.mark.gpu

ITER_RTOL = 1e-10
K = 64


def rel_err(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def family_instances(count, seed=3, m=200):
    fam = lp_family(m, 2 * m, 8, seed=seed)
    return [fam.instance(t) for t in config4_thetas(count, seed)]


@pytest.mark.parametrize("lam", ["0", "0.05"])
@pytest.mark.parametrize("count,blocks", [(10, 1), (10, 6), (70, 3)])
def test_shared_engine_tracks_oracle(oracle_lp, gpu_lp, monkeypatch, lam, count, blocks):
    monkeypatch.setenv("CP_AL_LAMBDA", lam)
    insts = family_instances(count)
    params = dict(iteration_limit=blocks * K, termination_evaluation_frequency=K, eps_optimal_relative=1e-14,
                  eps_feasible_relative=1e-14)
    print(f"[case] shared-A batch lam={lam} count={count} blocks={blocks}", flush=True)
    with LpBatch(gpu_lp, insts, gpu_lp.default_params(**params)) as b:
        assert b.stats()["shared_a"]
        assert b.solve() == 0
        for k in range(0, count, max(1, count // 8)):
            oracle_lp.reset_rng()
            o = oracle_lp.solve(**insts[k], params=oracle_lp.default_params(**params))
            g = b.result(k)
            assert g.iterations == o.iterations == blocks * K
            assert rel_err(g.x, o.x) <= ITER_RTOL, k
            assert rel_err(g.y, o.y) <= ITER_RTOL, k
            assert rel_err(g.reduced_cost, o.reduced_cost) <= ITER_RTOL, k


def test_shared_engine_equals_block_diagonal(gpu_lp, monkeypatch):
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    insts = family_instances(40, seed=4)
    params = gpu_lp.default_params(termination_evaluation_frequency=K, eps_optimal_relative=1e-6,
                                   eps_feasible_relative=1e-6, iteration_limit=100000)
    runs = []
    for flag in ("0", "1"):
        monkeypatch.setenv("PDAL_BATCH_NO_SHARED", flag)
        print(f"[case] fixed-A solve, no_shared={flag}", flush=True)
        with LpBatch(gpu_lp, insts, params) as b:
            assert b.stats()["shared_a"] == (flag == "0")
            assert b.solve() == len(insts)
            runs.append([b.result(k) for k in range(len(insts))])
    for a, c in zip(*runs):
        assert a.status == c.status == "OPTIMAL"
        assert a.iterations == c.iterations
        assert rel_err(a.x, c.x) <= 1e-8 and rel_err(a.y, c.y) <= 1e-8
        assert abs(a.primal_objective - c.primal_objective) <= 1e-9 * (1 + abs(c.primal_objective))


def test_shared_engine_repeated_solves_and_starts(gpu_lp, monkeypatch):
    """Bit-identical repeated solves; warm starts and primal weights reach the
    shared layout (a restart from the own solutions converges fast)."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    insts = family_instances(12, seed=5)
    params = gpu_lp.default_params(termination_evaluation_frequency=K, eps_optimal_relative=1e-6,
                                   eps_feasible_relative=1e-6, iteration_limit=100000)
    with LpBatch(gpu_lp, insts, params) as b:
        assert b.stats()["shared_a"]
        b.solve()
        first = [b.result(k) for k in range(len(insts))]
        pw = b.primal_weights()
        b.solve()
        for k, r in enumerate(first):
            s = b.result(k)
            assert np.array_equal(s.x, r.x) and s.iterations == r.iterations
        b.set_primal_weights(pw)
        b.set_starts(np.concatenate([r.x for r in first]), np.concatenate([r.y for r in first]))
        assert b.solve() == len(insts)
        warm = [b.result(k) for k in range(len(insts))]
    assert sum(w.iterations for w in warm) < 0.3 * sum(r.iterations for r in first)


@pytest.mark.parametrize("count", [150, 300])
def test_shared_packing_is_bitwise_neutral(gpu_lp, monkeypatch, count):
    """Packing the running instances into narrower layouts as others stop
    (256 -> 128 -> 64 slots, or 512 -> ...) leaves every instance's arithmetic
    as it was: the same bits as the solve that never packs."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    from pdal_amd.generators import planted_family
    fam = planted_family(150, 300, 8, seed=6)
    insts = [fam.instance(t) for t in config4_thetas(count, 8)]
    params = gpu_lp.default_params(termination_evaluation_frequency=K, eps_optimal_relative=1e-6,
                                   eps_feasible_relative=1e-6, iteration_limit=100000)
    monkeypatch.setenv("PDAL_BATCH_NO_HANDOFF", "1")  # the block-diagonal engine sums in another order
    # lane counts pinned: the set-up's power iteration of two batches must not autotune apart
    monkeypatch.setenv("PDAL_BATCH_G_A", "16")
    monkeypatch.setenv("PDAL_BATCH_G_AT", "8")
    runs = []
    for flag in ("1", "0"):
        monkeypatch.setenv("PDAL_BATCH_NO_COMPACT", flag)
        print(f"[case] shared-A packing count={count} no_compact={flag}", flush=True)
        with LpBatch(gpu_lp, insts, params) as b:
            assert b.stats()["shared_a"]
            assert b.solve() == count
            st = b.stats()
            assert (st["compactions"] > 0) == (flag == "0")
            runs.append([b.result(k) for k in range(count)])
    for a, c in zip(*runs):
        assert a.status == c.status == "OPTIMAL" and a.iterations == c.iterations
        assert np.array_equal(a.x, c.x) and np.array_equal(a.y, c.y)
        assert np.array_equal(a.reduced_cost, c.reduced_cost)
    its = [r.iterations for r in runs[0]]
    assert max(its) > min(its)  # the instances stop at different blocks


@pytest.mark.parametrize("mixed", [False, True])
def test_equality_rows_are_bitwise_neutral(gpu_lp, monkeypatch, mixed):
    """Rows with l = u in every instance: the dual pass takes the projection
    onto [l, u] as l itself, with the same bits as the general projection.
    With one instance's row made a range (mixed) the batch must not, and both
    runs take the general projection."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    from pdal_amd.generators import planted_family
    fam = planted_family(150, 300, 8, seed=6)
    insts = [fam.instance(t) for t in config4_thetas(100, 10)]
    assert all(np.array_equal(i["con_lb"], i["con_ub"]) for i in insts)
    if mixed:
        u = insts[5]["con_ub"].copy()
        u[7] += 1.0
        insts[5] = dict(insts[5], con_ub=u)
    params = gpu_lp.default_params(termination_evaluation_frequency=K, eps_optimal_relative=1e-6,
                                   eps_feasible_relative=1e-6, iteration_limit=20000)
    monkeypatch.setenv("PDAL_BATCH_G_A", "16")
    monkeypatch.setenv("PDAL_BATCH_G_AT", "8")
    runs = []
    for flag in ("0", "1"):
        monkeypatch.setenv("PDAL_BATCH_NO_EQ_ROWS", flag)
        print(f"[case] equality rows mixed={mixed} no_eq_rows={flag}", flush=True)
        with LpBatch(gpu_lp, insts, params) as b:
            st = b.stats()
            assert st["shared_a"] and st["equality_rows"] == (flag == "0" and not mixed)
            b.solve()
            runs.append([b.result(k) for k in range(len(insts))])
    for a, c in zip(*runs):
        assert a.status == c.status and a.iterations == c.iterations
        assert np.array_equal(a.x, c.x) and np.array_equal(a.y, c.y)
    assert sum(r.status == "OPTIMAL" for r in runs[0]) >= len(insts) - 1


@pytest.mark.parametrize("mixed", [False, True])
def test_shared_bounds_are_bitwise_neutral(gpu_lp, monkeypatch, mixed):
    """Variable bounds held once for all instances (their scaled bounds equal
    bit for bit) give the same bits as the side-by-side layout.  With one
    instance's upper bound changed (mixed) the batch must not share them, and
    both runs take the side-by-side layout."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    from pdal_amd.generators import planted_family
    fam = planted_family(150, 300, 8, seed=6)
    insts = [fam.instance(t) for t in config4_thetas(100, 9)]
    if mixed:
        ub = insts[3]["var_ub"].copy()
        ub[: 20] = 0.5
        insts[3] = dict(insts[3], var_ub=ub)
    params = gpu_lp.default_params(termination_evaluation_frequency=K, eps_optimal_relative=1e-6,
                                   eps_feasible_relative=1e-6, iteration_limit=20000)
    monkeypatch.setenv("PDAL_BATCH_G_A", "16")
    monkeypatch.setenv("PDAL_BATCH_G_AT", "8")
    runs = []
    for flag in ("0", "1"):
        monkeypatch.setenv("PDAL_BATCH_NO_SHARED_BOUNDS", flag)
        print(f"[case] shared bounds mixed={mixed} no_shared_bounds={flag}", flush=True)
        with LpBatch(gpu_lp, insts, params) as b:
            st = b.stats()
            assert st["shared_a"] and st["shared_bounds"] == (flag == "0" and not mixed)
            b.solve()
            runs.append([b.result(k) for k in range(len(insts))])
    for a, c in zip(*runs):
        assert a.status == c.status and a.iterations == c.iterations
        assert np.array_equal(a.x, c.x) and np.array_equal(a.y, c.y)
    assert sum(r.status == "OPTIMAL" for r in runs[0]) >= len(insts) - 1


def test_shared_handoff_tracks_oracle(oracle_lp, gpu_lp, monkeypatch):
    """The last running instances (8 or fewer) finish on the block-diagonal
    engine: each instance still follows the oracle's solve (same iteration
    count, objective within 1e-6), those that finished on either side."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    from pdal_amd.generators import planted_family
    fam = planted_family(150, 300, 8, seed=6)
    count = 100
    insts = [fam.instance(t) for t in config4_thetas(count, 10)]
    params = dict(termination_evaluation_frequency=K, eps_optimal_relative=1e-6, eps_feasible_relative=1e-6,
                  iteration_limit=100000)
    print("[case] shared-A hand-off", flush=True)
    with LpBatch(gpu_lp, insts, gpu_lp.default_params(**params)) as b:
        assert b.solve() == count
        st = b.stats()
        assert st["shared_a"] and st["handoffs"] == 1
        res = [b.result(k) for k in range(count)]
    its = np.array([r.iterations for r in res])
    late = list(np.argsort(its)[-10:])  # the last to stop: 8 of them finished block-diagonally
    early = list(np.argsort(its)[:4])
    for k in late + early:
        oracle_lp.reset_rng()
        o = oracle_lp.solve(**insts[k], params=oracle_lp.default_params(**params))
        g = res[k]
        assert g.status == o.status == "OPTIMAL"
        assert g.iterations == o.iterations, k
        assert abs(g.primal_objective - o.primal_objective) <= 1e-6 * abs(o.primal_objective)
        assert rel_err(g.x, o.x) <= 1e-6
