"""Batched CP_AL (pdal_batch.h) vs the FP64 oracle solving each instance alone.

Every instance of a batch must follow the single-instance algorithm: the same
iterates within 1e-10 relative after a fixed number of evaluation blocks, the
same termination block and status when solved, the objective within 1e-6 of
the oracle and of the planted optimum.  The oracle's generator is reseeded
before each instance, which is what the batch does for every instance.
"""
from __future__ import annotations

import numpy as np
import pytest

from pdal_amd.generators import config3_instance, planted_lp, kkt_objective_bar
from pdal_amd.lp import LpBatch

pytestmark = pytest.mark.gpu

ITER_RTOL = 1e-10
OBJ_RTOL = 1e-6
K = 64


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def mixed_instances(count=10, seed=0):
    """Uneven sizes (m in [40, 600], n = 2m, 8 nonzeros per column)."""
    rng = np.random.default_rng(seed)
    ms = np.exp(rng.uniform(np.log(40), np.log(600), count)).astype(int)
    return [planted_lp(int(m), 2 * int(m), 8, seed=100 + i) for i, m in enumerate(ms)]


def oracle_solve(oracle_lp, inst, params, start=None):
    oracle_lp.reset_rng()
    kw = {} if start is None else dict(primal_start=start[0], dual_start=start[1])
    return oracle_lp.solve(**inst.lp_args(), params=params, **kw)


@pytest.mark.parametrize("lam", ["0", "0.05"])
@pytest.mark.parametrize("blocks", [1, 5, 16])
def test_batch_iterates_track_oracle(oracle_lp, gpu_lp, monkeypatch, lam, blocks):
    monkeypatch.setenv("CP_AL_LAMBDA", lam)
    insts = mixed_instances()
    params = dict(iteration_limit=blocks * K, termination_evaluation_frequency=K, eps_optimal_relative=1e-14,
                  eps_feasible_relative=1e-14)
    print(f"[case] batch lam={lam} blocks={blocks} sizes={[i.m for i in insts]}", flush=True)
    with LpBatch(gpu_lp, [i.lp_args() for i in insts], gpu_lp.default_params(**params)) as b:
        assert b.solve() == 0
        for k, inst in enumerate(insts):
            o = oracle_solve(oracle_lp, inst, oracle_lp.default_params(**params))
            g = b.result(k)
            assert g.status == o.status == "ITERATION_LIMIT"
            assert g.iterations == o.iterations == blocks * K
            assert rel_err(g.x, o.x) <= ITER_RTOL, k
            assert rel_err(g.y, o.y) <= ITER_RTOL, k
            assert rel_err(g.reduced_cost, o.reduced_cost) <= ITER_RTOL, k


@pytest.mark.parametrize("tol", [1e-4, 1e-7])
def test_batch_solves_match_oracle_and_planted(oracle_lp, gpu_lp, monkeypatch, tol):
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    insts = mixed_instances(count=8, seed=1)
    params = dict(termination_evaluation_frequency=K, eps_optimal_relative=tol, eps_feasible_relative=tol,
                  iteration_limit=200000)
    print(f"[case] batch solve tol={tol}", flush=True)
    with LpBatch(gpu_lp, [i.lp_args() for i in insts], gpu_lp.default_params(**params)) as b:
        assert b.solve() == len(insts)
        st = b.stats()
        assert st["optimal"] == len(insts) and st["instances"] == len(insts)
        for k, inst in enumerate(insts):
            o = oracle_solve(oracle_lp, inst, oracle_lp.default_params(**params))
            g = b.result(k)
            assert g.status == o.status == "OPTIMAL"
            assert g.iterations == o.iterations
            assert abs(g.primal_objective - o.primal_objective) <= OBJ_RTOL * abs(o.primal_objective)
            # the bar the KKT test at tol implies for the planted optimum (gap allowance plus what
            # the residuals can move an objective through the optimal pair); at 1e-7 also the 1e-6
            # relative bar of the planted optimum
            bar = kkt_objective_bar(g.primal_objective, g.dual_objective, inst.x_star, inst.y_star, inst.b,
                                    inst.c, tol)
            assert abs(g.primal_objective - inst.optimal_value) <= bar
            if tol <= 1e-7:
                assert abs(g.primal_objective - inst.optimal_value) <= 1e-6 * abs(inst.optimal_value)
            assert max(g.relative_primal_residual, g.relative_dual_residual, g.relative_objective_gap) < tol


def test_batch_equals_single_instance_gpu_solves(gpu_lp, monkeypatch):
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    insts = mixed_instances(count=6, seed=2)
    params = gpu_lp.default_params(termination_evaluation_frequency=K, eps_optimal_relative=1e-6,
                                   eps_feasible_relative=1e-6, iteration_limit=200000)
    with LpBatch(gpu_lp, [i.lp_args() for i in insts], params) as b:
        b.solve()
        for k, inst in enumerate(insts):
            gpu_lp.reset_rng()
            s = gpu_lp.solve(**inst.lp_args(), params=params)
            g = b.result(k)
            assert g.status == s.status and g.iterations == s.iterations
            assert rel_err(g.x, s.x) <= 1e-9


def test_repeated_batch_solves_are_bitwise_identical(gpu_lp, monkeypatch):
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    insts = mixed_instances(count=5, seed=3)
    params = gpu_lp.default_params(termination_evaluation_frequency=K, eps_optimal_relative=1e-6,
                                   eps_feasible_relative=1e-6)
    with LpBatch(gpu_lp, [i.lp_args() for i in insts], params) as b:
        b.solve()
        first = [b.result(k) for k in range(len(insts))]
        b.solve()
        for k in range(len(insts)):
            r = b.result(k)
            assert np.array_equal(r.x, first[k].x) and np.array_equal(r.y, first[k].y)
            assert r.iterations == first[k].iterations


def test_batch_warm_start_tracks_oracle(oracle_lp, gpu_lp, monkeypatch):
    """Starts given through set_starts (concatenated, unscaled) and through the
    problems' own primal_start/dual_start give the oracle's warm-started
    trajectories."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    insts = mixed_instances(count=6, seed=4)
    rng = np.random.default_rng(9)
    starts = [(i.x_star + 0.05 * rng.standard_normal(i.n), i.y_star + 0.05 * rng.standard_normal(i.m)) for i in insts]
    params = dict(termination_evaluation_frequency=K, iteration_limit=4 * K, eps_optimal_relative=1e-14,
                  eps_feasible_relative=1e-14)
    oracle = [oracle_solve(oracle_lp, inst, oracle_lp.default_params(**params), start=s)
              for inst, s in zip(insts, starts)]
    # (a) through set_starts
    with LpBatch(gpu_lp, [i.lp_args() for i in insts], gpu_lp.default_params(**params)) as b:
        b.set_starts(np.concatenate([s[0] for s in starts]), np.concatenate([s[1] for s in starts]))
        b.solve()
        for k in range(len(insts)):
            g = b.result(k)
            assert rel_err(g.x, oracle[k].x) <= ITER_RTOL and rel_err(g.y, oracle[k].y) <= ITER_RTOL
    # (b) through the problems' start values
    kws = [dict(i.lp_args(), primal_start=s[0], dual_start=s[1]) for i, s in zip(insts, starts)]
    with LpBatch(gpu_lp, kws, gpu_lp.default_params(**params)) as b:
        b.solve()
        for k in range(len(insts)):
            g = b.result(k)
            assert rel_err(g.x, oracle[k].x) <= ITER_RTOL and rel_err(g.y, oracle[k].y) <= ITER_RTOL


@pytest.mark.parametrize("norm", ["l2", "linf"])
def test_batch_norms_and_short_limit(oracle_lp, gpu_lp, monkeypatch, norm):
    """L-inf / L2 residuals; an iteration limit below one block still runs one
    whole block, as the single-instance loop does."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    insts = mixed_instances(count=4, seed=5)
    params = dict(termination_evaluation_frequency=K, iteration_limit=10, optimality_norm=norm)
    with LpBatch(gpu_lp, [i.lp_args() for i in insts], gpu_lp.default_params(**params)) as b:
        b.solve()
        for k, inst in enumerate(insts):
            o = oracle_solve(oracle_lp, inst, oracle_lp.default_params(**params))
            g = b.result(k)
            assert g.iterations == o.iterations == K
            assert rel_err(g.x, o.x) <= ITER_RTOL
            assert g.relative_primal_residual == pytest.approx(o.relative_primal_residual, rel=1e-8)


def test_batch_of_config3_family(oracle_lp, gpu_lp, monkeypatch):
    """A few instances of the 4096-instance family (m log-uniform, n = 2m,
    8 nonzeros per column, seed = instance index) solved to 1e-7, tight enough
    for the 1e-6 relative bar on the planted objective."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    ms = [200, 1000, 350]
    insts = [config3_instance(i, m) for i, m in enumerate(ms)]
    params = dict(termination_evaluation_frequency=K, eps_optimal_relative=1e-7, eps_feasible_relative=1e-7,
                  iteration_limit=500000)
    with LpBatch(gpu_lp, [i.lp_args() for i in insts], gpu_lp.default_params(**params)) as b:
        assert b.solve() == len(insts)
        for k, inst in enumerate(insts):
            o = oracle_solve(oracle_lp, inst, oracle_lp.default_params(**params))
            g = b.result(k)
            assert g.iterations == o.iterations
            assert abs(g.primal_objective - inst.optimal_value) <= 1e-6 * abs(inst.optimal_value)


def test_compaction_does_not_change_results(gpu_lp, monkeypatch):
    """Dropping stopped instances from the launch tables (and re-capturing the
    graph) is invisible in the results: identical bits with it switched off.
    The lane counts are pinned: the autotuner picks them by timing, and the
    lane count sets the summation order of a row."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    monkeypatch.setenv("PDAL_BATCH_G_A", "8")
    monkeypatch.setenv("PDAL_BATCH_G_AT", "4")
    insts = mixed_instances(count=12, seed=6)
    params = gpu_lp.default_params(termination_evaluation_frequency=K, eps_optimal_relative=1e-6,
                                   eps_feasible_relative=1e-6)
    runs = []
    for flag in ("0", "1"):
        monkeypatch.setenv("PDAL_BATCH_NO_COMPACT", flag)
        with LpBatch(gpu_lp, [i.lp_args() for i in insts], params) as b:
            b.solve()
            runs.append(([b.result(k) for k in range(len(insts))], b.stats()))
    (with_c, st_c), (without, st_n) = runs
    assert st_c["compactions"] >= 1 and st_n["compactions"] == 0
    for a, b_ in zip(with_c, without):
        assert a.iterations == b_.iterations and np.array_equal(a.x, b_.x) and np.array_equal(a.y, b_.y)


def test_direct_launches_match_graph_replays(gpu_lp, monkeypatch):
    """PDAL_NO_GRAPH=1 launches each block's kernels directly instead of
    replaying the captured graph: same kernels, same bits."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    monkeypatch.setenv("PDAL_BATCH_G_A", "8")
    monkeypatch.setenv("PDAL_BATCH_G_AT", "4")
    insts = mixed_instances(count=6, seed=7)
    params = gpu_lp.default_params(termination_evaluation_frequency=K, eps_optimal_relative=1e-6,
                                   eps_feasible_relative=1e-6)
    runs = []
    for flag in ("0", "1"):
        monkeypatch.setenv("PDAL_NO_GRAPH", flag)
        with LpBatch(gpu_lp, [i.lp_args() for i in insts], params) as b:
            b.solve()
            runs.append([b.result(k) for k in range(len(insts))])
    for a, b_ in zip(*runs):
        assert a.iterations == b_.iterations and np.array_equal(a.x, b_.x) and np.array_equal(a.y, b_.y)


def test_staged_kernels_equal_table_kernels(gpu_lp, monkeypatch):
    """The LDS-staged iteration kernels (one workgroup per instance, gathered
    segment in LDS) sum every row in the table kernels' order: identical bits,
    iterations and statuses with staging forced on for every instance and
    switched off.  Lanes are pinned (they set the summation order)."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    monkeypatch.setenv("PDAL_BATCH_G_A", "8")
    monkeypatch.setenv("PDAL_BATCH_G_AT", "4")
    monkeypatch.setenv("PDAL_BATCH_MIN_STAGED", "1")
    # small instances (two workgroups per CU) and one whose segments need more
    # than half the LDS (82400 B of y, 163200 B of x_bar)
    insts = mixed_instances(count=9, seed=11) + [planted_lp(10300, 20400, 8, seed=77)]
    params = gpu_lp.default_params(termination_evaluation_frequency=K, eps_optimal_relative=1e-6,
                                   eps_feasible_relative=1e-6, iteration_limit=20 * K)
    runs = []
    for flag in ("0", "1"):
        monkeypatch.setenv("PDAL_BATCH_NO_STAGE", flag)
        print(f"[case] staged={flag == '0'} sizes={[i.m for i in insts]}", flush=True)
        with LpBatch(gpu_lp, [i.lp_args() for i in insts], params) as b:
            b.solve()
            runs.append([b.result(k) for k in range(len(insts))])
            st = b.stats()
            if flag == "0":
                assert st["staged_at"] == len(insts) and st["staged_a"] == len(insts), st
            else:
                assert st["staged_at"] == st["staged_a"] == st["split_a"] == 0, st
    for a, b_ in zip(*runs):
        assert a.status == b_.status and a.iterations == b_.iterations
        assert np.array_equal(a.x, b_.x) and np.array_equal(a.y, b_.y)


@pytest.mark.parametrize("lam", ["0", "0.05"])
def test_staged_iterates_track_oracle(oracle_lp, gpu_lp, monkeypatch, lam):
    """Staging forced on: 5 evaluation blocks of every instance within 1e-10
    of the oracle (the predictor runs as a table kernel when lambda > 0)."""
    monkeypatch.setenv("CP_AL_LAMBDA", lam)
    monkeypatch.setenv("PDAL_BATCH_MIN_STAGED", "1")
    insts = mixed_instances(count=6, seed=12)
    params = dict(iteration_limit=5 * K, termination_evaluation_frequency=K, eps_optimal_relative=1e-14,
                  eps_feasible_relative=1e-14)
    with LpBatch(gpu_lp, [i.lp_args() for i in insts], gpu_lp.default_params(**params)) as b:
        assert b.solve() == 0
        for k, inst in enumerate(insts):
            o = oracle_solve(oracle_lp, inst, oracle_lp.default_params(**params))
            g = b.result(k)
            assert g.iterations == o.iterations == 5 * K
            assert rel_err(g.x, o.x) <= ITER_RTOL, k
            assert rel_err(g.y, o.y) <= ITER_RTOL, k


@pytest.mark.parametrize("lam", ["0", "0.05"])
def test_split_dual_tracks_oracle(oracle_lp, gpu_lp, monkeypatch, lam):
    """A.x_bar staged in two column halves (x_bar of 20600 doubles exceeds one
    160 KB LDS load): 5 evaluation blocks within 1e-10 of the oracle, next to
    small fully staged instances.  The streams fork: the split kernel and the
    two staged classes run concurrently."""
    monkeypatch.setenv("CP_AL_LAMBDA", lam)
    monkeypatch.setenv("PDAL_BATCH_MIN_STAGED", "1")
    monkeypatch.setenv("PDAL_BATCH_SPLIT", "1")  # opt-in path
    insts = [planted_lp(10300, 20600, 8, seed=78)] + mixed_instances(count=3, seed=13)
    params = dict(iteration_limit=5 * K, termination_evaluation_frequency=K, eps_optimal_relative=1e-14,
                  eps_feasible_relative=1e-14)
    print(f"[case] split dual lam={lam} sizes={[i.m for i in insts]}", flush=True)
    with LpBatch(gpu_lp, [i.lp_args() for i in insts], gpu_lp.default_params(**params)) as b:
        st = b.stats()
        assert st["split_a"] == 1 and st["staged_a"] == len(insts) - 1, st
        assert b.solve() == 0
        for k, inst in enumerate(insts):
            o = oracle_solve(oracle_lp, inst, oracle_lp.default_params(**params))
            g = b.result(k)
            assert g.iterations == o.iterations == 5 * K
            assert rel_err(g.x, o.x) <= ITER_RTOL, k
            assert rel_err(g.y, o.y) <= ITER_RTOL, k


def test_overlapped_streams_match_one_stream(gpu_lp, monkeypatch):
    """The staged classes and the table kernel of one pass run on forked
    streams; the results are the bits of the one-stream run, with graphs and
    with direct launches."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    monkeypatch.setenv("PDAL_BATCH_G_A", "8")
    monkeypatch.setenv("PDAL_BATCH_G_AT", "4")
    monkeypatch.setenv("PDAL_BATCH_MIN_STAGED", "1")
    # small (half-LDS class), full-LDS class, and too large to stage for A.x_bar
    insts = mixed_instances(count=6, seed=14) + [planted_lp(10300, 20400, 8, seed=79),
                                                 planted_lp(10400, 20800, 8, seed=80)]
    params = gpu_lp.default_params(termination_evaluation_frequency=K, eps_optimal_relative=1e-6,
                                   eps_feasible_relative=1e-6, iteration_limit=12 * K)
    runs = []
    for overlap, graph in (("1", "0"), ("0", "0"), ("0", "1")):
        monkeypatch.setenv("PDAL_BATCH_NO_OVERLAP", overlap)
        monkeypatch.setenv("PDAL_NO_GRAPH", graph)
        print(f"[case] no_overlap={overlap} no_graph={graph}", flush=True)
        with LpBatch(gpu_lp, [i.lp_args() for i in insts], params) as b:
            b.solve()
            runs.append([b.result(k) for k in range(len(insts))])
    for other in runs[1:]:
        for a, b_ in zip(runs[0], other):
            assert a.status == b_.status and a.iterations == b_.iterations
            assert np.array_equal(a.x, b_.x) and np.array_equal(a.y, b_.y)


@pytest.mark.parametrize("cr_iters", ["5", "20"])
def test_batch_curtis_reid_tracks_oracle(oracle_lp, gpu_lp, monkeypatch, cr_iters):
    """Curtis-Reid scaling before Ruiz (CP_AL_CR_SCALING, preconditioner.cu:162-268)
    over the block-diagonal batch: each instance's iterates equal the oracle's
    single-instance CR solve."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    monkeypatch.setenv("CP_AL_CR_SCALING", "1")
    monkeypatch.setenv("CP_AL_CR_ITERS", cr_iters)
    insts = mixed_instances(count=6, seed=4)
    params = dict(iteration_limit=5 * K, termination_evaluation_frequency=K, eps_optimal_relative=1e-14,
                  eps_feasible_relative=1e-14)
    print(f"[case] batch curtis-reid iters={cr_iters}", flush=True)
    with LpBatch(gpu_lp, [i.lp_args() for i in insts], gpu_lp.default_params(**params)) as b:
        b.solve()
        for k, inst in enumerate(insts):
            o = oracle_solve(oracle_lp, inst, oracle_lp.default_params(**params))
            g = b.result(k)
            assert g.iterations == o.iterations == 5 * K
            assert rel_err(g.x, o.x) <= ITER_RTOL, k
            assert rel_err(g.y, o.y) <= ITER_RTOL, k
