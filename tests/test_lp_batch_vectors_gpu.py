"""New vectors for a batch's matrices (pdal_cp_al_batch_set_vectors).

A batch whose objective and bound vectors are replaced must solve exactly as a
batch created with the new vectors: the same bits of x and y, the same
iteration counts and statuses.  That holds only if the update replays the
set-up's scaling passes in order, recomputes the norms and the bound and
objective scales, resets the step schedule and redetects the shared-A
engine's layouts (shared bounds, equality rows).  Lane counts are pinned so
both batches launch the same kernels.  One case also checks the updated batch
against the oracle solving each new instance alone.
"""
from __future__ import annotations

import numpy as np
import pytest

from pdal_amd.generators import planted_family, planted_lp
from pdal_amd.lp import LpBatch

pytestmark = pytest.mark.gpu

K = 64


def family_instances(fam, count, seed, var_ub=None):
    th = np.random.default_rng(seed).standard_normal((count, 2 * fam.p))
    out = [fam.instance(t) for t in th]
    if var_ub is not None:
        for k, kw in enumerate(out):
            kw["var_ub"] = var_ub(k, th[k])
    return th, out


def assert_same_bits(a: LpBatch, b: LpBatch, count: int):
    for k in range(count):
        ra, rb = a.result(k), b.result(k)
        assert ra.status == rb.status and ra.iterations == rb.iterations, k
        assert np.array_equal(ra.x, rb.x) and np.array_equal(ra.y, rb.y), k
        assert ra.primal_objective == rb.primal_objective, k


@pytest.fixture
def pinned(monkeypatch):
    monkeypatch.setenv("PDAL_BATCH_G_A", "4")
    monkeypatch.setenv("PDAL_BATCH_G_AT", "4")
    return monkeypatch


@pytest.mark.parametrize("case", ["shared", "block_diagonal", "curtis_reid", "linf", "no_bound_objective",
                                  "no_ruiz"])
def test_set_vectors_equals_fresh_batch(gpu_lp, pinned, case):
    fam = planted_family(240, 480, 6, seed=7, p=4)
    count = 12
    kw = dict(termination_evaluation_frequency=K, eps_optimal_relative=1e-7, eps_feasible_relative=1e-7,
              iteration_limit=200000)
    if case == "block_diagonal":
        pinned.setenv("PDAL_BATCH_NO_SHARED", "1")
    if case == "curtis_reid":
        pinned.setenv("CP_AL_CR_SCALING", "1")
    if case == "linf":
        kw["optimality_norm"] = "linf"
    if case == "no_bound_objective":
        kw["bound_objective_rescaling"] = False
    if case == "no_ruiz":
        kw["l_inf_ruiz_iterations"] = 0
        kw["has_pock_chambolle_alpha"] = False
    params = gpu_lp.default_params(**kw)
    _, first = family_instances(fam, count, seed=11)
    _, second = family_instances(fam, count, seed=12)
    with LpBatch(gpu_lp, first, params) as upd, LpBatch(gpu_lp, second, params) as fresh:
        assert upd.stats()["shared_a"] == (case != "block_diagonal")
        upd.solve()
        upd.set_instances(second)
        upd.solve()
        fresh.solve()
        assert_same_bits(upd, fresh, count)
        assert sum(upd.result(k).status == "OPTIMAL" for k in range(count)) == count
        # and back: a second update equals a fresh batch of the first vectors
        upd.set_instances(first)
        upd.solve()
        with LpBatch(gpu_lp, first, params) as fresh1:
            fresh1.solve()
            assert_same_bits(upd, fresh1, count)


def test_set_vectors_changes_the_bound_layout(gpu_lp, pinned):
    """Variable bounds that differ per instance switch the shared-A engine
    from bounds held once to bounds side by side, and back."""
    fam = planted_family(200, 400, 6, seed=8, p=4)
    count = 8
    params = gpu_lp.default_params(termination_evaluation_frequency=K, eps_optimal_relative=1e-7,
                                   eps_feasible_relative=1e-7, iteration_limit=200000)

    def ub(k, th):  # loose finite upper bounds, different per instance: the planted optimum stays inside
        return 3.0 * fam.solution(th)[0] + 1.0 + 0.1 * k

    _, plain = family_instances(fam, count, seed=21)
    _, boxed = family_instances(fam, count, seed=22, var_ub=ub)
    with LpBatch(gpu_lp, plain, params) as upd:
        assert upd.stats()["shared_bounds"]
        upd.solve()
        upd.set_instances(boxed)
        with LpBatch(gpu_lp, boxed, params) as fresh:
            assert not fresh.stats()["shared_bounds"]
            assert not upd.stats()["shared_bounds"]
            upd.solve()
            fresh.solve()
            assert_same_bits(upd, fresh, count)
        upd.set_instances(plain)
        assert upd.stats()["shared_bounds"]
        upd.solve()
        with LpBatch(gpu_lp, plain, params) as fresh:
            fresh.solve()
            assert_same_bits(upd, fresh, count)


def test_set_vectors_of_distinct_matrices(gpu_lp, pinned):
    """A block-diagonal batch of different matrices: the factors of every
    instance are replayed (five blocks, bits compared)."""
    insts = [planted_lp(60 + 10 * i, 120 + 20 * i, 6, seed=300 + i) for i in range(5)]
    rng = np.random.default_rng(5)
    params = gpu_lp.default_params(termination_evaluation_frequency=K, iteration_limit=5 * K)
    first = [i.lp_args() for i in insts]
    second = []
    for a in first:
        kw = dict(a)
        kw["c"] = a["c"] * (1.0 + 0.05 * rng.standard_normal(a["c"].size))
        kw["con_lb"] = kw["con_ub"] = a["con_ub"] * (1.0 + 0.05 * rng.standard_normal(a["con_ub"].size))
        second.append(kw)
    with LpBatch(gpu_lp, first, params) as upd, LpBatch(gpu_lp, second, params) as fresh:
        assert not upd.stats()["shared_a"]
        upd.set_instances(second)
        upd.solve()
        fresh.solve()
        assert_same_bits(upd, fresh, len(insts))


def test_set_vectors_device_and_warm_starts(gpu_lp, pinned):
    """Device vectors equal host vectors; starts set after the update apply to
    the new vectors as on a fresh batch."""
    from pdal_amd.devbuf import DeviceArray
    fam = planted_family(200, 400, 6, seed=9, p=4)
    count = 6
    params = gpu_lp.default_params(termination_evaluation_frequency=K, eps_optimal_relative=1e-7,
                                   eps_feasible_relative=1e-7, iteration_limit=200000)
    th0, first = family_instances(fam, count, seed=31)
    th1, second = family_instances(fam, count, seed=32)
    c = DeviceArray.from_host(np.concatenate([kw["c"] for kw in second]))
    b = DeviceArray.from_host(np.concatenate([kw["con_lb"] for kw in second]))
    rng = np.random.default_rng(3)
    xs = np.concatenate([fam.solution(t)[0] for t in th1]) * (1.0 + 0.01 * rng.standard_normal(count * fam.n))
    ys = np.concatenate([fam.solution(t)[1] for t in th1]) * (1.0 + 0.01 * rng.standard_normal(count * fam.m))
    with LpBatch(gpu_lp, first, params) as upd, LpBatch(gpu_lp, second, params) as fresh:
        upd.set_vectors_device(c=c, con_lb=b, con_ub=b)
        c.close()
        b.close()
        upd.set_starts(xs, ys)
        fresh.set_starts(xs, ys)
        upd.solve()
        fresh.solve()
        assert_same_bits(upd, fresh, count)
        warm = np.mean([upd.result(k).iterations for k in range(count)])
        fresh.set_vectors()  # nothing replaced: zero starts, same vectors
        fresh.solve()
        cold = np.mean([fresh.result(k).iterations for k in range(count)])
        assert warm < cold


def test_updated_batch_tracks_oracle(oracle_lp, gpu_lp, pinned):
    """After an update every instance terminates in the oracle's block with
    the oracle's status, its objective within 1e-6 of the planted optimum."""
    fam = planted_family(160, 320, 6, seed=10, p=4)
    count = 4
    params = gpu_lp.default_params(termination_evaluation_frequency=K, eps_optimal_relative=1e-7,
                                   eps_feasible_relative=1e-7, iteration_limit=200000)
    _, first = family_instances(fam, count, seed=41)
    th, second = family_instances(fam, count, seed=42)
    with LpBatch(gpu_lp, first, params) as upd:
        upd.set_instances(second)
        upd.solve()
        for k in range(count):
            oracle_lp.reset_rng()
            o = oracle_lp.solve(**second[k], params=params)
            g = upd.result(k)
            assert g.status == o.status == "OPTIMAL" and g.iterations == o.iterations
            opt = fam.optimal_value(th[k])
            assert abs(g.primal_objective - opt) <= 1e-6 * (1.0 + abs(opt))
