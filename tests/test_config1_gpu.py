"""Plain CP_AL on the large planted LP family: Zipf(1.5) column counts
clipped to [2, 64], uniform rows, N(0,1) values, planted x*, y*, s.

* At a size the oracle finishes in seconds (m = 20,000, n = 40,000), the HIP
  iterates track the oracle within 1e-10 relative through 1000 iterations,
  with the default SpMV passes and with the column-sliced passes forced on
  (small slices, so every row spans many of them).
* At the full size (m = 2e6, n = 4e6, 48.3M nonzeros) one solve to relative
  KKT 1e-6, checked by properties recomputed here in numpy on the unscaled
  problem: bounds, primal residual, the negative part of the reduced cost,
  the duality gap and the planted objective.
"""
from __future__ import annotations

import numpy as np
import pytest

from pdal_amd.generators import config1, zipf_planted_lp

pytestmark = pytest.mark.gpu

ITER_RTOL = 1e-10


def rel_err(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def zipf_mid():
    return zipf_planted_lp(20000, 40000, seed=1)


@pytest.mark.parametrize("sliced", ["auto", "forced"])
@pytest.mark.parametrize("iters", [200, 1000])
def test_zipf_iterates_track_oracle(oracle_lp, gpu_lp, monkeypatch, zipf_mid, sliced, iters):
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    if sliced == "forced":
        monkeypatch.setenv("PDAL_SPMV_SLICED", "1")
        monkeypatch.setenv("PDAL_SPMV_SLICE_COLS", "1500")  # 14 slices of y, 27 of x_bar
        monkeypatch.setenv("PDAL_SPMV_SLICE_RPW", "64")
    out = []
    for lib in (oracle_lp, gpu_lp):
        lib.reset_rng()
        p = lib.default_params(iteration_limit=iters, eps_optimal_relative=1e-14, eps_feasible_relative=1e-14)
        print(f"[case] zipf m=20000 sliced={sliced} iters={iters} lib={lib.path.name}", flush=True)
        out.append(lib.solve(**zipf_mid.lp_args(), params=p))
    o, g = out
    assert o.iterations == g.iterations == iters
    assert rel_err(g.x, o.x) <= ITER_RTOL
    assert rel_err(g.y, o.y) <= ITER_RTOL
    assert rel_err(g.reduced_cost, o.reduced_cost) <= ITER_RTOL


def test_sliced_passes_are_chosen_and_equal_on_large_vectors(gpu_lp, monkeypatch):
    """On the full-size LP the automatic choice takes the sliced passes (the
    gathered vectors are 16 and 32 MB), and forcing them off changes the
    iterates only by summation order.  The generator is reseeded before each
    set-up and the CSR lane counts are pinned: the power iteration and the
    evaluation passes use them, and another start vector or lane count moves
    the power iteration's stopping step and with it the step size by ~1e-8."""
    from pdal_amd.lp import LpSession
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    monkeypatch.setenv("PDAL_SPMV_G_AT", "16")
    monkeypatch.setenv("PDAL_SPMV_G_A", "32")
    inst = config1(seed=1)
    res = {}
    for mode in ("auto", "0"):
        if mode == "0":
            monkeypatch.setenv("PDAL_SPMV_SLICED", "0")
        print(f"[case] config1 sliced={mode}", flush=True)
        p = gpu_lp.default_params(iteration_limit=400, eps_optimal_relative=1e-14, eps_feasible_relative=1e-14)
        gpu_lp.reset_rng()  # the session's power iteration draws its start vector at set-up
        with LpSession(gpu_lp, **inst.lp_args(), params=p) as sess:
            kt = sess.time_kernels(3)
            res[mode] = (kt, sess.solve())
        print(f"[kernels] {mode}: {kt}", flush=True)
    assert res["auto"][0]["lanes_at"] < 0 and res["auto"][0]["lanes_a"] < 0  # sliced passes in use
    assert res["0"][0]["lanes_at"] > 0 and res["0"][0]["lanes_a"] > 0
    a, b = res["auto"][1], res["0"][1]
    assert a.iterations == b.iterations == 400
    assert rel_err(a.x, b.x) <= ITER_RTOL
    assert rel_err(a.y, b.y) <= ITER_RTOL


def test_config1_full_solve_properties(gpu_lp, monkeypatch):
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    tol = 1e-6
    inst = config1(seed=1)
    A, b, c = inst.A, inst.b, inst.c
    print(f"[case] config1 full solve m={inst.m} n={inst.n} nnz={A.nnz} tol={tol}", flush=True)
    gpu_lp.reset_rng()
    p = gpu_lp.default_params(eps_optimal_relative=tol, eps_feasible_relative=tol, time_sec_limit=600.0)
    r = gpu_lp.solve(**inst.lp_args(), params=p)
    print(f"[solve] {r.status} iterations={r.iterations} time={r.solve_time_sec:.1f}s kkt={r.kkt:.3e}", flush=True)
    assert r.status == "OPTIMAL"
    assert r.kkt < tol
    x, y = r.x, r.y
    # bounds hold exactly (the iterate is a projection onto x >= 0)
    assert x.min() >= 0.0
    # unscaled primal residual of the equality rows, in the solver's L2 form
    rp = np.linalg.norm(A @ x - b) / (1.0 + np.linalg.norm(b))
    # dual residual: the reduced cost c - A^T y must be >= 0 on x >= 0 (ub = inf)
    lam = c - A.T @ y
    rd = np.linalg.norm(np.minimum(lam, 0.0)) / (1.0 + np.linalg.norm(c))
    # the reported reduced cost is that vector clamped at 0 (ub = inf, solver.cu:3828-3909)
    assert np.linalg.norm(r.reduced_cost - np.maximum(lam, 0.0)) <= 1e-9 * np.linalg.norm(c)
    pobj, dobj = float(c @ x), float(b @ y)
    gap = abs(pobj - dobj) / (1.0 + abs(pobj) + abs(dobj))
    print(f"[numpy] rel primal {rp:.3e} rel dual {rd:.3e} gap {gap:.3e}", flush=True)
    # the solver tests these in its scaled norms; unscaled they agree to within
    # the row / column scaling spread, which stays below a factor 10 here
    assert rp < 10 * tol and rd < 10 * tol and gap < 10 * tol
    assert abs(r.primal_objective - pobj) <= 1e-9 * abs(pobj)
    # planted optimum: |p - opt| <= |p - d| + |d - opt|, and both are bounded by
    # the gap and the residual terms; the KKT tolerance implies 2 * tol relative
    opt = inst.optimal_value
    err = abs(pobj - opt) / abs(opt)
    print(f"[planted] objective {pobj:.12e} planted {opt:.12e} rel err {err:.3e}", flush=True)
    assert err <= 2 * tol
