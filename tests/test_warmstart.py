"""Warm-start MLP kernels (libpdal_ws.so) against the CPU oracle
(oracle/mlp_oracle.py), and the oracle against finite differences."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import mlp_oracle as MO

ROOT = Path(__file__).resolve().parent.parent
WIDTHS = [64, 512, 512, 512, 256]


def test_oracle_gradient_matches_finite_differences():
    """FP64 oracle (rounding off): analytic gradient = central differences on
    a small network, for weights and biases of every layer."""
    rng = np.random.default_rng(0)
    widths = [5, 7, 6, 3]
    p = MO.init_params(widths, rng) + 0.05 * rng.standard_normal(sum(
        widths[l + 1] * widths[l] + widths[l + 1] for l in range(len(widths) - 1)))
    X = rng.standard_normal((9, 5))
    Y = rng.standard_normal((9, 3))
    _, g = MO.loss_grad(p, widths, X, Y, bf16=False)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (MO.loss_grad(p + e, widths, X, Y, bf16=False)[0] - MO.loss_grad(p - e, widths, X, Y, bf16=False)[0]) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9)


def test_bf16_rounding_ties_to_even():
    vals = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, 1.0 + 2**-9, -2.5], dtype=np.float32)
    got = MO.bf16_round(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1.0); 1 + 3*2^-8 -> 1 + 2^-6 (even)
    np.testing.assert_array_equal(got, [1.0, 1.0, 1.0 + 2**-6, 1.0, -2.5])


def test_library_exports_every_declared_symbol():
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "pdal_warmstart.h").read_text(), flags=re.S)
    declared = set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text))
    out = subprocess.run(["nm", "-D", "--defined-only", str(ROOT / "pdal_amd" / "lib" / "libpdal_ws.so")],
                         check=True, capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert declared and not (declared - exported)


@pytest.fixture(scope="module")
def ws():
    from pdal_amd.warmstart import load_ws
    return load_ws()


@pytest.mark.gpu
def test_forward_and_gradient_match_oracle(ws):
    rng = np.random.default_rng(1)
    mlp = ws.create(WIDTHS, seed=7)
    try:
        p = mlp.params()
        ref = MO.init_params(WIDTHS, np.random.default_rng(2))
        mlp.set_params(ref.astype(np.float32))
        assert np.array_equal(mlp.params(), ref.astype(np.float32))
        p = ref.astype(np.float32).astype(np.float64)
        X = rng.standard_normal((256, 64)).astype(np.float32)
        Y = rng.standard_normal((256, 256)).astype(np.float32)
        out = mlp.forward(X)
        H = MO.forward(p, WIDTHS, X)
        assert np.linalg.norm(out - H[-1]) <= 1e-4 * np.linalg.norm(H[-1])
        loss, g = mlp.loss_grad(X, Y)
        loss_o, g_o = MO.loss_grad(p, WIDTHS, X, Y)
        assert loss == pytest.approx(loss_o, rel=1e-4)
        for (gW, gb), (oW, ob) in zip(MO.split_params(g, WIDTHS), MO.split_params(g_o, WIDTHS)):
            for a, b in ((gW, oW), (gb, ob)):
                assert np.linalg.norm(a - b) <= 2e-3 * np.linalg.norm(b)
                assert float(a.ravel() @ b.ravel()) / (np.linalg.norm(a) * np.linalg.norm(b)) >= 0.99999
    finally:
        mlp.close()


@pytest.mark.gpu
@pytest.mark.parametrize("batch", [1, 37, 100])
def test_ragged_batches_match_oracle(ws, batch):
    """Batch sizes that are not multiples of the 64-row tile."""
    rng = np.random.default_rng(batch)
    mlp = ws.create([64, 96, 40], seed=3)
    try:
        p = mlp.params().astype(np.float64)
        X = rng.standard_normal((batch, 64)).astype(np.float32)
        Y = rng.standard_normal((batch, 40)).astype(np.float32)
        ref = MO.forward(p, [64, 96, 40], X)[-1]
        out = mlp.forward(X)
        assert np.linalg.norm(out - ref) <= 1e-4 * np.linalg.norm(ref)
        _, g = mlp.loss_grad(X, Y)
        _, g_o = MO.loss_grad(p, [64, 96, 40], X, Y)
        assert np.linalg.norm(g - g_o) <= 2e-3 * np.linalg.norm(g_o)
    finally:
        mlp.close()


@pytest.mark.gpu
def test_training_fits_a_teacher(ws):
    """Targets produced by a random teacher network: Adam on the device drives
    the training loss down by more than 10x."""
    rng = np.random.default_rng(4)
    widths = [64, 256, 256, 32]
    teacher = MO.init_params(widths, rng) * 0.5
    X = rng.standard_normal((2048, 64)).astype(np.float32)
    Y = MO.forward(teacher, widths, X, bf16=False)[-1].astype(np.float32)
    mlp = ws.create(widths, seed=11)
    try:
        hist = mlp.train(X, Y, batch=256, epochs=60, lr=1e-3, seed=5)
        assert hist[-1] < 0.1 * hist[0]
        assert np.all(np.isfinite(hist))
    finally:
        mlp.close()


@pytest.mark.gpu
def test_pod_expand_matches_numpy(ws):
    rng = np.random.default_rng(6)
    k, dim, batch = 40, 3001, 37
    basis = rng.standard_normal((k, dim)).astype(np.float32)
    mean = rng.standard_normal(dim).astype(np.float32)
    coef = rng.standard_normal((batch, k)).astype(np.float32)
    out = ws.pod_expand(basis, mean, coef)
    ref = mean.astype(np.float64) + coef.astype(np.float64) @ basis.astype(np.float64)
    assert out.dtype == np.float64 and out.shape == (batch, dim)
    assert np.max(np.abs(out - ref)) <= 1e-5 * np.max(np.abs(ref))


def test_pod_modes_are_orthonormal_and_capture_a_low_rank_family():
    """Snapshots of rank 5 plus a mean: 5 modes reconstruct them exactly."""
    from pdal_amd.warmstart import LearnedWarmStart
    rng = np.random.default_rng(8)
    basis = rng.standard_normal((5, 300))
    S = rng.standard_normal((40, 5)) @ basis + rng.standard_normal(300)
    mean, B, energy = LearnedWarmStart._pod(S, 5)
    np.testing.assert_allclose(B @ B.T, np.eye(5), atol=1e-10)
    recon = mean + ((S - mean) @ B.T) @ B
    np.testing.assert_allclose(recon, S, atol=1e-9)
    assert energy == pytest.approx(1.0, abs=1e-12)


# Objective agreement between a cold and a warm solve of one instance, both
# stopped by the relative KKT test at 1e-6: the primal residual (~1e-7
# relative) times the dual magnitude moves the objective of these
# ill-conditioned random instances by up to ~1e-5 relative.
OBJ_AGREEMENT = 5e-5


@pytest.mark.gpu
def test_learned_warm_start_pipeline(ws, monkeypatch):
    """The fixed-A family (b = b0 + U theta_b, c = c0 + V theta_c, U and V
    Gaussian x 0.1) at m=300: 512 training instances solved cold in one batch,
    POD + MLP fitted on their solutions and final primal weights, then 32
    held-out instances solved cold and from the network's starts.

    What the trained network gains over the method without it: its held-out
    prediction error is below that of the training-mean solution (the best
    start without a network).  Every warm answer passes the same FP64 KKT test
    (or falls back to a cold solve) and matches the cold objective.  The
    iteration counts are printed, not asserted: on this family the solution
    moves across active sets over theta and the network's starts are tens of
    percent off, where a PDHG warm start saves little."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    from pdal_amd.generators import config4_thetas, lp_family
    from pdal_amd.lp import LpBatch, load_cp_al
    from pdal_amd.warmstart import LearnedWarmStart, solve_warm
    lib = load_cp_al()
    fam = lp_family(300, 600, 8, seed=21)
    tr, te = config4_thetas(512, 1), config4_thetas(32, 2)
    params = lib.default_params(termination_evaluation_frequency=64, eps_optimal_relative=1e-6,
                                eps_feasible_relative=1e-6, iteration_limit=200000)
    with LpBatch(lib, [fam.instance(t) for t in tr], params) as b:
        assert b.solve() == len(tr)
        sols = [b.result(k) for k in range(len(tr))]
        pw = b.primal_weights()
    assert np.all(np.isfinite(pw)) and np.all(pw > 0)
    X, Y = np.stack([s.x for s in sols]), np.stack([s.y for s in sols])
    lws = LearnedWarmStart(ws, kx=128, seed=3)
    try:
        info = lws.fit(tr, X, Y, pw=pw, lb=np.zeros(fam.n), epochs=150)
        assert info["loss_last"] < 0.5 * info["loss_first"]
        xh, yh, pwh = lws.predict(te)
    finally:
        lws.close()
    test_inst = [fam.instance(t) for t in te]
    with LpBatch(lib, test_inst, params) as b:
        assert b.solve() == len(te)
        cold = [b.result(k) for k in range(len(te))]
    Zc = np.hstack([np.stack([c.x for c in cold]), np.stack([c.y for c in cold])])
    err_net = np.linalg.norm(np.hstack([xh, yh]) - Zc) / np.linalg.norm(Zc)
    err_mean = np.linalg.norm(np.hstack([X.mean(0), Y.mean(0)]) - Zc) / np.linalg.norm(Zc)
    assert err_net < err_mean
    assert np.all(np.isfinite(pwh)) and np.all(pwh > 0)
    warm, winfo = solve_warm(lib, test_inst, params, xh, yh, pwh)
    assert all(w.status == "OPTIMAL" for w in warm)
    for c, w in zip(cold, warm):  # two points that pass the 1e-6 KKT test from different starts
        assert abs(w.primal_objective - c.primal_objective) <= OBJ_AGREEMENT * (1 + abs(c.primal_objective))
    print(f"\nprediction error net {err_net:.3f} / mean {err_mean:.3f}; iterations cold "
          f"{sum(c.iterations for c in cold)} warm {sum(w.iterations for w in warm)} "
          f"(fallbacks {winfo['fallback']})")


@pytest.mark.gpu
def test_learned_warm_start_on_planted_family(ws, monkeypatch):
    """The planted fixed-A family (generators.planted_family) at m=1000: 256
    training instances solved in one batch to 1e-8 (accurate snapshots), the
    head + network fitted on their solutions and final primal weights, 32
    held-out instances solved cold, from the training-mean start (the start
    available without any model) and from the learned start.

    The learned start must cut the cold iterations by 20x and beat the mean
    start; every warm answer passes the solver's FP64 KKT test at 1e-6 and
    meets the planted optimum.  The network's own share over the affine head
    is printed: on a family whose optimal basis does not move the solution
    is affine in theta and the head already fits it."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    # lane counts pinned: two batches compared bit for bit must not autotune apart
    monkeypatch.setenv("PDAL_BATCH_G_A", "16")
    monkeypatch.setenv("PDAL_BATCH_G_AT", "8")
    from pdal_amd.generators import config4_thetas, planted_family
    from pdal_amd.lp import LpBatch, load_cp_al
    from pdal_amd.warmstart import LearnedWarmStart, solve_warm
    lib = load_cp_al()
    fam = planted_family(1000, 2000, 8, seed=4)
    tr, te = config4_thetas(256, 1), config4_thetas(32, 2)
    p_train = lib.default_params(termination_evaluation_frequency=64, eps_optimal_relative=1e-8,
                                 eps_feasible_relative=1e-8, iteration_limit=60000)
    params = lib.default_params(termination_evaluation_frequency=64, eps_optimal_relative=1e-6,
                                eps_feasible_relative=1e-6, iteration_limit=200000)
    print("[case] planted family m=1000: training solves", flush=True)
    with LpBatch(lib, [fam.instance(t) for t in tr], p_train) as b:
        b.solve()
        sols = [b.result(k) for k in range(len(tr))]
        pw = b.primal_weights()
    X, Y = np.stack([s.x for s in sols]), np.stack([s.y for s in sols])
    test_inst = [fam.instance(t) for t in te]

    def batch_its(xh=None, yh=None, pwh=None):
        with LpBatch(lib, test_inst, params) as b:
            if xh is not None:
                b.set_starts(np.ascontiguousarray(xh).ravel(), np.ascontiguousarray(yh).ravel())
            if pwh is not None:
                b.set_primal_weights(pwh)
            assert b.solve() == len(te)
            return [b.result(k) for k in range(len(te))]

    cold = batch_its()
    n = len(te)
    mean_start = batch_its(np.tile(X.mean(0), (n, 1)), np.tile(Y.mean(0), (n, 1)), np.full(n, np.exp(np.log(pw).mean())))
    lws = LearnedWarmStart(ws, kx=128, seed=3)
    try:
        info = lws.fit(tr, X, Y, pw=pw, lb=np.zeros(fam.n), epochs=150)
        xh, yh, pwh = lws.predict(te)
        xa, ya, _ = lws.predict(te, mlp=False)
        # the same starts expanded on the device straight into the batch's start buffers
        with LpBatch(lib, test_inst, params) as b:
            pw_dev = lws.predict_into(b, te)
            b.solve()
            dev = [b.result(k) for k in range(len(te))]
    finally:
        lws.close()
    np.testing.assert_array_equal(pw_dev, pwh)
    warm, winfo = solve_warm(lib, test_inst, params, xh, yh, pwh)
    for d, w in zip(dev, warm):
        assert d.iterations == w.iterations and np.array_equal(d.x, w.x) and np.array_equal(d.y, w.y)
    assert winfo["fallback"] == 0 and all(w.status == "OPTIMAL" for w in warm)
    for w, t in zip(warm, te):
        opt = fam.optimal_value(t)
        assert abs(w.primal_objective - opt) <= 1e-6 * (1 + abs(opt))
    its = {k: sum(r.iterations for r in v) for k, v in (("cold", cold), ("mean", mean_start), ("learned", warm))}
    assert its["learned"] * 20 <= its["cold"]
    assert its["learned"] < its["mean"]
    Xp = np.stack([fam.solution(t)[0] for t in te])
    err = lambda a: np.linalg.norm(a - Xp) / np.linalg.norm(Xp)  # noqa: E731
    print(f"\niterations {its}; head residual {info['head_rel_resid']:.2e}; x error vs planted: head {err(xa):.2e}, "
          f"head + network {err(xh):.2e}, mean {err(X.mean(0)):.2e}")


@pytest.mark.gpu
def test_network_beats_affine_head_on_moving_active_set(ws, monkeypatch):
    """On a family whose optimal active set moves with theta
    (generators.active_set_family: x = max(0, z(theta)), a kink per column)
    the affine head cannot follow the solution map.  The network on top of it
    must cut both the held-out start error against the planted optimum and the
    held-out iteration count; every warm answer meets the planted optimum."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    monkeypatch.setenv("PDAL_BATCH_G_A", "16")
    monkeypatch.setenv("PDAL_BATCH_G_AT", "8")
    from pdal_amd.generators import active_set_family, kkt_objective_bar
    from pdal_amd.lp import LpBatch, load_cp_al
    from pdal_amd.warmstart import LearnedWarmStart
    lib = load_cp_al()
    fam = active_set_family(1000, 2000, 8, seed=4)
    tr = np.random.default_rng(1).standard_normal((512, 2 * fam.p))
    te = np.random.default_rng(2).standard_normal((64, 2 * fam.p))
    p_train = lib.default_params(termination_evaluation_frequency=64, eps_optimal_relative=1e-8,
                                 eps_feasible_relative=1e-8, iteration_limit=60000)
    params = lib.default_params(termination_evaluation_frequency=64, eps_optimal_relative=1e-6,
                                eps_feasible_relative=1e-6, iteration_limit=200000)
    print("[case] active-set family m=1000: training solves", flush=True)
    with LpBatch(lib, [fam.instance(t) for t in tr], p_train) as b:
        b.solve()
        sols = [b.result(k) for k in range(len(tr))]
        pw = b.primal_weights()
    X, Y = np.stack([s.x for s in sols]), np.stack([s.y for s in sols])
    test_inst = [fam.instance(t) for t in te]
    lws = LearnedWarmStart(ws, widths=(2 * fam.p, 512, 512, 512, 256), kx=128, seed=3)
    try:
        info = lws.fit(tr, X, Y, pw=pw, lb=np.zeros(fam.n), epochs=300)
        starts = {"head": lws.predict(te, mlp=False), "head+network": lws.predict(te)}
    finally:
        lws.close()
    print(f"\nfit {info}", flush=True)
    Xp = np.stack([fam.solution(t)[0] for t in te])
    err = {k: float(np.linalg.norm(v[0] - Xp) / np.linalg.norm(Xp)) for k, v in starts.items()}
    print(f"\nx error vs planted {err}; predicted primal weights "
          f"{ {k: (float(np.min(v[2])), float(np.max(v[2]))) for k, v in starts.items()} }", flush=True)
    its = {}
    with LpBatch(lib, test_inst, params) as b:
        for k, (xh, yh, pwh) in starts.items():
            assert np.all(np.isfinite(xh)) and np.all(np.isfinite(yh)) and np.all(np.isfinite(pwh)), k
            b.set_starts(np.ascontiguousarray(xh).ravel(), np.ascontiguousarray(yh).ravel())
            b.set_primal_weights(pwh)
            n_opt = b.solve()
            res = [b.result(i) for i in range(len(te))]
            assert n_opt == len(te), (k, sorted({r.status for r in res}), [r.iterations for r in res][:8])
            for r, t in zip(res, te):  # within the bar the KKT test at 1e-6 implies
                opt = fam.optimal_value(t)
                xs, ys, _ = fam.solution(t)
                inst_t = fam.instance(t)
                bar = kkt_objective_bar(r.primal_objective, r.dual_objective, xs, ys, inst_t["con_lb"], inst_t["c"],
                                        1e-6)
                assert abs(r.primal_objective - opt) <= bar
            its[k] = sum(r.iterations for r in res)
        b.set_vectors()  # the same vectors: zero starts, default primal weights
        b.solve()
        its["cold"] = sum(b.result(i).iterations for i in range(len(te)))
    print(f"\nx error vs planted {err}; iterations {its}")
    assert err["head+network"] < 0.75 * err["head"]
    assert its["head+network"] < its["head"] < its["cold"]


@pytest.mark.gpu
def test_batch_primal_weights_roundtrip_and_exact_restart(monkeypatch):
    """set_primal_weights is read back before a solve; a batch restarted from
    its own solutions with the primal weights it ended with converges within a
    few evaluation blocks, far below the cold count."""
    monkeypatch.setenv("CP_AL_LAMBDA", "0")
    from pdal_amd.generators import config4_thetas, lp_family
    from pdal_amd.lp import LpBatch, load_cp_al
    lib = load_cp_al()
    fam = lp_family(300, 600, 8, seed=21)
    inst = [fam.instance(t) for t in config4_thetas(8, 5)]
    params = lib.default_params(termination_evaluation_frequency=64, eps_optimal_relative=1e-6,
                                eps_feasible_relative=1e-6, iteration_limit=200000)
    with LpBatch(lib, inst, params) as b:
        assert b.solve() == len(inst)
        cold = [b.result(k) for k in range(len(inst))]
        pw = b.primal_weights()
    with LpBatch(lib, inst, params) as b:
        b.set_primal_weights(pw)
        np.testing.assert_array_equal(b.primal_weights(), pw)
        b.set_starts(np.concatenate([c.x for c in cold]), np.concatenate([c.y for c in cold]))
        assert b.solve() == len(inst)
        warm = [b.result(k) for k in range(len(inst))]
    assert sum(w.iterations for w in warm) < 0.2 * sum(c.iterations for c in cold)
    for c, w in zip(cold, warm):  # two points that pass the 1e-6 KKT test from different starts
        assert abs(w.primal_objective - c.primal_objective) <= OBJ_AGREEMENT * (1 + abs(c.primal_objective))


def test_network_rejects_rows_of_the_wrong_width():
    """Inputs or targets whose rows are not the network's widths would be read
    past their ends by the library; the wrapper refuses them first."""
    from pdal_amd.warmstart import Mlp
    net = object.__new__(Mlp)
    net.widths = [8, 16, 5]
    net._check(np.zeros((3, 8), dtype=np.float32), np.zeros((3, 5), dtype=np.float32))
    with pytest.raises(ValueError):
        net._check(np.zeros((3, 64), dtype=np.float32))
    with pytest.raises(ValueError):
        net._check(np.zeros((3, 8), dtype=np.float32), np.zeros((3, 4), dtype=np.float32))
    with pytest.raises(ValueError):
        net._check(np.zeros(8, dtype=np.float32))
