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
