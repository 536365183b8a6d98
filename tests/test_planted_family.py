"""The fixed-A LP family with a planted optimum (pdal_amd.generators
.planted_family, the warm-start workload at m=5e4 / n=1e5): b and c are
affine in theta through fixed dense U, V, and (x, y, s)(theta) is a KKT point
for every theta.  Pinned here by the KKT conditions in numpy and by HiGHS's
optimal value, and the affine head of the warm start by exact recovery of an
affine map (a stub stands in for the network, which runs on the GPU)."""
from __future__ import annotations

import numpy as np
import pytest
from scipy.optimize import linprog

from pdal_amd.generators import config4_thetas, planted_family
from pdal_amd.warmstart import LearnedWarmStart


@pytest.fixture(scope="module")
def fam():
    return planted_family(200, 400, 8, seed=9)


def test_planted_point_is_kkt_for_every_theta(fam):
    for theta in config4_thetas(5, 3):
        d = fam.instance(theta)
        x, y, s = fam.solution(theta)
        A, b, c = d["A"], d["con_lb"], d["c"]
        assert np.linalg.norm(A @ x - b) <= 1e-12 * (1 + np.linalg.norm(b))
        assert np.all(x >= 0) and np.all(s >= 0)
        np.testing.assert_allclose(c - A.T @ y, s, atol=1e-12 * (1 + np.abs(c).max()))
        assert np.all((x == 0) | (s == 0))  # complementary supports
        # the affine form of the data
        p = fam.p
        np.testing.assert_allclose(b, fam.base.b + fam.U @ theta[:p], rtol=0, atol=1e-12 * np.abs(b).max())
        np.testing.assert_allclose(c, fam.base.c + fam.V @ theta[p:], rtol=0, atol=1e-12 * np.abs(c).max())
        # x and y move by about 10 % of the base point
        assert 0.02 < np.linalg.norm(x - fam.base.x_star) / np.linalg.norm(fam.base.x_star) < 0.3


@pytest.mark.parametrize("k", [0, 1, 2])
def test_planted_value_is_highs_optimum(fam, k):
    theta = config4_thetas(3, 4)[k]
    d = fam.instance(theta)
    r = linprog(d["c"], A_eq=d["A"], b_eq=d["con_lb"], bounds=(0, None), method="highs")
    assert r.status == 0
    opt = fam.optimal_value(theta)
    assert abs(r.fun - opt) <= 1e-9 * (1 + abs(opt))


def test_region_check_rejects_far_theta(fam):
    with pytest.raises(ValueError):
        fam.instance(np.full(2 * fam.p, 1e3))


class _StubMlp:
    def __init__(self, out):
        self.out = out

    def train(self, X, Y, **kw):
        return np.array([1.0, 0.5], dtype=np.float32)

    def forward(self, X):
        return np.zeros((X.shape[0], self.out), dtype=np.float32)

    def close(self):
        pass


class _StubLib:
    """Network that predicts zero residual; the POD expansion in numpy."""

    def create(self, widths, seed=0):
        return _StubMlp(widths[-1])

    @staticmethod
    def pod_expand(basis, mean, coef):
        return mean.astype(np.float64) + coef.astype(np.float64) @ basis.astype(np.float64)


def test_affine_head_recovers_affine_solutions():
    """Snapshots affine in theta (as a family with a fixed optimal basis has):
    the head alone predicts held-out snapshots to the FP32 of the expansion,
    the primal weight's logarithm too; without the head the stub network's
    zero output gives the training mean."""
    rng = np.random.default_rng(5)
    p, n, m, N = 4, 60, 30, 40
    Mx, My = rng.standard_normal((p, n)), rng.standard_normal((p, m))
    x0, y0 = rng.random(n) + 1.0, rng.standard_normal(m)
    th = rng.standard_normal((N, 2 * p))
    te = rng.standard_normal((6, 2 * p))
    X, Y = x0 + 0.1 * th[:, :p] @ Mx, y0 + 0.1 * th[:, p:] @ My
    pw = np.exp(0.3 * th[:, 0] - 0.2 * th[:, p])
    lws = LearnedWarmStart(_StubLib(), widths=(2 * p, 16, 24), kx=10, head=True)
    info = lws.fit(th, X, Y, pw=pw, lb=np.zeros(n))
    assert info["head_rel_resid"] <= 1e-12
    xh, yh, pwh = lws.predict(te)
    np.testing.assert_allclose(xh, x0 + 0.1 * te[:, :p] @ Mx, rtol=1e-6)
    np.testing.assert_allclose(yh, y0 + 0.1 * te[:, p:] @ My, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(pwh, np.exp(0.3 * te[:, 0] - 0.2 * te[:, p]), rtol=1e-10)
    plain = LearnedWarmStart(_StubLib(), widths=(2 * p, 16, 24), kx=10, head=False)
    plain.fit(th, X, Y, pw=pw)
    xm, _, _ = plain.predict(te)
    np.testing.assert_allclose(xm, np.broadcast_to(X.mean(0), xm.shape), rtol=1e-6)
    with pytest.raises(ValueError):
        plain.predict(te, mlp=False)


# ---------------------------------------------------------------------------
# The family whose active set moves with theta (generators.active_set_family)

@pytest.fixture(scope="module")
def afam():
    from pdal_amd.generators import active_set_family
    return active_set_family(200, 400, 8, seed=9)


def test_active_set_point_is_kkt_for_every_theta(afam):
    rng = np.random.default_rng(3)
    for theta in rng.standard_normal((5, 2 * afam.p)):
        d = afam.instance(theta)
        x, y, s = afam.solution(theta)
        A, b, c = d["A"], d["con_lb"], d["c"]
        assert np.linalg.norm(A @ x - b) <= 1e-12 * (1 + np.linalg.norm(b))
        assert np.all(x >= 0) and np.all(s >= 0)
        np.testing.assert_allclose(c - A.T @ y, s, atol=1e-12 * (1 + np.abs(c).max()))
        assert np.all((x == 0) | (s == 0))


@pytest.mark.parametrize("k", [0, 1])
def test_active_set_value_is_highs_optimum(afam, k):
    theta = np.random.default_rng(4).standard_normal((2, 2 * afam.p))[k]
    d = afam.instance(theta)
    r = linprog(d["c"], A_eq=d["A"], b_eq=d["con_lb"], bounds=(0, None), method="highs")
    assert r.status == 0
    opt = afam.optimal_value(theta)
    assert abs(r.fun - opt) <= 1e-9 * (1 + abs(opt))


def test_active_set_moves_and_data_is_not_affine(afam):
    """About a tenth of the columns change sides, and a least-squares affine
    fit of x(theta) over many parameters leaves a large residual, where the
    affine family's would be zero."""
    th = np.random.default_rng(5).standard_normal((400, 2 * afam.p))
    X = np.stack([afam.solution(t)[0] for t in th])
    moved = np.mean((X > 0) != (afam.base.x_star > 0))
    assert 0.05 < moved < 0.2
    D = np.hstack([th, np.ones((len(th), 1))])
    coef = np.linalg.lstsq(D, X, rcond=None)[0]
    rel = np.linalg.norm(X - D @ coef) / np.linalg.norm(X - X.mean(0))
    assert rel > 0.2
    # the supports of x and s change, the dual y does not leave its affine form
    Y = np.stack([afam.solution(t)[1] for t in th])
    cy = np.linalg.lstsq(D, Y, rcond=None)[0]
    assert np.linalg.norm(Y - D @ cy) <= 1e-10 * np.linalg.norm(Y)
