"""The twelve matrix features of the LP runner's policy layer
(compute_feature_matrix_stats, LP/CP_AL/src/cli.c:91-214), recomputed here in
numpy from their definitions and compared with what the runner prints in its
[feature-auto-policy] audit line (CP_AL_FEATURE_AUDIT_ONLY=1 returns before any
GPU work, cli.c:1057-1062, so this is a CPU test).

Definitions, over the CSR matrix as read from the MPS file: entries count
toward the row counts as stored and toward the column counts and the
magnitude statistics only when finite and nonzero; a row or column is
"nonempty" when it has such an entry; standard deviations and coefficients of
variation are population ones (divide by the count)."""
from __future__ import annotations

import re

import numpy as np
import pytest
import scipy.sparse as sp

from test_lp_cli import CLI, _run, write_mps

FEATURES = ["coeff_log_range", "row_logmax_std", "col_logmax_std", "row_logspan_mean", "col_logspan_mean",
            "row_span2_frac", "col_span2_frac", "row_nnz_cv", "col_nnz_cv", "equality_frac",
            "two_sided_var_frac", "objective_density"]


def numpy_features(A: sp.csr_matrix, c, l, u, lb, ub) -> dict:
    A = sp.csr_matrix(A)
    m, n = A.shape
    mag = np.abs(A.data)
    ok = np.isfinite(mag) & (mag > 0)
    rows = np.repeat(np.arange(m), np.diff(A.indptr))
    cols = A.indices
    rmax = np.zeros(m)
    rmin = np.full(m, np.inf)
    np.maximum.at(rmax, rows[ok], mag[ok])
    np.minimum.at(rmin, rows[ok], mag[ok])
    cmax = np.zeros(n)
    cmin = np.full(n, np.inf)
    np.maximum.at(cmax, cols[ok], mag[ok])
    np.minimum.at(cmin, cols[ok], mag[ok])
    rne, cne = rmax > 0, cmax > 0
    rspan = np.log10(rmax[rne] / rmin[rne])
    cspan = np.log10(cmax[cne] / cmin[cne])

    def cv(counts):
        counts = np.asarray(counts, dtype=np.float64)
        return 0.0 if counts.sum() <= 0 else float(np.std(counts) / np.mean(counts))

    return {
        "coeff_log_range": float(np.log10(mag[ok].max() / mag[ok].min())),
        "row_logmax_std": float(np.std(np.log10(rmax[rne]))),
        "col_logmax_std": float(np.std(np.log10(cmax[cne]))),
        "row_logspan_mean": float(rspan.mean()),
        "col_logspan_mean": float(cspan.mean()),
        "row_span2_frac": float(np.mean(rspan >= 2.0)),
        "col_span2_frac": float(np.mean(cspan >= 2.0)),
        "row_nnz_cv": cv(np.diff(A.indptr)),
        "col_nnz_cv": cv(np.bincount(cols[ok], minlength=n)),
        "equality_frac": float(np.mean(np.isfinite(l) & np.isfinite(u) & (l == u))),
        "two_sided_var_frac": float(np.mean(np.isfinite(lb) & np.isfinite(ub))),
        "objective_density": float(np.mean(np.abs(c) > 0)),
    }


def _instance(seed: int):
    rng = np.random.default_rng(seed)
    m, n = 70 + 10 * seed, 90 + 7 * seed
    A = sp.random(m, n, density=0.06 + 0.02 * seed, random_state=seed, format="lil")
    A[3, :] = 0.0   # an empty row
    A[:, 5] = 0.0   # an empty column
    A = sp.csr_matrix(A)
    A.eliminate_zeros()
    # magnitudes over several decades, both signs, so the span statistics are not trivial
    A.data = np.sign(rng.standard_normal(A.nnz)) * 10.0 ** rng.uniform(-3, 3 + seed, A.nnz)
    c = rng.standard_normal(n) * (rng.random(n) < 0.6)
    ax = A @ rng.random(n)
    kind = rng.integers(0, 4, m)
    l = np.where(kind == 0, ax, np.where(kind == 1, ax - 1.0, np.where(kind == 2, ax - 1.0, -np.inf)))
    u = np.where(kind == 0, ax, np.where(kind == 1, ax + 1.0, np.where(kind == 2, np.inf, ax + 1.0)))
    vk = rng.integers(0, 4, n)
    lb = np.where(vk == 0, -np.inf, np.where(vk == 1, 0.0, -1.0))
    ub = np.where(vk == 2, 2.0, np.where(vk == 3, 5.0, np.inf))
    lb[vk == 0] = -np.inf
    return A, c, l, u, lb, ub


@pytest.mark.skipif(not CLI.exists(), reason="LP runner not built")
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_printed_features_equal_numpy_definitions(tmp_path, seed):
    A, c, l, u, lb, ub = _instance(seed)
    write_mps(tmp_path / "f.mps", A, c, l, u, lb, ub)
    r = _run([str(tmp_path / "f.mps"), str(tmp_path)],
             {"CP_AL_FEATURE_AUTO_POLICY": "1", "CP_AL_FEATURE_AUDIT_ONLY": "1"})
    assert r.returncode == 0, r.stderr
    line = next(ln for ln in r.stderr.splitlines() if ln.startswith("[feature-auto-policy] version=2"))
    printed = {k: float(re.search(rf"\b{k}=(\S+)", line).group(1)) for k in FEATURES}
    want = numpy_features(A, c, l, u, lb, ub)
    for k in FEATURES:
        assert printed[k] == pytest.approx(want[k], rel=2e-6, abs=1e-9), (k, printed[k], want[k])
    # the instance exercises every statistic away from its trivial value
    assert 0 < want["row_span2_frac"] < 1 and 0 < want["col_span2_frac"] < 1
    assert 0 < want["equality_frac"] < 1 and 0 < want["two_sided_var_frac"] < 1 and 0 < want["objective_density"] < 1
