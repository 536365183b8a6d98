"""Seeded planted-optimum LP instance generators.

Every instance is min c^T x s.t. A x = b, x >= 0 with a known optimal
primal-dual pair built in:

  * A (m x n) has k_j nonzeros in column j at distinct uniform-random rows,
    values N(0,1);
  * x*: exactly floor(n * zero_frac) entries are 0, the rest U(0,1);
  * y* ~ N(0,1)^m; s_j ~ U(0,1) where x*_j = 0, else 0;
  * b = A x*, c = A^T y* + s.

(x*, y*, s) satisfies the KKT conditions (A x* = b, x* >= 0, s = c - A^T y*
>= 0, s^T x* = 0), so c^T x* = b^T y* is the optimal value.  Column counts:
a fixed k ("planted_lp") or Zipf(1.5) clipped to [2, 64] ("zipf_planted_lp").
All randomness comes from numpy's PCG64 seeded with `seed`, in a fixed order.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp


@dataclass
class PlantedLP:
    A: sp.csr_matrix
    b: np.ndarray
    c: np.ndarray
    x_star: np.ndarray
    y_star: np.ndarray
    s_star: np.ndarray
    seed: int

    @property
    def m(self) -> int:
        return self.A.shape[0]

    @property
    def n(self) -> int:
        return self.A.shape[1]

    @property
    def optimal_value(self) -> float:
        return float(self.c @ self.x_star)

    def lp_args(self) -> dict:
        """Keyword arguments of CpAl.solve for this instance."""
        n = self.n
        return dict(A=self.A, c=self.c, con_lb=self.b, con_ub=self.b, var_lb=np.zeros(n),
                    var_ub=np.full(n, np.inf))


def _distinct_rows(rng: np.random.Generator, cols: np.ndarray, m: int) -> np.ndarray:
    rows = rng.integers(0, m, size=cols.size, dtype=np.int64)
    while True:
        key = cols.astype(np.int64) * m + rows
        order = np.argsort(key, kind="stable")
        ks = key[order]
        dup = np.nonzero(ks[1:] == ks[:-1])[0]
        if dup.size == 0:
            return rows
        idx = order[dup + 1]
        rows[idx] = rng.integers(0, m, size=idx.size, dtype=np.int64)


def _planted_from_counts(rng: np.random.Generator, m: int, n: int, counts: np.ndarray, seed: int,
                         zero_frac: float) -> PlantedLP:
    counts = np.asarray(counts, dtype=np.int64)
    if counts.max(initial=0) > m:
        raise ValueError("a column has more nonzeros than rows")
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=indptr[1:])
    cols = np.repeat(np.arange(n, dtype=np.int64), counts)
    rows = _distinct_rows(rng, cols, m)
    vals = rng.standard_normal(cols.size)
    A = sp.csc_matrix((vals, rows, indptr), shape=(m, n)).tocsr()
    A.sort_indices()
    x = rng.random(n)
    zero = rng.permutation(n)[: int(n * zero_frac)]
    x[zero] = 0.0
    y = rng.standard_normal(m)
    s = np.zeros(n)
    s[zero] = rng.random(zero.size)
    b = A @ x
    c = A.T @ y + s
    return PlantedLP(A=A, b=b, c=c, x_star=x, y_star=y, s_star=s, seed=seed)


def planted_lp(m: int, n: int, nnz_per_col: int, seed: int, zero_frac: float = 0.5) -> PlantedLP:
    """Planted LP with a fixed number of nonzeros per column."""
    rng = np.random.default_rng(seed)
    return _planted_from_counts(rng, m, n, np.full(n, nnz_per_col), seed, zero_frac)


def zipf_planted_lp(m: int, n: int, seed: int, a: float = 1.5, lo: int = 2, hi: int = 64,
                    zero_frac: float = 0.5) -> PlantedLP:
    """Planted LP whose column counts are Zipf(a) clipped to [lo, hi]."""
    rng = np.random.default_rng(seed)
    counts = np.clip(rng.zipf(a, size=n), lo, hi)
    return _planted_from_counts(rng, m, n, counts, seed, zero_frac)


def config0(seed: int = 0) -> PlantedLP:
    """m=1000, n=2000, 10 nonzeros per column."""
    return planted_lp(1000, 2000, 10, seed)


def config1(seed: int = 1, scale: float = 1.0) -> PlantedLP:
    """m=2e6, n=4e6, Zipf(1.5) column counts in [2,64]; `scale` shrinks m and n."""
    m = int(round(2_000_000 * scale))
    return zipf_planted_lp(m, 2 * m, seed)


def config3_sizes(count: int = 4096, seed: int = 12345, lo: int = 200, hi: int = 20000) -> np.ndarray:
    """Row counts of the batched family: log-uniform on [lo, hi]."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=count)).astype(np.int64)


def config3_instance(i: int, m: int) -> PlantedLP:
    """Instance i of the batched family: n = 2m, 8 nonzeros per column, seed i."""
    return planted_lp(int(m), 2 * int(m), 8, seed=i)
