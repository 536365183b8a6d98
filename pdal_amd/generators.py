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


@dataclass
class PlantedQP:
    """min 1/2 x^T (Q + R^T R) x + c^T x  s.t. l <= A x <= u, lb <= x <= ub with
    a planted KKT point in the solver's sign convention (dual residual
    Qx + c - A^T y - lam = 0, y >= 0 on lower-active rows, y <= 0 on
    upper-active rows, lam >= 0 at x = lb, lam <= 0 at x = ub)."""
    A: sp.csr_matrix
    M: sp.csr_matrix          # low-rank factor: Q_total = M^T M + reg I
    reg: float
    c: np.ndarray
    l: np.ndarray
    u: np.ndarray
    lb: np.ndarray
    ub: np.ndarray
    x_star: np.ndarray
    y_star: np.ndarray
    lam_star: np.ndarray
    seed: int

    @property
    def n(self) -> int:
        return self.A.shape[1]

    @property
    def m(self) -> int:
        return self.A.shape[0]

    def q_explicit(self) -> sp.csr_matrix:
        Q = (self.M.T @ self.M + self.reg * sp.identity(self.n, format="csr")).tocsr()
        Q.sort_indices()
        return Q

    def q_diag_part(self) -> sp.csr_matrix:
        return (self.reg * sp.identity(self.n, format="csr")).tocsr()

    @property
    def optimal_value(self) -> float:
        x = self.x_star
        return float(0.5 * (np.sum((self.M @ x) ** 2) + self.reg * x @ x) + self.c @ x)

    def qp_args(self, form: str = "explicit", equality: bool = False) -> dict:
        """Keyword arguments of FaCp.solve: "explicit" passes Q = M^T M + reg I;
        "lowrank" passes Q = reg I and R = M.  equality=True gives the equality-
        constrained problem A x = A x* (the same planted KKT point: y is free)."""
        l, u = (self.A @ self.x_star,) * 2 if equality else (self.l, self.u)
        base = dict(c=self.c, A=self.A, con_lb=l, con_ub=u, var_lb=self.lb, var_ub=self.ub)
        if form == "explicit":
            return dict(base, Q=self.q_explicit())
        return dict(base, Q=self.q_diag_part(), R=self.M)


def planted_qp(n: int, m: int, r: int, seed: int, a_nnz_col: int = 8, m_nnz_col: int = 5, reg: float = 1e-3,
               box: float = 10.0, bound_frac: float = 0.3) -> PlantedQP:
    """Planted convex QP of the config-2 family (n=1e5, m=5e4, r=2e4 at full size)."""
    rng = np.random.default_rng(seed)

    def sparse_cols(rows, cols, k):
        counts = np.full(cols, k, dtype=np.int64)
        indptr = np.zeros(cols + 1, dtype=np.int64)
        np.cumsum(counts, out=indptr[1:])
        cidx = np.repeat(np.arange(cols, dtype=np.int64), counts)
        ridx = _distinct_rows(rng, cidx, rows)
        vals = rng.standard_normal(cidx.size)
        S = sp.csc_matrix((vals, ridx, indptr), shape=(rows, cols)).tocsr()
        S.sort_indices()
        return S

    A = sparse_cols(m, n, a_nnz_col)
    M = sparse_cols(r, n, m_nnz_col)
    x = rng.uniform(0.0, box, n)
    perm = rng.permutation(n)
    nb = int(n * bound_frac)
    at_lo, at_hi = perm[: nb // 2], perm[nb // 2: nb]
    x[at_lo] = 0.0
    x[at_hi] = box
    lam = np.zeros(n)
    lam[at_lo] = rng.uniform(0.1, 1.0, at_lo.size)
    lam[at_hi] = -rng.uniform(0.1, 1.0, at_hi.size)
    ax = A @ x
    kind = rng.integers(0, 3, m)  # 0: lower-active (y>0), 1: upper-active (y<0), 2: slack (y=0)
    y = np.zeros(m)
    y[kind == 0] = rng.uniform(0.1, 1.0, int(np.sum(kind == 0)))
    y[kind == 1] = -rng.uniform(0.1, 1.0, int(np.sum(kind == 1)))
    width = rng.uniform(0.5, 2.0, m)
    l = np.where(kind == 0, ax, ax - width)
    u = np.where(kind == 1, ax, ax + width)
    qx = M.T @ (M @ x) + reg * x
    c = -qx + A.T @ y + lam
    return PlantedQP(A=A, M=M, reg=reg, c=c, l=l, u=u, lb=np.zeros(n), ub=np.full(n, box), x_star=x, y_star=y,
                     lam_star=lam, seed=seed)


def config2(seed: int = 2, scale: float = 1.0) -> PlantedQP:
    """n=1e5, m=5e4, M 2e4 x n (5 nnz/col), A 8 nnz/col; `scale` shrinks the sizes."""
    n = int(round(100_000 * scale))
    return planted_qp(n, n // 2, max(n // 5, 1), seed)


@dataclass
class LpFamily:
    """Fixed-A LP family: b = b0 + U theta_b, c = c0 + V theta_c (x >= 0, A x = b),
    theta = [theta_b, theta_c] of length 2p; base (A, b0, c0) planted as planted_lp."""
    base: PlantedLP
    U: np.ndarray   # m x p
    V: np.ndarray   # n x p

    @property
    def m(self) -> int:
        return self.base.m

    @property
    def n(self) -> int:
        return self.base.n

    @property
    def p(self) -> int:
        return self.U.shape[1]

    def instance(self, theta) -> dict:
        """Keyword arguments of CpAl.make_problem / LpBatch for parameters theta."""
        theta = np.asarray(theta, dtype=np.float64)
        b = self.base.b + self.U @ theta[: self.p]
        c = self.base.c + self.V @ theta[self.p:]
        n = self.n
        return dict(A=self.base.A, c=c, con_lb=b, con_ub=b, var_lb=np.zeros(n), var_ub=np.full(n, np.inf))


def lp_family(m: int, n: int, nnz_per_col: int, seed: int, p: int = 32, scale: float = 0.1) -> LpFamily:
    base = planted_lp(m, n, nnz_per_col, seed)
    rng = np.random.default_rng(seed + 1_000_003)
    return LpFamily(base=base, U=scale * rng.standard_normal((m, p)), V=scale * rng.standard_normal((n, p)))


@dataclass
class PlantedFamily:
    """Fixed-A LP family with a planted optimum for every parameter vector.

    b(theta) = b0 + U theta_b and c(theta) = c0 + V theta_c as in LpFamily, with
    U and V dense but built from the planted optimum of the base LP:

      x(theta) = x* . (1 + W theta_b)        U = A diag(x*) W
      y(theta) = y* + G theta_c              V = A^T G + diag(s*) H
      s(theta) = s* . (1 + H theta_c)

    W (n x p), G (m x p), H (n x p) have N(0, scale^2 / p) entries, so x, y and
    s move by `scale` relative per entry.  x(theta) >= 0, s(theta) >= 0 keep
    their supports (disjoint, so s^T x = 0), A x(theta) = b(theta) and
    c(theta) - A^T y(theta) = s(theta): (x, y, s)(theta) is optimal for every
    theta with 1 + W theta_b > 0 and 1 + H theta_c > 0 (a 10-sigma event to
    fail at scale 0.1; `instance` checks it).  Independent Gaussian U, V cannot
    be used at this size: rows of A whose few entries share one sign force a
    sign on b_i, and an additive shift of b flips it (infeasible LPs).
    """
    base: PlantedLP
    W: np.ndarray   # n x p
    G: np.ndarray   # m x p
    H: np.ndarray   # n x p
    U: np.ndarray   # m x p
    V: np.ndarray   # n x p

    @property
    def m(self) -> int:
        return self.base.m

    @property
    def n(self) -> int:
        return self.base.n

    @property
    def p(self) -> int:
        return self.U.shape[1]

    def solution(self, theta):
        """The planted (x, y, s) of parameters theta."""
        theta = np.asarray(theta, dtype=np.float64)
        tb, tc = theta[: self.p], theta[self.p:]
        fx, fs = 1.0 + self.W @ tb, 1.0 + self.H @ tc
        if fx.min() <= 0.0 or fs.min() <= 0.0:
            raise ValueError("theta outside the family's planted region")
        return self.base.x_star * fx, self.base.y_star + self.G @ tc, self.base.s_star * fs

    def optimal_value(self, theta) -> float:
        x, _, _ = self.solution(theta)
        return float(self.instance(theta)["c"] @ x)

    def instance(self, theta) -> dict:
        """Keyword arguments of CpAl.make_problem / LpBatch for parameters theta."""
        theta = np.asarray(theta, dtype=np.float64)
        self.solution(theta)  # the region check
        b = self.base.b + self.U @ theta[: self.p]
        c = self.base.c + self.V @ theta[self.p:]
        n = self.n
        return dict(A=self.base.A, c=c, con_lb=b, con_ub=b, var_lb=np.zeros(n), var_ub=np.full(n, np.inf))


def planted_family(m: int, n: int, nnz_per_col: int, seed: int, p: int = 32, scale: float = 0.1) -> PlantedFamily:
    base = planted_lp(m, n, nnz_per_col, seed)
    rng = np.random.default_rng(seed + 2_000_003)
    sd = scale / np.sqrt(p)
    W = sd * rng.standard_normal((n, p))
    G = sd * rng.standard_normal((m, p))
    H = sd * rng.standard_normal((n, p))
    U = base.A @ (base.x_star[:, None] * W)
    V = base.A.T @ G + base.s_star[:, None] * H
    return PlantedFamily(base=base, W=W, G=G, H=H, U=np.ascontiguousarray(U), V=np.ascontiguousarray(V))


@dataclass
class ActiveSetFamily:
    """Fixed-A LP family whose optimal active set moves with the parameters.

      z(theta) = (x* - s*) + W theta_b
      x(theta) = max(z, 0),  s(theta) = max(-z, 0),  y(theta) = y* + G theta_c
      b(theta) = A x(theta),  c(theta) = A^T y(theta) + s(theta)

    x and s are nonnegative with disjoint supports for every theta, A x = b and
    c - A^T y = s: (x, y, s)(theta) is a KKT point, so x(theta) is optimal and
    c^T x(theta) the optimal value, with no region to leave.  An entry of z
    that crosses zero moves its column between the primal support and the
    dual slacks: b and c are piecewise linear in theta, the solution map has a
    kink per column, and an affine predictor of the solution cannot follow it
    (a ReLU network can).  W (n x p), G (m x p) have N(0, scale^2 / p)
    entries; the supports of x* and s* have U(0, 1) entries, so about a tenth
    of the columns change sides at scale 0.3.  When the support of x(theta)
    is larger than m the optimal x need not be unique; the optimal value is.
    """
    base: PlantedLP
    W: np.ndarray   # n x p
    G: np.ndarray   # m x p

    @property
    def m(self) -> int:
        return self.base.m

    @property
    def n(self) -> int:
        return self.base.n

    @property
    def p(self) -> int:
        return self.W.shape[1]

    def solution(self, theta):
        """The planted (x, y, s) of parameters theta."""
        theta = np.asarray(theta, dtype=np.float64)
        z = self.base.x_star - self.base.s_star + self.W @ theta[: self.p]
        return np.maximum(z, 0.0), self.base.y_star + self.G @ theta[self.p:], np.maximum(-z, 0.0)

    def optimal_value(self, theta) -> float:
        x, y, _ = self.solution(theta)
        return float(self.instance(theta)["c"] @ x)

    def instance(self, theta) -> dict:
        """Keyword arguments of CpAl.make_problem / LpBatch for parameters theta."""
        x, y, s = self.solution(theta)
        A = self.base.A
        b = A @ x
        c = A.T @ y + s
        n = self.n
        return dict(A=A, c=c, con_lb=b, con_ub=b, var_lb=np.zeros(n), var_ub=np.full(n, np.inf))


def active_set_family(m: int, n: int, nnz_per_col: int, seed: int, p: int = 4,
                      scale: float = 0.3) -> ActiveSetFamily:
    base = planted_lp(m, n, nnz_per_col, seed)
    rng = np.random.default_rng(seed + 3_000_003)
    sd = scale / np.sqrt(p)
    return ActiveSetFamily(base=base, W=sd * rng.standard_normal((n, p)), G=sd * rng.standard_normal((m, p)))


def config4_family(seed: int = 4, scale: float = 1.0) -> PlantedFamily:
    """m=5e4, n=1e5, 8 nonzeros per column; b = b0 + U theta_b, c = c0 + V theta_c
    with dense U (m x 32), V (n x 32) around a planted optimum, 0.1 relative."""
    m = int(round(50_000 * scale))
    return planted_family(m, 2 * m, 8, seed)


def config4_thetas(count: int, seed: int, p: int = 32) -> np.ndarray:
    """theta ~ N(0, I_2p), one row per instance."""
    return np.random.default_rng(seed).standard_normal((count, 2 * p))


def kkt_objective_bar(primal_obj: float, dual_obj: float, x_star, y_star, b, c, tol: float) -> float:
    """How far a solve that passes the relative KKT test at `tol` (L2 norms,
    the reference's definitions: |p - d| < tol (1 + |p| + |d|), |Ax - b| < tol
    (1 + |b|), |c - A^T y - lambda| < tol (1 + |c|)) may put its primal
    objective from the optimum of an LP with the optimal pair (x*, y*): the gap
    allowance plus what the feasibility residuals can move an objective,
    |y*| tol (1 + |b|) on the primal side and |x*| tol (1 + |c|) on the dual
    side (weak duality through the optimal pair)."""
    gap = tol * (1.0 + abs(primal_obj) + abs(dual_obj))
    side = tol * max(float(np.linalg.norm(y_star)) * (1.0 + float(np.linalg.norm(b))),
                     float(np.linalg.norm(x_star)) * (1.0 + float(np.linalg.norm(c))))
    return gap + side
