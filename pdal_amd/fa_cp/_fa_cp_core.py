"""Core entry points of the FA_CP Python interface over libfa_cp.so.

Mirrors the reference's compiled module `fa_cp._fa_cp_core`
(QP/FA_CP/python_bindings/_core_bindings.cpp:206-600): `get_default_params()`
and `solve_once(Q, R, A, objective_vector, ...)` with the same argument
meanings, parameter keys, result keys, status strings and codes, and the same
exception types (ValueError for bad arguments, OverflowError for indices
beyond int32, RuntimeError when the C layer returns NULL).  Matrices may be
numpy arrays (dense) or scipy CSR / CSC / COO matrices and are handed to
create_qp_problem in that format; the solve itself runs in libfa_cp.so on the
GPU (ctypes releases the GIL for the call).
"""
from __future__ import annotations

import ctypes as C
from typing import Any, Optional

import numpy as np

from .. import qp as _qp

_LIB: Optional[_qp.FaCp] = None

# parameters the interface reads from a params dict (_core_bindings.cpp:302-337);
# value kind: f float, i int, b bool, n norm string
_PARAM_KEYS = {
    "verbose": ("i", None),
    "termination_evaluation_frequency": ("i", None),
    "eps_optimal_relative": ("f", "termination_criteria"),
    "eps_feasible_relative": ("f", "termination_criteria"),
    "time_sec_limit": ("f", "termination_criteria"),
    "iteration_limit": ("i", "termination_criteria"),
    "l_inf_ruiz_iterations": ("i", None),
    "has_pock_chambolle_alpha": ("b", None),
    "pock_chambolle_alpha": ("f", None),
    "bound_objective_rescaling": ("b", None),
    "artificial_restart_threshold": ("f", "restart_params"),
    "sufficient_reduction_for_restart": ("f", "restart_params"),
    "necessary_reduction_for_restart": ("f", "restart_params"),
    "k_p": ("f", "restart_params"),
    "reflection_coefficient": ("f", None),
    "feasibility_polishing": ("b", None),
    "eps_feas_polish_relative": ("f", "termination_criteria"),
    "optimality_norm": ("n", None),
    "sv_max_iter": ("i", None),
    "sv_tol": ("f", None),
}

_STATUS = {0: "UNSPECIFIED", 1: "OPTIMAL", 2: "PRIMAL_INFEASIBLE", 3: "DUAL_INFEASIBLE",
           4: "INFEASIBLE_OR_UNBOUNDED", 5: "TIME_LIMIT", 6: "ITERATION_LIMIT", 7: "FEAS_POLISH_SUCCESS"}
_CODE = {1: 0, 2: 1, 3: 2, 5: 3, 6: 4, 4: 5}


def _lib() -> _qp.FaCp:
    global _LIB
    if _LIB is None:
        _LIB = _qp.load_fa_cp()  # raises when libfa_cp.so is not built
    return _LIB


def _defaults() -> _qp.FaCpParameters:
    p = _qp.FaCpParameters()
    _lib().lib.set_default_parameters(C.byref(p))
    return p


def get_default_params() -> dict:
    """Default parameters as a dict (keys as the reference's interface)."""
    p = _defaults()
    d: dict[str, Any] = {}
    for key, (kind, group) in _PARAM_KEYS.items():
        src = getattr(p, group) if group else p
        v = getattr(src, key)
        if kind == "n":
            d[key] = "linf" if v == 1 else "l2"
        elif kind == "b":
            d[key] = bool(v)
        else:
            d[key] = v
    return d


def _parse_norm(v) -> int:
    if not isinstance(v, str):
        raise ValueError("optimality_norm must be a string ('l2'/'linf')")
    s = v.lower()
    if s == "l2":
        return 0
    if s == "linf":
        return 1
    raise ValueError(f"Unknown norm type: {v}. Use 'l2' or 'linf'.")


def _params_from_dict(params) -> _qp.FaCpParameters:
    p = _defaults()
    if params is None:
        return p
    for key, (kind, group) in _PARAM_KEYS.items():
        if key not in params:
            continue
        v = params[key]
        dst = getattr(p, group) if group else p
        if kind == "n":
            setattr(dst, key, _parse_norm(v))
        elif kind == "i":
            setattr(dst, key, int(v))
        elif kind == "b":
            setattr(dst, key, bool(v))
        else:
            setattr(dst, key, float(v))
    return p  # unknown keys are ignored, as in the reference


def _index_array(a, name: str) -> np.ndarray:
    a = np.asarray(a)
    if a.ndim != 1:
        raise ValueError(f"{name} must be 1D.")
    if a.dtype == np.int32:
        return np.ascontiguousarray(a)
    if a.dtype == np.int64:
        if a.size and (a.min() < 0 or a.max() > np.iinfo(np.int32).max):
            raise OverflowError(f"{name} has value out of int32 range; backend currently supports only "
                                f"32-bit indices.")
        return np.ascontiguousarray(a, dtype=np.int32)
    raise ValueError(f"{name} must be int32 or int64.")


def _matrix_desc(M, zero_tol: float, keep: list) -> _qp.MatrixDesc:
    if not hasattr(M, "shape"):
        raise ValueError("matrix A must be numpy.ndarray or scipy.sparse matrix (no .shape attr)")
    if len(M.shape) != 2:
        raise ValueError("matrix A must be 2D")
    d = _qp.MatrixDesc()
    d.zero_tolerance = float(zero_tol)
    d.m, d.n = int(M.shape[0]), int(M.shape[1])
    f64p = C.POINTER(C.c_double)
    i32p = C.POINTER(C.c_int)
    if isinstance(M, np.ndarray):
        D = np.ascontiguousarray(M, dtype=np.float64)
        keep.append(D)
        d.fmt = _qp.MATRIX_DENSE
        d.data.dense.A = D.ctypes.data_as(f64p)
        return d
    fmt = getattr(M, "format", "unknown")
    if fmt not in ("csr", "csc", "coo"):
        raise ValueError("Unsupported matrix A: expected numpy.ndarray or scipy.sparse (csr/csc/coo)")
    vals = np.ascontiguousarray(M.data, dtype=np.float64)
    if fmt == "coo":
        a, b = _index_array(M.row, "coo.row"), _index_array(M.col, "coo.col")
    else:
        a, b = _index_array(M.indptr, f"{fmt}.indptr"), _index_array(M.indices, f"{fmt}.indices")
    keep.extend([vals, a, b])
    if fmt == "csr":
        d.fmt = _qp.MATRIX_CSR
        d.data.csr.nnz, d.data.csr.row_ptr, d.data.csr.col_ind, d.data.csr.vals = (
            vals.size, a.ctypes.data_as(i32p), b.ctypes.data_as(i32p), vals.ctypes.data_as(f64p))
    elif fmt == "csc":
        d.fmt = _qp.MATRIX_CSC
        d.data.csc.nnz, d.data.csc.col_ptr, d.data.csc.row_ind, d.data.csc.vals = (
            vals.size, a.ctypes.data_as(i32p), b.ctypes.data_as(i32p), vals.ctypes.data_as(f64p))
    else:
        d.fmt = _qp.MATRIX_COO
        d.data.coo.nnz, d.data.coo.row_ind, d.data.coo.col_ind, d.data.coo.vals = (
            vals.size, a.ctypes.data_as(i32p), b.ctypes.data_as(i32p), vals.ctypes.data_as(f64p))
    return d


def _vector(v, name: str, n: int, keep: list):
    if v is None:
        return None
    a = np.asarray(v)
    if a.ndim != 1:
        raise ValueError(f"{name} must be 1D.")
    if a.size != n:
        raise ValueError(f"{name} length mismatch: expect {n}, got {a.size}")
    a = np.ascontiguousarray(a, dtype=np.float64)
    keep.append(a)
    return a.ctypes.data_as(C.POINTER(C.c_double))


def solve_once(Q, R, A, objective_vector, objective_constant=None, variable_lower_bound=None,
               variable_upper_bound=None, constraint_lower_bound=None, constraint_upper_bound=None,
               zero_tolerance: float = 0.0, params=None, primal_start=None, dual_start=None) -> dict:
    """Build the QP, solve it once and return the result dict."""
    lib = _lib()
    keep: list = []
    dA = None if A is None else _matrix_desc(A, zero_tolerance, keep)
    dQ = None if Q is None else _matrix_desc(Q, zero_tolerance, keep)
    dR = None if R is None else _matrix_desc(R, zero_tolerance, keep)
    n = next((d.n for d in (dA, dQ, dR) if d is not None and d.n > 0), 0)
    m = dA.m if dA is not None and dA.m > 0 else 0
    c = _vector(objective_vector, "objective_vector", n, keep)
    lb = _vector(variable_lower_bound, "variable_lower_bound", n, keep)
    ub = _vector(variable_upper_bound, "variable_upper_bound", n, keep)
    lo = _vector(constraint_lower_bound, "constraint_lower_bound", m, keep)
    up = _vector(constraint_upper_bound, "constraint_upper_bound", m, keep)
    c0 = None if objective_constant is None else C.byref(C.c_double(float(objective_constant)))
    ref = lambda d: None if d is None else C.byref(d)  # noqa: E731
    prob = lib.lib.create_qp_problem(c, ref(dQ), ref(dR), ref(dA), lo, up, lb, ub, c0)
    if not prob:
        raise RuntimeError("create_qp_problem failed.")
    try:
        if primal_start is not None or dual_start is not None:
            ps = _vector(primal_start, "primal_start", n, keep)
            ds = _vector(dual_start, "dual_start", m, keep)
            lib.lib.set_start_values(prob, ps, ds)
        p = _params_from_dict(params)
        res = lib.lib.solve_qp_problem(prob, C.byref(p))
    finally:
        lib.lib.qp_problem_free(prob)
    if not res:
        raise RuntimeError("solve_qp_problem returned NULL.")
    try:
        r = res.contents
        nx, my = r.num_variables, r.num_constraints
        x = np.ctypeslib.as_array(r.primal_solution, shape=(nx,)).copy() if nx else np.zeros(0)
        y = np.ctypeslib.as_array(r.dual_solution, shape=(my,)).copy() if my else np.zeros(0)
        reason = int(r.termination_reason)
        return {
            "X": x, "Pi": y,
            "PrimalObj": r.primal_objective_value, "DualObj": r.dual_objective_value,
            "ObjectiveGap": r.objective_gap, "RelativeObjectiveGap": r.relative_objective_gap,
            "Status": _STATUS.get(reason, "UNKNOWN"), "StatusCode": _CODE.get(reason, -1),
            "Iterations": r.total_count, "RescalingTimeSec": r.rescaling_time_sec,
            "RuntimeSec": r.cumulative_time_sec,
            "RelativePrimalResidual": r.relative_primal_residual,
            "RelativeDualResidual": r.relative_dual_residual,
            "MaxPrimalRayInfeas": r.max_primal_ray_infeasibility,
            "MaxDualRayInfeas": r.max_dual_ray_infeasibility,
            "PrimalRayLinObj": r.primal_ray_linear_objective, "DualRayObj": r.dual_ray_objective,
        }
    finally:
        lib.lib.fa_cp_result_free(res)
