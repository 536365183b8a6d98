"""ctypes binding of the cp_al.h C ABI.

`CpAl(path)` wraps any shared object that exports the reference's LP entry
points (LP/CP_AL/include/cp_al.h:28-47) with the struct layouts of
include/cp_al_types.h.  `load_cp_al()` returns the MI355X product library
(pdal_amd/lib/libcp_al.so) and raises if it has not been built: there is no
fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"

# termination_reason_t (cp_al_types.h)
TERMINATION_REASONS = [
    "UNSPECIFIED", "OPTIMAL", "PRIMAL_INFEASIBLE", "DUAL_INFEASIBLE", "INFEASIBLE_OR_UNBOUNDED",
    "TIME_LIMIT", "ITERATION_LIMIT", "FEAS_POLISH_SUCCESS",
]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class LpProblem(C.Structure):
    _fields_ = [
        ("num_variables", C.c_int),
        ("num_constraints", C.c_int),
        ("variable_lower_bound", _dp),
        ("variable_upper_bound", _dp),
        ("objective_vector", _dp),
        ("objective_constant", C.c_double),
        ("constraint_matrix_row_pointers", _ip),
        ("constraint_matrix_col_indices", _ip),
        ("constraint_matrix_values", _dp),
        ("constraint_matrix_num_nonzeros", C.c_int),
        ("constraint_lower_bound", _dp),
        ("constraint_upper_bound", _dp),
        ("primal_start", _dp),
        ("dual_start", _dp),
    ]


class RestartParameters(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "artificial_restart_threshold", "sufficient_reduction_for_restart",
        "necessary_reduction_for_restart", "k_p", "k_i", "k_d", "i_smooth")]


class TerminationCriteria(C.Structure):
    _fields_ = [
        ("eps_optimal_relative", C.c_double),
        ("eps_feasible_relative", C.c_double),
        ("eps_feas_polish_relative", C.c_double),
        ("time_sec_limit", C.c_double),
        ("iteration_limit", C.c_int),
    ]


class CpAlParameters(C.Structure):
    _fields_ = [
        ("l_inf_ruiz_iterations", C.c_int),
        ("has_pock_chambolle_alpha", C.c_bool),
        ("pock_chambolle_alpha", C.c_double),
        ("bound_objective_rescaling", C.c_bool),
        ("verbose", C.c_bool),
        ("termination_evaluation_frequency", C.c_int),
        ("sv_max_iter", C.c_int),
        ("sv_tol", C.c_double),
        ("termination_criteria", TerminationCriteria),
        ("restart_params", RestartParameters),
        ("reflection_coefficient", C.c_double),
        ("feasibility_polishing", C.c_bool),
        ("optimality_norm", C.c_int),
        ("presolve", C.c_bool),
        ("matrix_zero_tol", C.c_double),
    ]


class CpAlResult(C.Structure):
    _fields_ = [
        ("num_variables", C.c_int),
        ("num_constraints", C.c_int),
        ("num_nonzeros", C.c_int),
        ("num_reduced_variables", C.c_int),
        ("num_reduced_constraints", C.c_int),
        ("num_reduced_nonzeros", C.c_int),
        ("primal_solution", _dp),
        ("dual_solution", _dp),
        ("reduced_cost", _dp),
        ("total_count", C.c_int),
        ("rescaling_time_sec", C.c_double),
        ("cumulative_time_sec", C.c_double),
        ("presolve_time", C.c_double),
        ("presolve_status", C.c_int),
        ("absolute_primal_residual", C.c_double),
        ("relative_primal_residual", C.c_double),
        ("absolute_dual_residual", C.c_double),
        ("relative_dual_residual", C.c_double),
        ("primal_objective_value", C.c_double),
        ("dual_objective_value", C.c_double),
        ("objective_gap", C.c_double),
        ("relative_objective_gap", C.c_double),
        ("max_primal_ray_infeasibility", C.c_double),
        ("max_dual_ray_infeasibility", C.c_double),
        ("primal_ray_linear_objective", C.c_double),
        ("dual_ray_objective", C.c_double),
        ("termination_reason", C.c_int),
        ("feasibility_polishing_time", C.c_double),
        ("feasibility_iteration", C.c_int),
    ]


class _Dense(C.Structure):
    _fields_ = [("A", _dp)]


class _Csr(C.Structure):
    _fields_ = [("nnz", C.c_int), ("row_ptr", _ip), ("col_ind", _ip), ("vals", _dp)]


class _Csc(C.Structure):
    _fields_ = [("nnz", C.c_int), ("col_ptr", _ip), ("row_ind", _ip), ("vals", _dp)]


class _Coo(C.Structure):
    _fields_ = [("nnz", C.c_int), ("row_ind", _ip), ("col_ind", _ip), ("vals", _dp)]


class _MatrixData(C.Union):
    _fields_ = [("dense", _Dense), ("csr", _Csr), ("csc", _Csc), ("coo", _Coo)]


class MatrixDesc(C.Structure):
    _fields_ = [("m", C.c_int), ("n", C.c_int), ("fmt", C.c_int), ("data", _MatrixData)]


MATRIX_DENSE, MATRIX_CSR, MATRIX_CSC, MATRIX_COO = 0, 1, 2, 3


def _f64(a) -> Optional[np.ndarray]:
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i32(a) -> np.ndarray:
    a = np.asarray(a)
    if a.size and (a.max(initial=0) > np.iinfo(np.int32).max or a.min(initial=0) < np.iinfo(np.int32).min):
        raise OverflowError("index does not fit int32")
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


@dataclass
class LpResult:
    x: np.ndarray
    y: np.ndarray
    reduced_cost: np.ndarray
    status: str
    iterations: int
    primal_objective: float
    dual_objective: float
    relative_primal_residual: float
    relative_dual_residual: float
    relative_objective_gap: float
    absolute_primal_residual: float
    absolute_dual_residual: float
    objective_gap: float
    rescaling_time_sec: float
    solve_time_sec: float
    extra: dict = field(default_factory=dict)

    @property
    def kkt(self) -> float:
        return max(self.relative_primal_residual, self.relative_dual_residual, self.relative_objective_gap)


class CpAl:
    """A library exporting the cp_al.h entry points."""

    def __init__(self, path: os.PathLike | str):
        path = Path(path)
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built")
        self.path = path
        self.lib = C.CDLL(str(path), mode=os.RTLD_LOCAL | os.RTLD_NOW)
        L = self.lib
        L.create_lp_problem.restype = C.POINTER(LpProblem)
        L.create_lp_problem.argtypes = [_dp, C.POINTER(MatrixDesc), _dp, _dp, _dp, _dp, _dp]
        L.set_start_values.restype = None
        L.set_start_values.argtypes = [C.POINTER(LpProblem), _dp, _dp]
        L.solve_lp_problem.restype = C.POINTER(CpAlResult)
        L.solve_lp_problem.argtypes = [C.POINTER(LpProblem), C.POINTER(CpAlParameters)]
        L.set_default_parameters.restype = None
        L.set_default_parameters.argtypes = [C.POINTER(CpAlParameters)]
        L.cp_al_result_free.restype = None
        L.cp_al_result_free.argtypes = [C.POINTER(CpAlResult)]
        L.lp_problem_free.restype = None
        L.lp_problem_free.argtypes = [C.POINTER(LpProblem)]
        self._reset = None
        for name in ("pdal_cp_al_reset_rng", "cp_al_oracle_reset_rng"):
            if hasattr(L, name):
                self._reset = getattr(L, name)
                self._reset.restype = None
                self._reset.argtypes = []
                break
        if hasattr(L, "cp_al_oracle_solve_many"):
            L.cp_al_oracle_solve_many.restype = None
            L.cp_al_oracle_solve_many.argtypes = [C.POINTER(C.POINTER(LpProblem)), C.c_int,
                                                  C.POINTER(CpAlParameters), C.c_int,
                                                  C.POINTER(C.POINTER(CpAlResult))]
            L.cp_al_oracle_max_threads.restype = C.c_int
            L.cp_al_oracle_max_threads.argtypes = []
        if hasattr(L, "pdal_cp_al_time_kernels"):
            L.pdal_cp_al_time_kernels.restype = C.c_int
            L.pdal_cp_al_time_kernels.argtypes = [C.POINTER(LpProblem), C.POINTER(CpAlParameters), C.c_int, _dp]
        if hasattr(L, "pdal_cp_al_session_create"):
            L.pdal_cp_al_session_create.restype = C.c_void_p
            L.pdal_cp_al_session_create.argtypes = [C.POINTER(LpProblem), C.POINTER(CpAlParameters)]
            L.pdal_cp_al_session_solve.restype = C.POINTER(CpAlResult)
            L.pdal_cp_al_session_solve.argtypes = [C.c_void_p]
            L.pdal_cp_al_session_time_kernels.restype = C.c_int
            L.pdal_cp_al_session_time_kernels.argtypes = [C.c_void_p, C.c_int, _dp]
            L.pdal_cp_al_session_setup_seconds.restype = C.c_double
            L.pdal_cp_al_session_setup_seconds.argtypes = [C.c_void_p]
            L.pdal_cp_al_session_free.restype = None
            L.pdal_cp_al_session_free.argtypes = [C.c_void_p]
        if hasattr(L, "pdal_gather_calibration"):
            L.pdal_gather_calibration.restype = C.c_double
            L.pdal_gather_calibration.argtypes = [C.c_longlong, C.c_longlong, C.c_int]
        if hasattr(L, "pdal_copy_calibration"):
            L.pdal_copy_calibration.restype = C.c_double
            L.pdal_copy_calibration.argtypes = [C.c_longlong, C.c_int]
        if hasattr(L, "pdal_cp_al_batch_create"):
            L.pdal_cp_al_batch_create.restype = C.c_void_p
            L.pdal_cp_al_batch_create.argtypes = [C.POINTER(C.POINTER(LpProblem)), C.c_int,
                                                  C.POINTER(CpAlParameters)]
            L.pdal_cp_al_batch_set_starts.restype = C.c_int
            L.pdal_cp_al_batch_set_starts.argtypes = [C.c_void_p, _dp, _dp]
            L.pdal_cp_al_batch_solve.restype = C.c_int
            L.pdal_cp_al_batch_solve.argtypes = [C.c_void_p]
            L.pdal_cp_al_batch_result.restype = C.POINTER(CpAlResult)
            L.pdal_cp_al_batch_result.argtypes = [C.c_void_p, C.c_int]
            L.pdal_cp_al_batch_stats.restype = None
            L.pdal_cp_al_batch_stats.argtypes = [C.c_void_p, _dp]
            L.pdal_cp_al_batch_free.restype = None
            L.pdal_cp_al_batch_free.argtypes = [C.c_void_p]
            L.pdal_cp_al_batch_primal_weights.restype = C.c_int
            L.pdal_cp_al_batch_primal_weights.argtypes = [C.c_void_p, _dp]
            L.pdal_cp_al_batch_set_primal_weights.restype = C.c_int
            L.pdal_cp_al_batch_set_primal_weights.argtypes = [C.c_void_p, _dp]
            L.pdal_cp_al_batch_time_kernels.restype = C.c_int
            L.pdal_cp_al_batch_time_kernels.argtypes = [C.c_void_p, C.c_int, _dp]
            L.pdal_cp_al_batch_start_buffers.restype = C.c_int
            L.pdal_cp_al_batch_start_buffers.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]
            L.pdal_cp_al_batch_set_starts_from_buffers.restype = C.c_int
            L.pdal_cp_al_batch_set_starts_from_buffers.argtypes = [C.c_void_p]
            if hasattr(L, "pdal_cp_al_batch_set_vectors"):
                for name in ("pdal_cp_al_batch_set_vectors", "pdal_cp_al_batch_set_vectors_device"):
                    getattr(L, name).restype = C.c_int
                    getattr(L, name).argtypes = [C.c_void_p] + [C.c_void_p] * 5

    # ------------------------------------------------------------------
    def default_params(self, **overrides) -> CpAlParameters:
        p = CpAlParameters()
        self.lib.set_default_parameters(C.byref(p))
        p.presolve = False
        for k, v in overrides.items():
            set_param(p, k, v)
        return p

    def reset_rng(self) -> None:
        if self._reset is None:
            raise AttributeError(f"{self.path.name} has no generator reset entry point")
        self._reset()

    def make_problem(self, A, c=None, con_lb=None, con_ub=None, var_lb=None, var_ub=None, obj_const=None,
                     primal_start=None, dual_start=None, fmt: str = "csr"):
        """Build an lp_problem_t from a matrix passed to create_lp_problem in
        format `fmt` ("csr", "csc", "coo" or "dense")."""
        import scipy.sparse as sp
        if fmt == "dense":
            dense = np.ascontiguousarray(A.toarray() if sp.issparse(A) else A, dtype=np.float64)
            m, n = dense.shape
        else:
            S = {"csr": sp.csr_matrix, "csc": sp.csc_matrix, "coo": sp.coo_matrix}[fmt](A)
            m, n = S.shape
        keep = dict(c=_f64(c), l=_f64(con_lb), u=_f64(con_ub), lb=_f64(var_lb), ub=_f64(var_ub))
        desc = MatrixDesc()
        desc.m, desc.n = m, n
        if fmt == "dense":
            desc.fmt = MATRIX_DENSE
            keep["dense"] = dense
            desc.data.dense.A = _ptr(dense, C.c_double)
        elif fmt == "csr":
            desc.fmt = MATRIX_CSR
            keep.update(p=_i32(S.indptr), i=_i32(S.indices), v=_f64(S.data))
            desc.data.csr.nnz = int(S.nnz)
            desc.data.csr.row_ptr = _ptr(keep["p"], C.c_int)
            desc.data.csr.col_ind = _ptr(keep["i"], C.c_int)
            desc.data.csr.vals = _ptr(keep["v"], C.c_double)
        elif fmt == "csc":
            desc.fmt = MATRIX_CSC
            keep.update(p=_i32(S.indptr), i=_i32(S.indices), v=_f64(S.data))
            desc.data.csc.nnz = int(S.nnz)
            desc.data.csc.col_ptr = _ptr(keep["p"], C.c_int)
            desc.data.csc.row_ind = _ptr(keep["i"], C.c_int)
            desc.data.csc.vals = _ptr(keep["v"], C.c_double)
        else:
            desc.fmt = MATRIX_COO
            keep.update(r=_i32(S.row), i=_i32(S.col), v=_f64(S.data))
            desc.data.coo.nnz = int(S.nnz)
            desc.data.coo.row_ind = _ptr(keep["r"], C.c_int)
            desc.data.coo.col_ind = _ptr(keep["i"], C.c_int)
            desc.data.coo.vals = _ptr(keep["v"], C.c_double)
        oc = None if obj_const is None else C.byref(C.c_double(float(obj_const)))
        prob = self.lib.create_lp_problem(_ptr(keep["c"], C.c_double), C.byref(desc), _ptr(keep["l"], C.c_double),
                                          _ptr(keep["u"], C.c_double), _ptr(keep["lb"], C.c_double),
                                          _ptr(keep["ub"], C.c_double), oc)
        if not prob:
            raise ValueError("create_lp_problem returned NULL")
        if primal_start is not None or dual_start is not None:
            ps, ds = _f64(primal_start), _f64(dual_start)
            self.lib.set_start_values(prob, _ptr(ps, C.c_double), _ptr(ds, C.c_double))
        return prob

    def free_problem(self, prob) -> None:
        self.lib.lp_problem_free(prob)

    def solve_problem(self, prob, params: Optional[CpAlParameters] = None) -> LpResult:
        res = self.lib.solve_lp_problem(prob, None if params is None else C.byref(params))
        if not res:
            raise RuntimeError("solve_lp_problem returned NULL")
        return self._take_result(res, prob)

    def solve_many(self, instances: list, params: CpAlParameters, threads: int) -> list:
        """Solve a list of problems (keyword dicts of make_problem) on `threads`
        host threads through the oracle's parallel entry point (CPU baseline
        only); every solve starts from a freshly seeded generator."""
        if not hasattr(self.lib, "cp_al_oracle_solve_many"):
            raise AttributeError(f"{self.path.name} has no parallel solve entry point")
        probs = [self.make_problem(**kw) for kw in instances]
        arr = (C.POINTER(LpProblem) * len(probs))(*probs)
        out = (C.POINTER(CpAlResult) * len(probs))()
        try:
            self.lib.cp_al_oracle_solve_many(arr, len(probs), C.byref(params), int(threads), out)
            res = []
            for k, p in enumerate(probs):
                if not out[k]:
                    raise RuntimeError(f"oracle solve {k} returned NULL")
                res.append(self._take_result(out[k], p))
            return res
        finally:
            for p in probs:
                self.free_problem(p)

    def _take_result(self, res, prob) -> LpResult:
        try:
            r = res.contents
            n, m = prob.contents.num_variables, prob.contents.num_constraints
            out = LpResult(
                x=np.ctypeslib.as_array(r.primal_solution, shape=(n,)).copy() if n else np.zeros(0),
                y=np.ctypeslib.as_array(r.dual_solution, shape=(m,)).copy() if m else np.zeros(0),
                reduced_cost=np.ctypeslib.as_array(r.reduced_cost, shape=(n,)).copy() if n else np.zeros(0),
                status=TERMINATION_REASONS[r.termination_reason],
                iterations=r.total_count,
                primal_objective=r.primal_objective_value,
                dual_objective=r.dual_objective_value,
                relative_primal_residual=r.relative_primal_residual,
                relative_dual_residual=r.relative_dual_residual,
                relative_objective_gap=r.relative_objective_gap,
                absolute_primal_residual=r.absolute_primal_residual,
                absolute_dual_residual=r.absolute_dual_residual,
                objective_gap=r.objective_gap,
                rescaling_time_sec=r.rescaling_time_sec,
                solve_time_sec=r.cumulative_time_sec,
                extra={"feasibility_polishing_time": r.feasibility_polishing_time,
                       "feasibility_iteration": r.feasibility_iteration,
                       "presolve_status": r.presolve_status, "presolve_time": r.presolve_time,
                       "num_reduced_variables": r.num_reduced_variables,
                       "num_reduced_constraints": r.num_reduced_constraints,
                       "num_reduced_nonzeros": r.num_reduced_nonzeros},
            )
        finally:
            self.lib.cp_al_result_free(res)
        return out

    def solve(self, A, c=None, con_lb=None, con_ub=None, var_lb=None, var_ub=None, obj_const=None,
              params: Optional[CpAlParameters] = None, primal_start=None, dual_start=None,
              fmt: str = "csr") -> LpResult:
        prob = self.make_problem(A, c, con_lb, con_ub, var_lb, var_ub, obj_const, primal_start, dual_start, fmt)
        try:
            return self.solve_problem(prob, params)
        finally:
            self.free_problem(prob)

    def time_kernels(self, A, c, con_lb, con_ub, var_lb, var_ub, params=None, iters: int = 20) -> dict:
        if not hasattr(self.lib, "pdal_cp_al_time_kernels"):
            raise AttributeError("library has no kernel timing entry point")
        prob = self.make_problem(A, c, con_lb, con_ub, var_lb, var_ub)
        out = (C.c_double * 8)()
        try:
            rc = self.lib.pdal_cp_al_time_kernels(prob, None if params is None else C.byref(params), iters, out)
        finally:
            self.free_problem(prob)
        if rc != 0:
            raise RuntimeError(f"pdal_cp_al_time_kernels failed ({rc})")
        return dict(primal_ms=out[0], dual_ms=out[1], predictor_ms=out[2], iteration_ms=out[3], eval_ms=out[4],
                    lanes_at=int(out[5]), lanes_a=int(out[6]))


class LpSession:
    """Device-resident solve session of the product library (pdal_lp.h)."""

    def __init__(self, lib: CpAl, A, c=None, con_lb=None, con_ub=None, var_lb=None, var_ub=None,
                 params: Optional[CpAlParameters] = None, primal_start=None, dual_start=None):
        if not hasattr(lib.lib, "pdal_cp_al_session_create"):
            raise AttributeError(f"{lib.path.name} has no session entry points")
        self.lib = lib
        self.prob = lib.make_problem(A, c, con_lb, con_ub, var_lb, var_ub, None, primal_start, dual_start)
        self.params = params if params is not None else lib.default_params()
        self.h = lib.lib.pdal_cp_al_session_create(self.prob, C.byref(self.params))
        if not self.h:
            lib.free_problem(self.prob)
            raise RuntimeError("pdal_cp_al_session_create failed")

    @property
    def setup_seconds(self) -> float:
        return float(self.lib.lib.pdal_cp_al_session_setup_seconds(self.h))

    def solve(self) -> LpResult:
        res = self.lib.lib.pdal_cp_al_session_solve(self.h)
        if not res:
            raise RuntimeError("pdal_cp_al_session_solve failed")
        return self.lib._take_result(res, self.prob)

    def time_kernels(self, iters: int = 20) -> dict:
        out = (C.c_double * 8)()
        rc = self.lib.lib.pdal_cp_al_session_time_kernels(self.h, iters, out)
        if rc != 0:
            raise RuntimeError(f"pdal_cp_al_session_time_kernels failed ({rc})")
        return dict(primal_ms=out[0], dual_ms=out[1], predictor_ms=out[2], iteration_ms=out[3], eval_ms=out[4],
                    lanes_at=int(out[5]), lanes_a=int(out[6]))

    def close(self) -> None:
        if self.h:
            self.lib.lib.pdal_cp_al_session_free(self.h)
            self.h = None
            self.lib.free_problem(self.prob)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class LpBatch:
    """Batched device-resident solve of many LPs (pdal_batch.h).  `instances`
    is a list of keyword dicts accepted by CpAl.make_problem."""

    def __init__(self, lib: CpAl, instances: list, params: Optional[CpAlParameters] = None):
        if not hasattr(lib.lib, "pdal_cp_al_batch_create"):
            raise AttributeError(f"{lib.path.name} has no batch entry points")
        self.lib = lib
        self.probs = [lib.make_problem(**kw) for kw in instances]
        arr = (C.POINTER(LpProblem) * len(self.probs))(*self.probs)
        self.params = params if params is not None else lib.default_params()
        self.h = lib.lib.pdal_cp_al_batch_create(arr, len(self.probs), C.byref(self.params))
        if not self.h:
            self._free_probs()
            raise RuntimeError("pdal_cp_al_batch_create failed")
        self.sizes = [(p.contents.num_constraints, p.contents.num_variables) for p in self.probs]

    def _free_probs(self) -> None:
        for p in self.probs:
            self.lib.free_problem(p)
        self.probs = []

    def set_starts(self, x=None, y=None) -> None:
        """x: the members' primal starts concatenated (sum n_i doubles), y:
        their dual starts (sum m_i); None keeps the current one."""
        xs, ys = _f64(x), _f64(y)
        if xs is not None and xs.size != sum(ni for _, ni in self.sizes):
            raise ValueError(f"x has {xs.size} entries, the batch has {sum(ni for _, ni in self.sizes)} columns")
        if ys is not None and ys.size != sum(mi for mi, _ in self.sizes):
            raise ValueError(f"y has {ys.size} entries, the batch has {sum(mi for mi, _ in self.sizes)} rows")
        rc = self.lib.lib.pdal_cp_al_batch_set_starts(self.h, _ptr(xs, C.c_double), _ptr(ys, C.c_double))
        if rc != 0:
            raise RuntimeError(f"pdal_cp_al_batch_set_starts failed ({rc})")

    def start_buffers(self):
        """Device addresses (ints) of the unscaled start buffers (x: sum n_i,
        y: sum m_i doubles, concatenated) a device-side producer writes."""
        px, py = C.c_void_p(), C.c_void_p()
        if self.lib.lib.pdal_cp_al_batch_start_buffers(self.h, C.byref(px), C.byref(py)) != 0:
            raise RuntimeError("pdal_cp_al_batch_start_buffers failed")
        return px.value, py.value

    def set_starts_from_buffers(self) -> None:
        if self.lib.lib.pdal_cp_al_batch_set_starts_from_buffers(self.h) != 0:
            raise RuntimeError("pdal_cp_al_batch_set_starts_from_buffers failed")

    def set_vectors(self, c=None, con_lb=None, con_ub=None, var_lb=None, var_ub=None) -> None:
        """New objective / bound vectors for the same matrices, each the
        members' vectors concatenated (c, var_lb, var_ub: sum n_i; con_lb,
        con_ub: sum m_i); None keeps the current one.  The solves that follow
        equal those of a batch created with these vectors; starts become zero
        and primal weights their defaults."""
        n, m = sum(ni for _, ni in self.sizes), sum(mi for mi, _ in self.sizes)
        arrs = []
        for name, v, size in (("c", c, n), ("con_lb", con_lb, m), ("con_ub", con_ub, m), ("var_lb", var_lb, n),
                              ("var_ub", var_ub, n)):
            a = _f64(v)
            if a is not None and a.size != size:
                raise ValueError(f"{name} has {a.size} entries, the batch needs {size}")
            arrs.append(a)
        rc = self.lib.lib.pdal_cp_al_batch_set_vectors(self.h, *[None if a is None else a.ctypes.data for a in arrs])
        if rc != 0:
            raise RuntimeError(f"pdal_cp_al_batch_set_vectors failed ({rc})")

    def set_vectors_device(self, c=None, con_lb=None, con_ub=None, var_lb=None, var_ub=None) -> None:
        """As set_vectors from device memory: each argument a device address
        (int) or a devbuf.DeviceArray of the right size; None keeps."""
        n, m = sum(ni for _, ni in self.sizes), sum(mi for mi, _ in self.sizes)
        ptrs = []
        for name, v, size in (("c", c, n), ("con_lb", con_lb, m), ("con_ub", con_ub, m), ("var_lb", var_lb, n),
                              ("var_ub", var_ub, n)):
            if v is None or isinstance(v, int):
                ptrs.append(v)
                continue
            if v.size != size:
                raise ValueError(f"{name} has {v.size} entries, the batch needs {size}")
            ptrs.append(v.ptr)
        rc = self.lib.lib.pdal_cp_al_batch_set_vectors_device(self.h, *ptrs)
        if rc != 0:
            raise RuntimeError(f"pdal_cp_al_batch_set_vectors_device failed ({rc})")

    def set_instances(self, instances: list) -> None:
        """set_vectors from keyword dicts (as for the constructor) whose
        matrices are the batch's: only their vectors are read."""
        if len(instances) != len(self.sizes):
            raise ValueError(f"{len(instances)} instances for a batch of {len(self.sizes)}")
        cat = {}
        defaults = {"c": 0.0, "con_lb": -np.inf, "con_ub": np.inf, "var_lb": -np.inf, "var_ub": np.inf}
        for key, default in defaults.items():  # create_lp_problem's values for an absent vector
            parts = []
            for (mi, ni), kw in zip(self.sizes, instances):
                size = ni if key in ("c", "var_lb", "var_ub") else mi
                v = kw.get(key)
                parts.append(np.full(size, default) if v is None else np.asarray(v, dtype=np.float64).ravel())
            cat[key] = np.concatenate(parts) if parts else np.zeros(0)
        self.set_vectors(**cat)

    def solve(self) -> int:
        n_opt = self.lib.lib.pdal_cp_al_batch_solve(self.h)
        if n_opt < 0:
            raise RuntimeError("pdal_cp_al_batch_solve failed")
        return n_opt

    def result(self, i: int) -> LpResult:
        res = self.lib.lib.pdal_cp_al_batch_result(self.h, i)
        if not res:
            raise IndexError(i)
        return self.lib._take_result(res, self.probs[i])

    def stats(self) -> dict:
        out = (C.c_double * 18)()
        self.lib.lib.pdal_cp_al_batch_stats(self.h, out)
        return dict(setup_s=out[0], solve_s=out[1], replays=int(out[2]), instances=int(out[3]),
                    optimal=int(out[4]), iterations=int(out[5]), lanes_a=int(out[6]), lanes_at=int(out[7]),
                    compactions=int(out[8]), half_stopped_s=out[9], p99_stopped_s=out[10],
                    staged_at=int(out[11]), staged_a=int(out[12]), split_a=int(out[13]),
                    shared_a=bool(out[14]), handoffs=int(out[15]), shared_bounds=bool(out[16]),
                    equality_rows=bool(out[17]))

    def primal_weights(self) -> np.ndarray:
        out = np.empty(len(self.sizes), dtype=np.float64)
        self.lib.lib.pdal_cp_al_batch_primal_weights(self.h, out.ctypes.data_as(_dp))
        return out

    def set_primal_weights(self, pw) -> None:
        pw = np.ascontiguousarray(pw, dtype=np.float64)
        if pw.size != len(self.sizes) or self.lib.lib.pdal_cp_al_batch_set_primal_weights(
                self.h, pw.ctypes.data_as(_dp)) != 0:
            raise RuntimeError("pdal_cp_al_batch_set_primal_weights failed")

    def time_kernels(self, iters: int = 20) -> dict:
        out = (C.c_double * 8)()
        rc = self.lib.lib.pdal_cp_al_batch_time_kernels(self.h, iters, out)
        if rc != 0:
            raise RuntimeError(f"pdal_cp_al_batch_time_kernels failed ({rc})")
        return dict(primal_ms=out[0], dual_ms=out[1], lanes_at=int(out[2]), lanes_a=int(out[3]), nnz=int(out[4]),
                    n=int(out[5]), m=int(out[6]), primal_blocks=int(out[7]))

    def close(self) -> None:
        if self.h:
            self.lib.lib.pdal_cp_al_batch_free(self.h)
            self.h = None
        self._free_probs()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def set_param(p: CpAlParameters, key: str, value) -> None:
    """Set a parameter by its cp_al_parameters_t name (nested names with '.')."""
    if key in ("iteration_limit", "eps_optimal_relative", "eps_feasible_relative", "time_sec_limit",
               "eps_feas_polish_relative"):
        setattr(p.termination_criteria, key, value)
    elif key in ("artificial_restart_threshold", "sufficient_reduction_for_restart",
                 "necessary_reduction_for_restart", "k_p", "k_i", "k_d", "i_smooth"):
        setattr(p.restart_params, key, value)
    elif key == "optimality_norm" and isinstance(value, str):
        p.optimality_norm = {"l2": 0, "linf": 1}[value]
    else:
        if not hasattr(p, key):
            raise KeyError(key)
        setattr(p, key, value)


def load_cp_al() -> CpAl:
    """The MI355X libcp_al.so; raises FileNotFoundError if it is not built."""
    return CpAl(LIB_DIR / "libcp_al.so")
