"""ctypes binding of the fa_cp.h C ABI.

`FaCp(path)` wraps any shared object exporting the reference's QP entry points
(QP/FA_CP/include/fa_cp.h:28-49) with the struct layouts of
include/fa_cp_types.h.  `load_fa_cp()` returns the MI355X product library
(pdal_amd/lib/libfa_cp.so) and raises if it is not built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from pathlib import Path
from typing import Optional

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"

TERMINATION_REASONS = [
    "UNSPECIFIED", "OPTIMAL", "PRIMAL_INFEASIBLE", "DUAL_INFEASIBLE", "INFEASIBLE_OR_UNBOUNDED",
    "TIME_LIMIT", "ITERATION_LIMIT", "FEAS_POLISH_SUCCESS",
]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class CsrComponent(C.Structure):
    _fields_ = [("row_ptr", _ip), ("col_ind", _ip), ("val", _dp)]


class QpProblem(C.Structure):
    _fields_ = [
        ("num_variables", C.c_int),
        ("num_constraints", C.c_int),
        ("num_rank_lowrank_obj", C.c_int),
        ("variable_lower_bound", _dp),
        ("variable_upper_bound", _dp),
        ("objective_vector", _dp),
        ("objective_constant", C.c_double),
        ("constraint_matrix", C.POINTER(CsrComponent)),
        ("constraint_matrix_num_nonzeros", C.c_int),
        ("objective_sparse_matrix", C.POINTER(CsrComponent)),
        ("objective_sparse_matrix_num_nonzeros", C.c_int),
        ("objective_lowrank_matrix", C.POINTER(CsrComponent)),
        ("objective_lowrank_matrix_num_nonzeros", C.c_int),
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
        ("eps_infeasible", C.c_double),
        ("time_sec_limit", C.c_double),
        ("iteration_limit", C.c_int),
    ]


class FaCpParameters(C.Structure):
    _fields_ = [
        ("l_inf_ruiz_iterations", C.c_int),
        ("has_pock_chambolle_alpha", C.c_bool),
        ("pock_chambolle_alpha", C.c_double),
        ("bound_objective_rescaling", C.c_bool),
        ("trajectory_csv", C.c_char_p),
        ("use_al_qp", C.c_bool),
        ("al_sigma", C.c_double),
        ("sigma_update_rule", C.c_int),
        ("sigma_update_gain", C.c_double),
        ("sigma_clip_low", C.c_double),
        ("sigma_clip_high", C.c_double),
        ("sigma_min", C.c_double),
        ("sigma_max", C.c_double),
        ("sigma_rollback_tol", C.c_double),
        ("sigma_zero_threshold", C.c_double),
        ("sigma_activation_threshold", C.c_double),
        ("sigma_normal_motion_threshold", C.c_double),
        ("sigma_activation_value", C.c_double),
        ("verbose", C.c_int),
        ("termination_evaluation_frequency", C.c_int),
        ("sv_max_iter", C.c_int),
        ("sv_tol", C.c_double),
        ("termination_criteria", TerminationCriteria),
        ("restart_params", RestartParameters),
        ("reflection_coefficient", C.c_double),
        ("feasibility_polishing", C.c_bool),
        ("optimality_norm", C.c_int),
    ]


class FaCpResult(C.Structure):
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
        ("total_inner_count", C.c_int),
        ("rescaling_time_sec", C.c_double),
        ("cumulative_time_sec", C.c_double),
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
        ("final_al_sigma", C.c_double),
        ("sigma_update_count", C.c_int),
        ("sigma_update_rule", C.c_int),
        ("termination_reason", C.c_int),
        ("feasibility_polishing_time", C.c_double),
        ("feasibility_iteration", C.c_int),
    ]


class LinCpAlParameters(C.Structure):
    """cp_al_parameters_t of the linearized CP-AL QP solver (lin_cp_al_types.h)."""
    _fields_ = [
        ("l_inf_ruiz_iterations", C.c_int),
        ("has_pock_chambolle_alpha", C.c_bool),
        ("pock_chambolle_alpha", C.c_double),
        ("bound_objective_rescaling", C.c_bool),
        ("use_al_qp", C.c_bool),
        ("sigma_init", C.c_double),
        ("verbose", C.c_int),
        ("termination_evaluation_frequency", C.c_int),
        ("sv_max_iter", C.c_int),
        ("sv_tol", C.c_double),
        ("termination_criteria", TerminationCriteria),
        ("restart_params", RestartParameters),
        ("reflection_coefficient", C.c_double),
        ("feasibility_polishing", C.c_bool),
        ("optimality_norm", C.c_int),
    ]


class PlantedCpAlParameters(C.Structure):
    """cp_al_parameters_t of the planted-KKT CP-AL (planted_cp_al_types.h): the
    linearized solver's with trajectory_csv after sigma_init."""
    _fields_ = (LinCpAlParameters._fields_[:6] + [("trajectory_csv", C.c_char_p)]
                + LinCpAlParameters._fields_[6:])


class LinCpAlResult(C.Structure):
    """cp_al_result_t of the linearized CP-AL QP solver: FaCpResult without the sigma fields."""
    _fields_ = [f for f in FaCpResult._fields_
                if f[0] not in ("final_al_sigma", "sigma_update_count", "sigma_update_rule")]


class PlantedFaCpParameters(C.Structure):
    """fa_cp_parameters_t of the planted-KKT FA_CP (planted_fa_cp_types.h)."""
    _fields_ = [
        ("l_inf_ruiz_iterations", C.c_int),
        ("has_pock_chambolle_alpha", C.c_bool),
        ("pock_chambolle_alpha", C.c_double),
        ("bound_objective_rescaling", C.c_bool),
        ("use_al_qp", C.c_bool),
        ("al_sigma", C.c_double),
        ("adaptive_al_sigma", C.c_bool),
        ("adaptive_al_sigma_mode", C.c_int),
        ("adaptive_al_sigma_rule", C.c_int),
        ("al_sigma_min", C.c_double),
        ("al_sigma_max", C.c_double),
        ("al_sigma_increase_factor", C.c_double),
        ("al_sigma_decrease_factor", C.c_double),
        ("al_sigma_primal_trigger", C.c_double),
        ("al_sigma_gap_trigger", C.c_double),
        ("trajectory_csv", C.c_char_p),
        ("verbose", C.c_int),
        ("termination_evaluation_frequency", C.c_int),
        ("sv_max_iter", C.c_int),
        ("sv_tol", C.c_double),
        ("termination_criteria", TerminationCriteria),
        ("restart_params", RestartParameters),
        ("reflection_coefficient", C.c_double),
        ("feasibility_polishing", C.c_bool),
        ("optimality_norm", C.c_int),
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
    _fields_ = [("m", C.c_int), ("n", C.c_int), ("fmt", C.c_int), ("zero_tolerance", C.c_double),
                ("data", _MatrixData)]


MATRIX_DENSE, MATRIX_CSR, MATRIX_CSC, MATRIX_COO = 0, 1, 2, 3


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    a = np.asarray(a)
    if a.size and (a.max(initial=0) > np.iinfo(np.int32).max or a.min(initial=0) < np.iinfo(np.int32).min):
        raise OverflowError("index does not fit int32")
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def make_desc(M, fmt: str, keep: list, zero_tolerance: float = 0.0) -> MatrixDesc:
    """A matrix_desc_t for M (scipy.sparse or dense) in format `fmt`; arrays
    the descriptor points to are appended to `keep`."""
    import scipy.sparse as sp
    d = MatrixDesc()
    d.zero_tolerance = float(zero_tolerance)
    if fmt == "dense":
        D = np.ascontiguousarray(M.toarray() if sp.issparse(M) else M, dtype=np.float64)
        d.m, d.n, d.fmt = D.shape[0], D.shape[1], MATRIX_DENSE
        keep.append(D)
        d.data.dense.A = _ptr(D, C.c_double)
        return d
    S = {"csr": sp.csr_matrix, "csc": sp.csc_matrix, "coo": sp.coo_matrix}[fmt](M)
    d.m, d.n = S.shape
    if fmt == "csr":
        p, i, v = _i32(S.indptr), _i32(S.indices), _f64(S.data)
        d.fmt = MATRIX_CSR
        d.data.csr.nnz, d.data.csr.row_ptr, d.data.csr.col_ind, d.data.csr.vals = (
            int(S.nnz), _ptr(p, C.c_int), _ptr(i, C.c_int), _ptr(v, C.c_double))
    elif fmt == "csc":
        p, i, v = _i32(S.indptr), _i32(S.indices), _f64(S.data)
        d.fmt = MATRIX_CSC
        d.data.csc.nnz, d.data.csc.col_ptr, d.data.csc.row_ind, d.data.csc.vals = (
            int(S.nnz), _ptr(p, C.c_int), _ptr(i, C.c_int), _ptr(v, C.c_double))
    else:
        p, i, v = _i32(S.row), _i32(S.col), _f64(S.data)
        d.fmt = MATRIX_COO
        d.data.coo.nnz, d.data.coo.row_ind, d.data.coo.col_ind, d.data.coo.vals = (
            int(S.nnz), _ptr(p, C.c_int), _ptr(i, C.c_int), _ptr(v, C.c_double))
    keep.extend([p, i, v])
    return d


@dataclass
class QpResult:
    x: np.ndarray
    y: np.ndarray
    reduced_cost: np.ndarray
    status: str
    status_code: int
    iterations: int
    inner_iterations: int
    primal_objective: float
    dual_objective: float
    relative_primal_residual: float
    relative_dual_residual: float
    relative_objective_gap: float
    absolute_primal_residual: float
    absolute_dual_residual: float
    objective_gap: float
    max_primal_ray_infeasibility: float
    max_dual_ray_infeasibility: float
    primal_ray_linear_objective: float
    dual_ray_objective: float
    final_al_sigma: float
    sigma_update_count: int
    rescaling_time_sec: float
    solve_time_sec: float
    feasibility_polishing_time: float = 0.0
    feasibility_iteration: int = 0

    @property
    def kkt(self) -> float:
        return max(self.relative_primal_residual, self.relative_dual_residual, self.relative_objective_gap)


class FaCp:
    """A library exporting the fa_cp.h entry points."""
    Params = FaCpParameters
    Result = FaCpResult
    result_free = "fa_cp_result_free"
    reset_names = ("pdal_fa_cp_reset_rng", "fa_cp_oracle_reset_rng")

    def __init__(self, path):
        path = Path(path)
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built")
        self.path = path
        self.lib = C.CDLL(str(path), mode=os.RTLD_LOCAL | os.RTLD_NOW)
        L = self.lib
        L.create_qp_problem.restype = C.POINTER(QpProblem)
        L.create_qp_problem.argtypes = [_dp, C.POINTER(MatrixDesc), C.POINTER(MatrixDesc), C.POINTER(MatrixDesc),
                                        _dp, _dp, _dp, _dp, _dp]
        L.set_start_values.restype = None
        L.set_start_values.argtypes = [C.POINTER(QpProblem), _dp, _dp]
        L.solve_qp_problem.restype = C.POINTER(self.Result)
        L.solve_qp_problem.argtypes = [C.POINTER(QpProblem), C.POINTER(self.Params)]
        L.set_default_parameters.restype = None
        L.set_default_parameters.argtypes = [C.POINTER(self.Params)]
        self._free = getattr(L, self.result_free)
        self._free.restype = None
        self._free.argtypes = [C.POINTER(self.Result)]
        L.qp_problem_free.restype = None
        L.qp_problem_free.argtypes = [C.POINTER(QpProblem)]
        self._reset = None
        for name in self.reset_names:
            if hasattr(L, name):
                self._reset = getattr(L, name)
                self._reset.restype = None
                self._reset.argtypes = []
                break

    def default_params(self, **overrides):
        p = self.Params()
        self.lib.set_default_parameters(C.byref(p))
        p.verbose = 0
        for k, v in overrides.items():
            set_param(p, k, v)
        return p

    def reset_rng(self) -> None:
        self._reset()

    def make_problem(self, c, Q=None, R=None, A=None, con_lb=None, con_ub=None, var_lb=None, var_ub=None,
                     obj_const=None, fmt: str = "csr", zero_tolerance: float = 0.0):
        keep: list = []
        dq = make_desc(Q, fmt, keep, zero_tolerance) if Q is not None else None
        dr = make_desc(R, fmt, keep, zero_tolerance) if R is not None else None
        da = make_desc(A, fmt, keep, zero_tolerance) if A is not None else None
        vecs = [_f64(v) for v in (c, con_lb, con_ub, var_lb, var_ub)]
        oc = None if obj_const is None else C.byref(C.c_double(float(obj_const)))
        prob = self.lib.create_qp_problem(
            _ptr(vecs[0], C.c_double), None if dq is None else C.byref(dq), None if dr is None else C.byref(dr),
            None if da is None else C.byref(da), *[_ptr(v, C.c_double) for v in vecs[1:]], oc)
        if not prob:
            raise ValueError("create_qp_problem returned NULL")
        return prob

    def free_problem(self, prob) -> None:
        self.lib.qp_problem_free(prob)

    def take_result(self, res, n: int, m: int) -> QpResult:
        try:
            r = res.contents
            arr = lambda p, k: np.ctypeslib.as_array(p, shape=(k,)).copy() if k else np.zeros(0)  # noqa: E731
            return QpResult(
                x=arr(r.primal_solution, n), y=arr(r.dual_solution, m), reduced_cost=arr(r.reduced_cost, n),
                status=TERMINATION_REASONS[r.termination_reason], status_code=r.termination_reason,
                iterations=r.total_count, inner_iterations=r.total_inner_count,
                primal_objective=r.primal_objective_value, dual_objective=r.dual_objective_value,
                relative_primal_residual=r.relative_primal_residual,
                relative_dual_residual=r.relative_dual_residual,
                relative_objective_gap=r.relative_objective_gap,
                absolute_primal_residual=r.absolute_primal_residual,
                absolute_dual_residual=r.absolute_dual_residual, objective_gap=r.objective_gap,
                max_primal_ray_infeasibility=r.max_primal_ray_infeasibility,
                max_dual_ray_infeasibility=r.max_dual_ray_infeasibility,
                primal_ray_linear_objective=r.primal_ray_linear_objective, dual_ray_objective=r.dual_ray_objective,
                final_al_sigma=getattr(r, "final_al_sigma", 0.0),
                sigma_update_count=getattr(r, "sigma_update_count", 0),
                rescaling_time_sec=r.rescaling_time_sec, solve_time_sec=r.cumulative_time_sec,
                feasibility_polishing_time=r.feasibility_polishing_time,
                feasibility_iteration=r.feasibility_iteration)
        finally:
            self._free(res)

    def solve(self, c, Q=None, R=None, A=None, con_lb=None, con_ub=None, var_lb=None, var_ub=None,
              obj_const=None, params=None, primal_start=None, dual_start=None,
              fmt: str = "csr", zero_tolerance: float = 0.0) -> QpResult:
        prob = self.make_problem(c, Q, R, A, con_lb, con_ub, var_lb, var_ub, obj_const, fmt, zero_tolerance)
        try:
            if primal_start is not None or dual_start is not None:
                ps, ds = _f64(primal_start), _f64(dual_start)
                self.lib.set_start_values(prob, _ptr(ps, C.c_double), _ptr(ds, C.c_double))
            n, m = prob.contents.num_variables, prob.contents.num_constraints
            res = self.lib.solve_qp_problem(prob, None if params is None else C.byref(params))
            if not res:
                raise RuntimeError("solve_qp_problem returned NULL")
            return self.take_result(res, n, m)
        finally:
            self.free_problem(prob)


class LinCpAl(FaCp):
    """A library exporting the linearized CP-AL QP entry points (lin_cp_al.h):
    the same problem constructors as FA_CP, its own parameter and result types."""
    Params = LinCpAlParameters
    Result = LinCpAlResult
    result_free = "cp_al_result_free"
    reset_names = ("pdal_lin_cp_al_reset_rng", "lin_cp_al_oracle_reset_rng")


class PlantedCpAl(FaCp):
    """A library exporting the planted-KKT CP-AL entry points (planted_cp_al.h):
    the linearized solver's names and result type, its own parameter type."""
    Params = PlantedCpAlParameters
    Result = LinCpAlResult
    result_free = "cp_al_result_free"
    reset_names = ("pdal_planted_cp_al_reset_rng", "planted_cp_al_oracle_reset_rng")


class PlantedFaCp(FaCp):
    """A library exporting the planted-KKT FA_CP entry points (planted_fa_cp.h):
    FA_CP's names, its own parameter type, the result without the sigma fields."""
    Params = PlantedFaCpParameters
    Result = LinCpAlResult  # the same layout: FA_CP's result without the sigma fields
    result_free = "fa_cp_result_free"
    reset_names = ("pdal_planted_fa_cp_reset_rng", "planted_fa_cp_oracle_reset_rng")


def set_param(p, key: str, value) -> None:
    if key in ("iteration_limit", "eps_optimal_relative", "eps_feasible_relative", "time_sec_limit",
               "eps_feas_polish_relative", "eps_infeasible"):
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


def load_fa_cp() -> FaCp:
    """The MI355X libfa_cp.so; raises FileNotFoundError if it is not built."""
    return FaCp(LIB_DIR / "libfa_cp.so")


def load_planted_fa_cp() -> PlantedFaCp:
    """The MI355X librhr_fa_cp.so; raises FileNotFoundError if it is not built."""
    return PlantedFaCp(LIB_DIR / "librhr_fa_cp.so")


def load_planted_cp_al() -> PlantedCpAl:
    """The MI355X librhr_cp_al.so; raises FileNotFoundError if it is not built."""
    return PlantedCpAl(LIB_DIR / "librhr_cp_al.so")


def load_lin_cp_al() -> LinCpAl:
    """The MI355X librhr_lin_cp_al.so; raises FileNotFoundError if it is not built."""
    return LinCpAl(LIB_DIR / "librhr_lin_cp_al.so")
