"""`Model`: the FA_CP modelling object (QP/FA_CP/python/fa_cp/model.py:83-645).

    min (or max) 1/2 x^T (Q + R^T R) x + c^T x + c0
    s.t. constraint_lower_bound <= A x <= constraint_upper_bound
         variable_lower_bound <= x <= variable_upper_bound

Same constructor, setters, parameter handling (aliases from FA_CP._PARAM_ALIAS,
the `Params` view), warm start, `optimize()` and result properties as the
reference, including its conventions:
  * dimensions come from A, else Q, else R; a model with none of them is left
    without data (the constructor returns early);
  * missing bounds stay None and reach the solver as -inf / +inf;
  * MAXIMIZE negates c and c0 (not Q) before the solve and negates ObjVal,
    DualObj, PrimalRayLinObj and DualRayObj after it;
  * a warm start of the wrong length is dropped with a RuntimeWarning.
"""
from __future__ import annotations

import warnings
from typing import Any, Optional, Union

import numpy as np
import scipy.sparse as sp

from . import FA_CP
from ._fa_cp_core import get_default_params, solve_once

ArrayLike = Union[np.ndarray, list, tuple]
_RESULT_FIELDS = ("_x", "_y", "_objval", "_dualobj", "_gap", "_rel_gap", "_status", "_status_code", "_iter",
                  "_runtime", "_rescale_time", "_rel_p_res", "_rel_d_res", "_max_p_ray", "_max_d_ray",
                  "_p_ray_lin_obj", "_d_ray_obj")


def _dense(a: ArrayLike) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _csr(M) -> sp.csr_matrix:
    out = sp.csr_matrix(M, dtype=np.float64)
    out.indptr = out.indptr.astype(np.int32, copy=False)
    out.indices = out.indices.astype(np.int32, copy=False)
    out.sort_indices()
    return out


def _check_2d(M, what: str) -> None:
    if not hasattr(M, "shape") or len(M.shape) != 2:
        raise ValueError(f"{what} must be a 2D numpy.ndarray or scipy.sparse matrix.")


class _ParamsView:
    """Attribute / item access to the model's parameters (aliases resolved)."""

    def __init__(self, model: "Model"):
        object.__setattr__(self, "_m", model)

    def __getattr__(self, name: str):
        key = FA_CP._PARAM_ALIAS.get(name, name)
        params = object.__getattribute__(self, "_m")._params
        if key in params:
            return params[key]
        raise AttributeError(f"Unknown parameter '{name}'")

    def __setattr__(self, name: str, value) -> None:
        self._m.setParam(name, value)

    def __getitem__(self, name: str):
        return getattr(self, name)

    def __setitem__(self, name: str, value) -> None:
        self._m.setParam(name, value)

    def keys(self):
        return self._m._params.keys()


class Model:
    """A QP (or LP when Q and R are absent) solved by FA_CP on the GPU."""

    def __init__(self, objective_vector: ArrayLike,
                 constraint_matrix: Optional[Union[np.ndarray, sp.spmatrix]] = None,
                 constraint_lower_bound: Optional[ArrayLike] = None,
                 constraint_upper_bound: Optional[ArrayLike] = None,
                 objective_matrix: Optional[Union[np.ndarray, sp.spmatrix]] = None,
                 objective_matrix_low_rank: Optional[Union[np.ndarray, sp.spmatrix]] = None,
                 variable_lower_bound: Optional[ArrayLike] = None,
                 variable_upper_bound: Optional[ArrayLike] = None,
                 objective_constant: float = 0.0):
        self.num_vars = 0
        self.num_constrs = 0
        if constraint_matrix is not None:
            _check_2d(constraint_matrix, "constraint_matrix")
            self.num_constrs, self.num_vars = (int(v) for v in constraint_matrix.shape)
        for M, what in ((objective_matrix, "objective_matrix"),
                        (objective_matrix_low_rank, "objective_matrix_low_rank")):
            if M is None:
                continue
            _check_2d(M, what)
            if self.num_vars == 0:
                self.num_vars = int(M.shape[1])
            elif M.shape[1] != self.num_vars:
                raise ValueError(f"{what} dimensions mismatch variables ({self.num_vars})")
        if constraint_matrix is None and objective_matrix is None and objective_matrix_low_rank is None:
            return  # nothing to size the model by
        self.ModelSense = FA_CP.MINIMIZE
        self._params: dict[str, Any] = dict(get_default_params())
        self.Params = _ParamsView(self)
        self._primal_start: Optional[np.ndarray] = None
        self._dual_start: Optional[np.ndarray] = None
        self._clear_solution_cache()
        self.setObjectiveVector(objective_vector)
        self.setObjectiveConstant(objective_constant)
        self.setObjectiveMatrix(objective_matrix)
        self.setObjectiveMatrixLowRank(objective_matrix_low_rank)
        self.setConstraintMatrix(constraint_matrix)
        self.setConstraintLowerBound(constraint_lower_bound)
        self.setConstraintUpperBound(constraint_upper_bound)
        self.setVariableLowerBound(variable_lower_bound)
        self.setVariableUpperBound(variable_upper_bound)

    # ------------------------------------------------------------------ data
    def setObjectiveVector(self, c: ArrayLike) -> None:
        c = _dense(c)
        if c.ndim != 1:
            raise ValueError(f"setObjectiveVector: c must be 1D, got shape {c.shape}")
        if c.size != self.num_vars:
            raise ValueError(f"setObjectiveVector: length {c.size} != self.num_vars ({self.num_vars})")
        self.c = c
        self._clear_solution_cache()

    def setObjectiveConstant(self, c0: float) -> None:
        self.c0 = float(c0)
        self._clear_solution_cache()

    def _matrix(self, M, who: str, letter: str):
        if not isinstance(M, (np.ndarray, sp.spmatrix)):
            raise TypeError(f"{who}: {letter} must be a numpy.ndarray or scipy.sparse matrix")
        if len(M.shape) != 2:
            raise ValueError(f"{who}: {letter} must be 2D, got shape {M.shape}")
        if M.shape[1] != self.num_vars:
            raise ValueError(f"{who}: {letter} shape {M.shape} does not match number of variables "
                             f"({self.num_vars})")
        return _csr(M) if sp.issparse(M) else _dense(M)

    def setObjectiveMatrix(self, Q_like) -> None:
        self.Q = None if Q_like is None else self._matrix(Q_like, "setObjectiveMatrix", "Q")
        self._clear_solution_cache()

    def setObjectiveMatrixLowRank(self, R_like) -> None:
        self.R = None if R_like is None else self._matrix(R_like, "setObjectiveMatrixLowRank", "R")
        self._clear_solution_cache()

    def setConstraintMatrix(self, A_like) -> None:
        if A_like is None:
            self.A = None
            self.num_constrs = 0
            self.constr_lb = self.constr_ub = None
            self._clear_solution_cache()
            return
        self.A = self._matrix(A_like, "setConstraintMatrix", "A")
        self.num_constrs = int(self.A.shape[0])
        for attr, name in (("constr_lb", "constraint_lower_bound"), ("constr_ub", "constraint_upper_bound")):
            v = getattr(self, attr, None)
            if v is not None and np.asarray(v).size != self.num_constrs:
                raise ValueError(f"setConstraintMatrix: {name} length {np.asarray(v).size} != rows "
                                 f"{self.num_constrs}. Update it with the matching setter.")
        self._clear_solution_cache()

    def _bound(self, v, who: str, size: int, count: str):
        if v is None:
            return None
        v = _dense(v).ravel()
        if v.size != size:
            raise ValueError(f"{who}: length {v.size} != self.{count} ({size})")
        return v

    def setConstraintLowerBound(self, constr_lb: Optional[ArrayLike]) -> None:
        self.constr_lb = self._bound(constr_lb, "setConstraintLowerBound", self.num_constrs, "num_constrs")
        self._clear_solution_cache()

    def setConstraintUpperBound(self, constr_ub: Optional[ArrayLike]) -> None:
        self.constr_ub = self._bound(constr_ub, "setConstraintUpperBound", self.num_constrs, "num_constrs")
        self._clear_solution_cache()

    def setVariableLowerBound(self, lb: Optional[ArrayLike]) -> None:
        self.lb = self._bound(lb, "setVariableLowerBound", self.num_vars, "num_vars")
        self._clear_solution_cache()

    def setVariableUpperBound(self, ub: Optional[ArrayLike]) -> None:
        self.ub = self._bound(ub, "setVariableUpperBound", self.num_vars, "num_vars")
        self._clear_solution_cache()

    # ------------------------------------------------------------ warm start
    def setWarmStart(self, primal: Optional[ArrayLike] = None, dual: Optional[ArrayLike] = None) -> None:
        self._primal_start = self._start(primal, self.num_vars, "primal")
        self._dual_start = self._start(dual, self.num_constrs, "dual")

    @staticmethod
    def _start(v, size: int, what: str):
        if v is None:
            return None
        v = _dense(v).ravel()
        if v.size != size:
            warnings.warn(f"Warm start {what} size mismatch (expected {size}, got {v.size}).", RuntimeWarning)
            return None
        return v

    def clearWarmStart(self) -> None:
        self.setWarmStart(None, None)

    # ------------------------------------------------------------ parameters
    def setParam(self, name: str, value: Any) -> None:
        self._params[FA_CP._PARAM_ALIAS.get(name, name)] = value

    def getParam(self, name: str) -> Any:
        return self._params.get(FA_CP._PARAM_ALIAS.get(name, name))

    def setParams(self, /, **kwargs) -> None:
        for k, v in kwargs.items():
            self.setParam(k, v)

    # ----------------------------------------------------------------- solve
    def optimize(self) -> None:
        self._clear_solution_cache()
        if self.ModelSense not in (FA_CP.MINIMIZE, FA_CP.MAXIMIZE):
            raise ValueError("model_sense must be FA_CP.MINIMIZE or FA_CP.MAXIMIZE")
        sign = 1 if self.ModelSense == FA_CP.MINIMIZE else -1
        info = solve_once(self.Q, self.R, self.A, sign * self.c, sign * self.c0, self.lb, self.ub, self.constr_lb,
                          self.constr_ub, zero_tolerance=0.0, params=self._params,
                          primal_start=self._primal_start, dual_start=self._dual_start)
        flip = lambda v: None if v is None else sign * v  # noqa: E731
        self._x = np.asarray(info["X"])
        self._y = np.asarray(info["Pi"])
        self._objval = flip(info["PrimalObj"])
        self._dualobj = flip(info["DualObj"])
        self._gap = info["ObjectiveGap"]
        self._rel_gap = info["RelativeObjectiveGap"]
        self._status = str(info["Status"])
        self._status_code = int(info["StatusCode"])
        self._iter = int(info["Iterations"])
        self._runtime = info["RuntimeSec"]
        self._rescale_time = info["RescalingTimeSec"]
        self._rel_p_res = info["RelativePrimalResidual"]
        self._rel_d_res = info["RelativeDualResidual"]
        self._max_p_ray = info["MaxPrimalRayInfeas"]
        self._max_d_ray = info["MaxDualRayInfeas"]
        self._p_ray_lin_obj = flip(info["PrimalRayLinObj"])
        self._d_ray_obj = flip(info["DualRayObj"])

    def _clear_solution_cache(self) -> None:
        for f in _RESULT_FIELDS:
            setattr(self, f, None)

    # ------------------------------------------------------------- results
    X = property(lambda self: self._x)
    Pi = property(lambda self: self._y)
    ObjVal = property(lambda self: self._objval)
    DualObj = property(lambda self: self._dualobj)
    Gap = property(lambda self: self._gap)
    RelGap = property(lambda self: self._rel_gap)
    Status = property(lambda self: self._status)
    StatusCode = property(lambda self: self._status_code)
    IterCount = property(lambda self: self._iter)
    Runtime = property(lambda self: self._runtime)
    RescalingTime = property(lambda self: self._rescale_time)
    RelPrimalResidual = property(lambda self: self._rel_p_res)
    RelDualResidual = property(lambda self: self._rel_d_res)
    MaxPrimalRayInfeas = property(lambda self: self._max_p_ray)
    MaxDualRayInfeas = property(lambda self: self._max_d_ray)
    PrimalRayLinObj = property(lambda self: self._p_ray_lin_obj)
    DualRayObj = property(lambda self: self._d_ray_obj)
    PrimalInfeas = property(lambda self: self._rel_p_res)
    DualInfeas = property(lambda self: self._rel_d_res)
