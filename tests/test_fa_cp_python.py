"""The FA_CP Python interface (pdal_amd.fa_cp): Model, FA_CP constants and
_fa_cp_core, against the reference package's behaviour
(QP/FA_CP/python/fa_cp/model.py, FA_CP.py; python_bindings/_core_bindings.cpp).
CPU tests cover argument handling; `gpu` tests solve."""
from __future__ import annotations

import json
import warnings
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from pdal_amd.fa_cp import FA_CP, Model
from pdal_amd.fa_cp import _fa_cp_core as core

GOLDEN = Path(__file__).resolve().parent / "golden" / "kat_qp.json"


def _f(a):
    return np.array([float(v) for v in a])


def kat(name):
    return next(c for c in json.loads(GOLDEN.read_text())["cases"] if c["name"] == name)


def test_default_params_keys_and_values():
    d = core.get_default_params()
    assert set(d) == {"verbose", "termination_evaluation_frequency", "eps_optimal_relative", "eps_feasible_relative",
                      "time_sec_limit", "iteration_limit", "l_inf_ruiz_iterations", "has_pock_chambolle_alpha",
                      "pock_chambolle_alpha", "bound_objective_rescaling", "artificial_restart_threshold",
                      "sufficient_reduction_for_restart", "necessary_reduction_for_restart", "k_p",
                      "reflection_coefficient", "feasibility_polishing", "eps_feas_polish_relative",
                      "optimality_norm", "sv_max_iter", "sv_tol"}
    assert d["optimality_norm"] == "linf" and d["verbose"] == 1 and d["termination_evaluation_frequency"] == 200
    assert d["necessary_reduction_for_restart"] == 0.8 and d["iteration_limit"] == 2**31 - 1


def test_params_parsing_and_norm_errors():
    p = core._params_from_dict({"optimality_norm": "L2", "iteration_limit": 7, "k_p": 0.5, "unknown_key": 1})
    assert p.optimality_norm == 0 and p.termination_criteria.iteration_limit == 7 and p.restart_params.k_p == 0.5
    with pytest.raises(ValueError):
        core._params_from_dict({"optimality_norm": 1})
    with pytest.raises(ValueError):
        core._params_from_dict({"optimality_norm": "l1"})


def test_aliases_and_params_view():
    c = kat("diag_q")
    m = Model(_f(c["c"]), sp.csr_matrix(np.array(c["A"])), objective_matrix=sp.csr_matrix(np.array(c["Q"])))
    m.setParams(OptimalityTol=1e-9, TimeLimit=12.0)
    m.Params.IterationLimit = 33
    m.Params["LogLevel"] = 0
    assert m.getParam("eps_optimal_relative") == 1e-9 and m.Params.time_sec_limit == 12.0
    assert m.getParam("IterationLimit") == 33 and m.Params.verbose == 0
    with pytest.raises(AttributeError):
        m.Params.NoSuchParam  # noqa: B018
    assert set(FA_CP._PARAM_ALIAS.values()) <= set(m.Params.keys())


def test_dimension_checks_and_errors():
    A = np.ones((2, 3))
    with pytest.raises(ValueError):
        Model(np.ones(2), A)  # c length
    with pytest.raises(ValueError):
        Model(np.ones(3), A, constraint_lower_bound=np.zeros(3))
    with pytest.raises(ValueError):
        Model(np.ones(3), A, objective_matrix=np.eye(2))
    m = Model(np.ones(3), A)
    with pytest.raises(TypeError):
        m.setObjectiveMatrix([[1.0]])
    assert (m.num_vars, m.num_constrs) == (3, 2)
    m = Model(np.ones(4), objective_matrix=sp.eye(4))  # n from Q, no constraints
    assert (m.num_vars, m.num_constrs) == (4, 0) and m.A is None
    empty = Model(np.ones(0))  # no matrix: left without data, as in the reference
    assert empty.num_vars == 0 and not hasattr(empty, "_params")


def test_warm_start_size_mismatch_warns():
    m = Model(np.ones(3), np.ones((2, 3)))
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        m.setWarmStart(primal=np.zeros(4), dual=np.zeros(2))
    assert any(issubclass(x.category, RuntimeWarning) for x in w)
    assert m._primal_start is None and m._dual_start is not None
    m.clearWarmStart()
    assert m._dual_start is None


def test_core_argument_errors():
    with pytest.raises(ValueError):
        core.solve_once(None, None, np.ones((2, 2)), np.ones(3))
    big = sp.csr_matrix(np.ones((1, 2)))
    big.indices = big.indices.astype(np.int64)
    big.indices[0] = 2**33
    with pytest.raises(OverflowError):
        core.solve_once(None, None, big, np.ones(2))
    with pytest.raises(ValueError):
        core.solve_once(None, None, sp.dia_matrix(np.eye(2)), np.ones(2))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["diag_q", "sparse_q_plus_low_rank"])
def test_model_solves_known_answers(name):
    c = kat(name)
    m = Model(_f(c["c"]), sp.csc_matrix(np.array(c["A"])), constraint_lower_bound=_f(c["l"]),
              constraint_upper_bound=_f(c["u"]), objective_matrix=sp.csc_matrix(np.array(c["Q"])),
              objective_matrix_low_rank=None if c["R"] is None else sp.csc_matrix(np.array(c["R"])),
              variable_lower_bound=_f(c["lb"]), variable_upper_bound=_f(c["ub"]))
    m.setParams(LogLevel=0, OptimalityTol=1e-9, FeasibilityTol=1e-9)
    m.optimize()
    assert m.Status == "OPTIMAL" and m.StatusCode == FA_CP.OPTIMAL
    np.testing.assert_allclose(m.X, _f(c["x"]), atol=1e-6)
    np.testing.assert_allclose(m.Pi, _f(c["y"]), atol=1e-6)
    assert m.ObjVal == pytest.approx(c["objective"], rel=1e-7)
    assert m.IterCount > 0 and m.Runtime >= 0.0 and m.RelPrimalResidual < 1e-9


@pytest.mark.gpu
def test_maximize_flips_linear_terms():
    """MAXIMIZE of -(1/2 x^T Q x) ... : the interface negates c and c0 only."""
    Q = sp.diags([2.0, 2.0]).tocsr()
    c = np.array([-2.0, -4.0])
    mmin = Model(c, objective_matrix=Q, variable_lower_bound=np.zeros(2), variable_upper_bound=np.full(2, 5.0),
                 objective_constant=1.0)
    mmin.setParams(LogLevel=0, OptimalityTol=1e-10, FeasibilityTol=1e-10)
    mmin.optimize()
    mmax = Model(-c, objective_matrix=Q, variable_lower_bound=np.zeros(2), variable_upper_bound=np.full(2, 5.0),
                 objective_constant=-1.0)
    mmax.ModelSense = FA_CP.MAXIMIZE
    mmax.setParams(LogLevel=0, OptimalityTol=1e-10, FeasibilityTol=1e-10)
    mmax.optimize()
    np.testing.assert_allclose(mmin.X, [1.0, 2.0], atol=1e-7)
    np.testing.assert_allclose(mmax.X, mmin.X, atol=1e-9)
    assert mmax.ObjVal == pytest.approx(-mmin.ObjVal, rel=1e-12)


@pytest.mark.gpu
def test_infeasible_status_code_and_warm_start():
    A = sp.csr_matrix(np.array([[1.0, 1.0]]))
    m = Model(np.ones(2), A, constraint_lower_bound=[-1.0], constraint_upper_bound=[-1.0],
              objective_matrix=sp.eye(2, format="csr"), variable_lower_bound=np.zeros(2),
              variable_upper_bound=np.full(2, np.inf))
    m.setParams(LogLevel=0, IterationLimit=2000)
    m.optimize()
    assert m.Status == "PRIMAL_INFEASIBLE" and m.StatusCode == FA_CP.PRIMAL_INFEASIBLE
    c = kat("diag_q")
    m = Model(_f(c["c"]), np.array(c["A"]), constraint_upper_bound=_f(c["u"]), objective_matrix=np.array(c["Q"]),
              variable_lower_bound=_f(c["lb"]))
    m.setParams(LogLevel=0, OptimalityTol=1e-9, FeasibilityTol=1e-9)
    m.optimize()
    cold = m.IterCount
    m.setWarmStart(primal=m.X, dual=m.Pi)
    m.optimize()
    assert m.Status == "OPTIMAL" and m.IterCount <= cold
