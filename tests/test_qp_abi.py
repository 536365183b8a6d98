"""CPU tests of the libfa_cp.so boundary: exported symbols, struct layouts,
defaults and host-side problem construction (QP/FA_CP/src/fa_cp.c:28-376,
utils.cu:494-537, 1135-1358).  No GPU call is made here."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from pdal_amd import qp as Q

ROOT = Path(__file__).resolve().parent.parent
INCLUDE = ROOT / "include"


def _declared_functions(header: Path) -> list[str]:
    text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    text = re.sub(r"//.*", "", text)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text)
    return [n for n in names if n not in ("if", "while", "return", "sizeof")]


def _exported(lib: Path) -> set[str]:
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib)], check=True, capture_output=True, text=True)
    return {ln.split()[-1] for ln in out.stdout.splitlines() if " T " in ln}


@pytest.mark.parametrize("header", ["fa_cp.h", "pdal_qp.h", "fa_cp_mps_parser.h"])
def test_library_exports_every_declared_symbol(gpu_qp, header):
    declared = _declared_functions(INCLUDE / header)
    assert declared, header
    missing = set(declared) - _exported(gpu_qp.path)
    assert not missing, f"{header}: {sorted(missing)} not exported"


def test_internal_symbols_are_hidden(gpu_qp):
    exported = _exported(gpu_qp.path)
    assert not [s for s in exported if "pdal" in s and not s.startswith("pdal_fa_cp_")]


def test_struct_layout_matches_header(tmp_path):
    """ctypes mirror == C layout of include/fa_cp_types.h (same field order and
    types as QP/FA_CP/include/fa_cp_types.h:41-222)."""
    structs = {"qp_problem_t": Q.QpProblem, "fa_cp_parameters_t": Q.FaCpParameters,
               "fa_cp_result_t": Q.FaCpResult, "matrix_desc_t": Q.MatrixDesc, "CsrComponent": Q.CsrComponent,
               "termination_criteria_t": Q.TerminationCriteria, "restart_parameters_t": Q.RestartParameters}
    lines = []
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    src = tmp_path / "layout.c"
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"fa_cp_types.h\"\nint main(void){\n"
                   + "\n".join(lines) + "\nreturn 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{INCLUDE}", str(src), "-o", str(exe)], check=True)
    got = dict(ln.rsplit(" ", 1) for ln in subprocess.run([str(exe)], check=True, capture_output=True,
                                                          text=True).stdout.splitlines())
    for cname, cls in structs.items():
        assert int(got[f"{cname} size"]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(cls, fname).offset, (cname, fname)


def test_default_parameters_match_reference(gpu_qp, oracle_qp):
    """QP/FA_CP/src/utils.cu:494-537."""
    for lib in (gpu_qp, oracle_qp):
        p = Q.FaCpParameters()
        lib.lib.set_default_parameters(C.byref(p))
        assert p.l_inf_ruiz_iterations == 10 and p.has_pock_chambolle_alpha and p.pock_chambolle_alpha == 1.0
        assert p.bound_objective_rescaling and p.use_al_qp and p.al_sigma == 0.0
        assert p.sigma_update_rule == 6  # guarded
        assert (p.sigma_update_gain, p.sigma_clip_low, p.sigma_clip_high) == (0.5, 0.6, 1.7)
        assert (p.sigma_min, p.sigma_max, p.sigma_rollback_tol, p.sigma_zero_threshold) == (1e-8, 10.0, 1.25, 1e-8)
        assert (p.sigma_activation_threshold, p.sigma_normal_motion_threshold, p.sigma_activation_value) == (
            2.0, 1e-10, 0.1)
        assert p.verbose == 1 and p.termination_evaluation_frequency == 200 and not p.feasibility_polishing
        assert p.reflection_coefficient == 1.0 and p.sv_max_iter == 5000 and p.sv_tol == 1e-4
        tc = p.termination_criteria
        assert (tc.eps_optimal_relative, tc.eps_feasible_relative, tc.eps_infeasible) == (1e-4, 1e-4, 1e-10)
        assert tc.time_sec_limit == 3600.0 and tc.iteration_limit == 2**31 - 1
        rp = p.restart_params
        assert (rp.artificial_restart_threshold, rp.sufficient_reduction_for_restart,
                rp.necessary_reduction_for_restart) == (0.36, 0.2, 0.8)
        assert (rp.k_p, rp.k_i, rp.k_d, rp.i_smooth) == (0.99, 0.01, 0.0, 0.3)
        assert p.optimality_norm == 1  # L_inf


def _component(comp, rows, nnz):
    rp = np.ctypeslib.as_array(comp.row_ptr, shape=(rows + 1,)).copy()
    if nnz == 0:
        return rp, np.zeros(0, np.int32), np.zeros(0)
    ci = np.ctypeslib.as_array(comp.col_ind, shape=(nnz,)).copy()
    v = np.ctypeslib.as_array(comp.val, shape=(nnz,)).copy()
    return rp, ci, v


def _read(prob):
    P = prob.contents
    n, m, r = P.num_variables, P.num_constraints, P.num_rank_lowrank_obj
    vec = lambda p, k: np.ctypeslib.as_array(p, shape=(k,)).copy() if k else np.zeros(0)  # noqa: E731
    return dict(n=n, m=m, r=r, A=_component(P.constraint_matrix.contents, m, P.constraint_matrix_num_nonzeros),
                Q=_component(P.objective_sparse_matrix.contents, n, P.objective_sparse_matrix_num_nonzeros),
                R=_component(P.objective_lowrank_matrix.contents, r, P.objective_lowrank_matrix_num_nonzeros),
                c=vec(P.objective_vector, n), lb=vec(P.variable_lower_bound, n), ub=vec(P.variable_upper_bound, n),
                l=vec(P.constraint_lower_bound, m), u=vec(P.constraint_upper_bound, m), c0=P.objective_constant)


def _csr_triplet(M):
    M = sp.csr_matrix(M)
    return M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data


@pytest.mark.parametrize("fmt", ["csr", "csc", "coo", "dense"])
def test_create_problem_formats(gpu_qp, oracle_qp, fmt):
    rng = np.random.default_rng(2)
    A = sp.random(5, 6, density=0.4, random_state=1, format="csr")
    Qm = sp.random(6, 6, density=0.3, random_state=2, format="csr")
    Qm = (Qm + Qm.T).tocsr()
    R = sp.random(2, 6, density=0.5, random_state=3, format="csr")
    A.data[0] = 1e-14  # below the dense default tolerance (1e-12), kept by CSR/CSC/COO at tolerance 0
    c = rng.standard_normal(6)
    outs = []
    for lib in (gpu_qp, oracle_qp):
        prob = lib.make_problem(c, Q=Qm, R=R, A=A, obj_const=1.5, fmt=fmt)
        try:
            outs.append(_read(prob))
        finally:
            lib.free_problem(prob)
    for got in outs:
        assert (got["n"], got["m"], got["r"]) == (6, 5, 2) and got["c0"] == 1.5
        Aref = sp.csr_matrix(A)
        if fmt == "dense":
            Aref = sp.csr_matrix(np.where(np.abs(A.toarray()) > 1e-12, A.toarray(), 0.0))
        for key, M in (("A", Aref), ("Q", Qm), ("R", R)):
            rp, ci, v = got[key]
            erp, eci, ev = _csr_triplet(M)
            np.testing.assert_array_equal(rp, erp)
            np.testing.assert_array_equal(ci, eci)
            np.testing.assert_array_equal(v, ev)
        np.testing.assert_array_equal(got["lb"], np.full(6, -np.inf))
        np.testing.assert_array_equal(got["ub"], np.full(6, np.inf))
        np.testing.assert_array_equal(got["l"], np.full(5, -np.inf))
        np.testing.assert_array_equal(got["u"], np.full(5, np.inf))
    for k in ("c", "lb", "ub", "l", "u"):
        np.testing.assert_array_equal(outs[0][k], outs[1][k])


@pytest.mark.parametrize("fmt", ["csc", "coo", "dense"])
def test_zero_tolerance_drops_small_entries(gpu_qp, fmt):
    A = np.array([[1.0, 1e-3, 0.0], [0.0, 2.0, -5e-4]])
    prob = gpu_qp.make_problem(np.zeros(3), A=sp.csr_matrix(A), fmt=fmt, zero_tolerance=1e-3)
    try:
        got = _read(prob)
    finally:
        gpu_qp.free_problem(prob)
    rp, ci, v = got["A"]
    np.testing.assert_array_equal(rp, [0, 1, 2])
    np.testing.assert_array_equal(ci, [0, 1])
    np.testing.assert_array_equal(v, [1.0, 2.0])


def test_dimensions_from_q_or_r_when_a_is_missing(gpu_qp):
    prob = gpu_qp.make_problem(np.ones(4), Q=sp.eye(4, format="csr"))
    try:
        got = _read(prob)
        assert (got["n"], got["m"], got["r"]) == (4, 0, 0)
        np.testing.assert_array_equal(got["A"][0], [0])
    finally:
        gpu_qp.free_problem(prob)
    prob = gpu_qp.make_problem(np.ones(3), R=sp.csr_matrix(np.ones((2, 3))))
    try:
        got = _read(prob)
        assert (got["n"], got["m"], got["r"]) == (3, 0, 2)
    finally:
        gpu_qp.free_problem(prob)


def test_no_matrix_is_rejected(gpu_qp):
    with pytest.raises(ValueError):
        gpu_qp.make_problem(np.ones(3))


def test_coo_out_of_range_is_rejected(gpu_qp):
    d = Q.MatrixDesc()
    d.m, d.n, d.fmt = 2, 2, Q.MATRIX_COO
    rows = np.array([0, 5], np.int32)
    cols = np.array([0, 1], np.int32)
    vals = np.array([1.0, 2.0])
    d.data.coo.nnz = 2
    d.data.coo.row_ind = rows.ctypes.data_as(C.POINTER(C.c_int))
    d.data.coo.col_ind = cols.ctypes.data_as(C.POINTER(C.c_int))
    d.data.coo.vals = vals.ctypes.data_as(C.POINTER(C.c_double))
    assert not gpu_qp.lib.create_qp_problem(None, None, None, C.byref(d), None, None, None, None, None)


def test_set_start_values_copies(gpu_qp):
    prob = gpu_qp.make_problem(np.ones(2), A=sp.csr_matrix(np.ones((1, 2))))
    try:
        x = np.array([1.0, 2.0])
        y = np.array([3.0])
        gpu_qp.lib.set_start_values(prob, x.ctypes.data_as(C.POINTER(C.c_double)),
                                    y.ctypes.data_as(C.POINTER(C.c_double)))
        x[:] = 0
        y[:] = 0
        P = prob.contents
        assert list(np.ctypeslib.as_array(P.primal_start, shape=(2,))) == [1.0, 2.0]
        assert list(np.ctypeslib.as_array(P.dual_start, shape=(1,))) == [3.0]
        gpu_qp.lib.set_start_values(prob, None, None)
        assert not P.primal_start and not P.dual_start
    finally:
        gpu_qp.free_problem(prob)


def test_solve_null_problem_returns_null(gpu_qp):
    assert not gpu_qp.lib.solve_qp_problem(None, None)
