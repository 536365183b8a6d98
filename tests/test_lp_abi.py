"""CPU tests of the libcp_al.so boundary: exported symbols, struct layouts,
host-side problem construction, defaults and the MPS reader.  No GPU call is
made here (create/free/read/default entry points are host code)."""
from __future__ import annotations

import ctypes as C
import gzip
import re
import subprocess
import textwrap
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from pdal_amd import lp as L

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


@pytest.mark.parametrize("header", ["cp_al.h", "mps_parser.h", "pdal_lp.h", "pdal_batch.h"])
def test_library_exports_every_declared_symbol(gpu_lp, header):
    declared = _declared_functions(INCLUDE / header)
    assert declared, header
    missing = set(declared) - _exported(gpu_lp.path)
    assert not missing, f"{header}: {sorted(missing)} not exported"


def test_struct_layout_matches_header(tmp_path):
    """ctypes mirror == C layout of include/cp_al_types.h (same field order and
    types as LP/CP_AL/include/cp_al_types.h:47-189)."""
    structs = {"lp_problem_t": L.LpProblem, "cp_al_parameters_t": L.CpAlParameters,
               "cp_al_result_t": L.CpAlResult, "matrix_desc_t": L.MatrixDesc,
               "termination_criteria_t": L.TerminationCriteria, "restart_parameters_t": L.RestartParameters}
    lines = []
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    src = tmp_path / "layout.c"
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"cp_al_types.h\"\nint main(void){\n"
                   + "\n".join(lines) + "\nreturn 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{INCLUDE}", str(src), "-o", str(exe)], check=True)
    got = dict(ln.rsplit(" ", 1) for ln in subprocess.run([str(exe)], check=True, capture_output=True,
                                                          text=True).stdout.splitlines())
    for cname, cls in structs.items():
        assert int(got[f"{cname} size"]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(cls, fname).offset, (cname, fname)


def test_default_parameters_match_reference(gpu_lp, oracle_lp):
    """LP/CP_AL/src/utils.cu:361-394."""
    for lib in (gpu_lp, oracle_lp):
        p = L.CpAlParameters()
        lib.lib.set_default_parameters(C.byref(p))
        assert p.l_inf_ruiz_iterations == 10 and p.has_pock_chambolle_alpha and p.pock_chambolle_alpha == 1.0
        assert p.bound_objective_rescaling and not p.verbose and p.termination_evaluation_frequency == 200
        assert not p.feasibility_polishing and p.reflection_coefficient == 1.0
        assert p.sv_max_iter == 5000 and p.sv_tol == 1e-4
        tc = p.termination_criteria
        assert (tc.eps_optimal_relative, tc.eps_feasible_relative, tc.time_sec_limit) == (1e-4, 1e-4, 3600.0)
        assert tc.iteration_limit == 2**31 - 1 and tc.eps_feas_polish_relative == 1e-6
        rp = p.restart_params
        assert (rp.artificial_restart_threshold, rp.sufficient_reduction_for_restart,
                rp.necessary_reduction_for_restart) == (0.36, 0.2, 0.5)
        assert (rp.k_p, rp.k_i, rp.k_d, rp.i_smooth) == (0.99, 0.01, 0.0, 0.3)
        assert p.optimality_norm == 0 and p.presolve and p.matrix_zero_tol == 1e-9


def test_reflection_coefficient_env_is_clamped(gpu_lp, monkeypatch):
    monkeypatch.setenv("CP_AL_REFLECTION_COEFFICIENT", "3.5")
    p = L.CpAlParameters()
    gpu_lp.lib.set_default_parameters(C.byref(p))
    assert p.reflection_coefficient == 2.0


def _read_problem(prob):
    P = prob.contents
    m, n, nnz = P.num_constraints, P.num_variables, P.constraint_matrix_num_nonzeros
    rp = np.ctypeslib.as_array(P.constraint_matrix_row_pointers, shape=(m + 1,)).copy()
    ci = np.ctypeslib.as_array(P.constraint_matrix_col_indices, shape=(max(nnz, 1),))[:nnz].copy()
    v = np.ctypeslib.as_array(P.constraint_matrix_values, shape=(max(nnz, 1),))[:nnz].copy()
    vec = lambda p, k: np.ctypeslib.as_array(p, shape=(k,)).copy() if k else np.zeros(0)  # noqa: E731
    return dict(rp=rp, ci=ci, v=v, c=vec(P.objective_vector, n), lb=vec(P.variable_lower_bound, n),
                ub=vec(P.variable_upper_bound, n), l=vec(P.constraint_lower_bound, m),
                u=vec(P.constraint_upper_bound, m), obj_const=P.objective_constant)


@pytest.mark.parametrize("fmt", ["csr", "csc", "coo", "dense"])
def test_create_problem_formats(gpu_lp, oracle_lp, fmt):
    rng = np.random.default_rng(4)
    A = sp.random(7, 9, density=0.35, random_state=3, format="csr")
    A.data[0] = 0.0  # an explicit stored zero: kept by CSR/CSC/COO, dropped by dense (cp_al.c, utils.cu:1042)
    c = rng.standard_normal(9)
    for lib in (gpu_lp, oracle_lp):
        prob = lib.make_problem(A, c=c, fmt=fmt)
        try:
            got = _read_problem(prob)
        finally:
            lib.free_problem(prob)
        ref = sp.csr_matrix(A)
        if fmt == "dense":
            ref = sp.csr_matrix(A.toarray())
            ref.eliminate_zeros()
        elif fmt == "coo":
            coo = A.tocoo()
            ref = sp.csr_matrix((coo.data, (coo.row, coo.col)), shape=A.shape)  # keeps stored zeros
        dense_got = sp.csr_matrix((got["v"], got["ci"], got["rp"]), shape=A.shape).toarray()
        np.testing.assert_array_equal(dense_got, A.toarray())
        if fmt != "dense":
            assert got["rp"][-1] == A.nnz
        np.testing.assert_array_equal(got["c"], c)
        # NULL vectors -> defaults (cp_al.c:107-113)
        assert np.all(got["lb"] == -np.inf) and np.all(got["ub"] == np.inf)
        assert np.all(got["l"] == -np.inf) and np.all(got["u"] == np.inf)
        assert got["obj_const"] == 0.0


def test_coo_out_of_range_returns_null(gpu_lp):
    desc = L.MatrixDesc()
    desc.m, desc.n, desc.fmt = 2, 2, L.MATRIX_COO
    r = np.array([0, 5], dtype=np.int32)
    c = np.array([0, 1], dtype=np.int32)
    v = np.array([1.0, 2.0])
    desc.data.coo.nnz = 2
    desc.data.coo.row_ind = r.ctypes.data_as(C.POINTER(C.c_int))
    desc.data.coo.col_ind = c.ctypes.data_as(C.POINTER(C.c_int))
    desc.data.coo.vals = v.ctypes.data_as(C.POINTER(C.c_double))
    assert not gpu_lp.lib.create_lp_problem(None, C.byref(desc), None, None, None, None, None)


def test_invalid_eval_frequency_returns_null(gpu_lp):
    """check_params_validity (solver.cu:1757-1766): eval_freq < 3 is rejected
    before any device work."""
    A = sp.identity(3, format="csr")
    p = gpu_lp.default_params(termination_evaluation_frequency=2)
    with pytest.raises(RuntimeError):
        gpu_lp.solve(A=A, c=np.ones(3), con_lb=np.ones(3), con_ub=np.ones(3), params=p)


KAT1_MPS = textwrap.dedent("""\
    NAME          KAT1
    ROWS
     N  COST
     E  R1
     L  R2
     L  R3
    COLUMNS
        X1        COST      1.0        R1        1.0
        X1        R3        3.0
        X2        COST      1.0        R1        2.0
        X2        R2        1.0        R3        2.0
    RHS
        RHS       R1        5.0        R2        2.0
        RHS       R3        8.0
    BOUNDS
     FR BND       X1
     FR BND       X2
    ENDATA
    """)


def _mps_reader(lib):
    f = lib.lib.read_mps_file
    f.restype = C.POINTER(L.LpProblem)
    f.argtypes = [C.c_char_p]
    return f


@pytest.mark.parametrize("gz", [False, True])
def test_read_mps_kat1(gpu_lp, tmp_path, gz):
    path = tmp_path / ("kat1.mps.gz" if gz else "kat1.mps")
    if gz:
        with gzip.open(path, "wt") as fh:
            fh.write(KAT1_MPS)
    else:
        path.write_text(KAT1_MPS)
    prob = _mps_reader(gpu_lp)(str(path).encode())
    assert prob
    got = _read_problem(prob)
    gpu_lp.free_problem(prob)
    A = sp.csr_matrix((got["v"], got["ci"], got["rp"]), shape=(3, 2)).toarray()
    np.testing.assert_array_equal(A, [[1, 2], [0, 1], [3, 2]])
    np.testing.assert_array_equal(got["c"], [1, 1])
    np.testing.assert_array_equal(got["l"], [5, -np.inf, -np.inf])
    np.testing.assert_array_equal(got["u"], [5, 2, 8])
    assert np.all(got["lb"] == -np.inf) and np.all(got["ub"] == np.inf)


def test_read_mps_ranges_bounds_objsense(gpu_lp, tmp_path):
    """RANGES / BOUNDS / OBJSENSE / objective RHS rules of mps_parser.c:740-830."""
    text = textwrap.dedent("""\
        NAME T
        OBJSENSE
            MAX
        ROWS
         N  OBJ
         G  G1
         L  L1
         E  E1
         E  E2
        COLUMNS
            MARKER    'MARKER'  'INTORG'
            A         OBJ       2.0       G1        1.0
            A         L1        1.0       E1        1.0
            B         OBJ       -1.0      E2        4.0
            MARKER    'MARKER'  'INTEND'
            C         G1        1.0       UNKNOWN   9.0
        RHS
            RHS       OBJ       3.0       G1        1.0
            RHS       L1        4.0       E1        2.0
            RHS       E2        6.0
        RANGES
            RNG       G1        -2.0      L1        3.0
            RNG       E1        1.5       E2        -0.5
        BOUNDS
         UP BND       A         -1.0
         MI BND       B
         BV BND       C
        ENDATA
        """)
    path = tmp_path / "t.mps"
    path.write_text(text)
    prob = _mps_reader(gpu_lp)(str(path).encode())
    got = _read_problem(prob)
    gpu_lp.free_problem(prob)
    # a lone "MAX" line is one alphabetic token, which the reference reader
    # takes as a section header (mps_parser.c:437-470): the sense stays "min"
    np.testing.assert_array_equal(got["c"], [2.0, -1.0, 0.0])
    assert got["obj_const"] == -3.0                                 # RHS on the objective row
    np.testing.assert_array_equal(got["l"], [1.0, 1.0, 2.0, 5.5])
    np.testing.assert_array_equal(got["u"], [3.0, 4.0, 3.5, 6.0])
    np.testing.assert_array_equal(got["lb"], [0.0, -np.inf, 0.0])  # UP<0 leaves the lower bound
    np.testing.assert_array_equal(got["ub"], [-1.0, np.inf, 1.0])
    assert len(got["v"]) == 5                                       # the UNKNOWN-row entry is dropped


def test_read_mps_missing_file(gpu_lp, tmp_path):
    assert not _mps_reader(gpu_lp)(str(tmp_path / "nope.mps").encode())


def test_read_mps_objsense_max_with_second_token(gpu_lp, tmp_path):
    """OBJSENSE data line with two tokens reaches the MAX branch: objective
    and constant are negated (mps_parser.c:473-477, 514, 535-541)."""
    text = textwrap.dedent("""\
        OBJSENSE
            MAX       X
        ROWS
         N  OBJ
         E  R
        COLUMNS
            A         OBJ       2.0       R         1.0
        RHS
            RHS       OBJ       3.0       R         1.0
        ENDATA
        """)
    path = tmp_path / "max.mps"
    path.write_text(text)
    prob = _mps_reader(gpu_lp)(str(path).encode())
    got = _read_problem(prob)
    gpu_lp.free_problem(prob)
    np.testing.assert_array_equal(got["c"], [-2.0])
    assert got["obj_const"] == 3.0
