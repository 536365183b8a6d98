"""RHR-CP-AL, the planted-KKT CP-AL QP solver (QP/PLANTED_CP_AL):
librhr_cp_al.so / planted_cp_al.h.  FA_CP's main loop, restart test and
Halpern step over the linearized solver's sigma / lambda steps, with its own
set-up (lambda = 2.001 ||A||^2, sigma = sigma_base / 2), update (equality
rows; LP and diagonal Q in closed form; FA_CP's BB inner solve with the
effective dual fixed for the other Q types; T-outputs in major-kernel
iterations only), fixed-point error on the reflection, and at restarts a
sigma controller (AL branch) or the primal-weight PID.

Pins of the oracle (CPU), independent of its code: solves reach the planted
optimum of equality-constrained QPs (explicit, low-rank and diagonal Q) and of
an equality LP, with the AL branch on and off.  The diagonal case plants its
own KKT point (c = -D x* + A'y* + lam*).  The GPU path is held to the oracle:
iterates within 1e-10 up to the first restart, and through restarts for the
LP and diagonal cases (no inner solve); after restarts the BB inner solve's
stopping tests make the trajectory rounding-sensitive, so it is held to 10x
the distance between the oracle and the oracle summing in reverse; full
solves to the oracle's status and objective and to the planted optimum.
"""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from pdal_amd import qp as Q
from pdal_amd.generators import planted_lp, planted_qp

ROOT = Path(__file__).resolve().parent.parent
INCLUDE = ROOT / "include"
TOL = dict(eps_optimal_relative=1e-7, eps_feasible_relative=1e-7, iteration_limit=300000)


@pytest.fixture(scope="module")
def oracle_pc():
    return Q.PlantedCpAl(ROOT / "oracle" / "_build" / "libplanted_cp_al_oracle.so")


@pytest.fixture(scope="module")
def oracle_pc_rev():
    return Q.PlantedCpAl(ROOT / "oracle" / "_build" / "libplanted_cp_al_oracle_rev.so")


@pytest.fixture(scope="module")
def gpu_pc():
    """The MI355X build; the test fails (no fallback) if it is missing."""
    return Q.load_planted_cp_al()


def qp_inst():
    return planted_qp(120, 60, 30, seed=5)


def lp_args():
    lp = planted_lp(80, 160, 6, seed=3)
    return lp, dict(c=lp.c, A=lp.A, con_lb=lp.b, con_ub=lp.b, var_lb=np.zeros(lp.n), var_ub=np.full(lp.n, np.inf))


def diag_case():
    """Diagonal Q with the planted point of qp_inst(): c = -D x* + A'y* + lam*
    keeps (x*, y*, lam*) a KKT point of min 1/2 x'Dx + c'x s.t. A x = A x*."""
    inst = qp_inst()
    d = np.random.default_rng(11).uniform(0.5, 2.0, inst.n)
    c = -d * inst.x_star + inst.A.T @ inst.y_star + inst.lam_star
    b = inst.A @ inst.x_star
    args = dict(c=c, Q=sp.diags(d).tocsr(), A=inst.A, con_lb=b, con_ub=b, var_lb=inst.lb, var_ub=inst.ub)
    opt = float(0.5 * inst.x_star @ (d * inst.x_star) + c @ inst.x_star)
    return args, opt


CASES = {  # name: (args factory, optimum factory, parameter overrides)
    "eq_explicit_al": (lambda: qp_inst().qp_args("explicit", True), lambda: qp_inst().optimal_value, {}),
    "eq_explicit_plain": (lambda: qp_inst().qp_args("explicit", True), lambda: qp_inst().optimal_value,
                          {"use_al_qp": False}),
    "eq_lowrank_al": (lambda: qp_inst().qp_args("lowrank", True), lambda: qp_inst().optimal_value, {}),
    "eq_lowrank_plain": (lambda: qp_inst().qp_args("lowrank", True), lambda: qp_inst().optimal_value,
                         {"use_al_qp": False}),
    "diag_al": (lambda: diag_case()[0], lambda: diag_case()[1], {}),
    "diag_plain": (lambda: diag_case()[0], lambda: diag_case()[1], {"use_al_qp": False}),
    "lp_al": (lambda: lp_args()[1], lambda: lp_args()[0].optimal_value, {}),
    "lp_plain": (lambda: lp_args()[1], lambda: lp_args()[0].optimal_value, {"use_al_qp": False}),
}
NO_INNER = ["diag_al", "diag_plain", "lp_al", "lp_plain"]
BB = ["eq_explicit_al", "eq_explicit_plain", "eq_lowrank_al", "eq_lowrank_plain"]


def solve(lib, args, **params):
    lib.reset_rng()
    return lib.solve(**args, params=lib.default_params(**params))


def rel_err(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def test_diag_case_is_planted():
    """The diagonal case's KKT point in numpy: stationarity, feasibility and
    the bound multipliers' signs."""
    inst = qp_inst()
    args, _ = diag_case()
    x, y, lam = inst.x_star, inst.y_star, inst.lam_star
    r = args["Q"] @ x + args["c"] - args["A"].T @ y - lam
    assert np.linalg.norm(r) <= 1e-9 * (1 + np.linalg.norm(args["c"]))
    assert np.allclose(args["A"] @ x, args["con_lb"])
    assert np.all(lam[x > inst.lb] <= 1e-12) and np.all(lam[x < inst.ub] >= -1e-12)


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_reaches_planted_optimum(oracle_pc, name):
    args_f, opt_f, extra = CASES[name]
    r = solve(oracle_pc, args_f(), **TOL, **extra)
    opt = opt_f()
    assert r.status == "OPTIMAL"
    assert abs(r.primal_objective - opt) <= 1e-6 * (1 + abs(opt))


def test_oracle_defaults(oracle_pc):
    """Defaults of QP/PLANTED_CP_AL/src/utils.cu:481-513."""
    p = oracle_pc.default_params()
    assert p.sigma_init == 0.0 and p.use_al_qp and p.trajectory_csv is None
    assert p.reflection_coefficient == 1.0 and p.termination_evaluation_frequency == 200


def test_oracle_sigma_init_changes_the_trajectory(oracle_pc):
    args = qp_inst().qp_args("explicit", True)
    a = solve(oracle_pc, args, iteration_limit=300)
    b = solve(oracle_pc, args, iteration_limit=300, sigma_init=1e-3)
    assert not np.allclose(a.x, b.x, rtol=0, atol=1e-9)


# ---------------------------------------------------------------------------
# Boundary of librhr_cp_al.so (no GPU call)

def _exported(lib: Path) -> set:
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib)], check=True, capture_output=True, text=True)
    return {ln.split()[-1] for ln in out.stdout.splitlines() if " T " in ln}


def test_library_exports_every_declared_symbol(gpu_pc):
    declared = []
    for h in ("planted_cp_al.h", "planted_cp_al_mps_parser.h"):
        text = re.sub(r"/\*.*?\*/", "", (INCLUDE / h).read_text(), flags=re.S)
        declared += [n for n in re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text)
                     if n not in ("if", "while", "return", "sizeof")]
    exported = _exported(gpu_pc.path)
    assert {"solve_qp_problem", "read_mps_file", "cp_al_result_free"} <= set(declared)
    assert not set(declared) - exported
    assert not [n for n in exported if "pdal" in n and not n.startswith("pdal_planted_cp_al_")]


def test_struct_layout_matches_header(tmp_path):
    structs = {"cp_al_parameters_t": Q.PlantedCpAlParameters, "cp_al_result_t": Q.LinCpAlResult}
    lines = []
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    src = tmp_path / "layout.c"
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"planted_cp_al_types.h\"\nint main(void){\n"
                   + "\n".join(lines) + "\nreturn 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{INCLUDE}", str(src), "-o", str(exe)], check=True)
    got = dict(ln.rsplit(" ", 1) for ln in subprocess.run([str(exe)], check=True, capture_output=True,
                                                          text=True).stdout.splitlines())
    for cname, cls in structs.items():
        assert int(got[f"{cname} size"]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(cls, fname).offset, (cname, fname)


def test_default_parameters_match_oracle(gpu_pc, oracle_pc):
    assert bytes(gpu_pc.default_params()) == bytes(oracle_pc.default_params())


# ---------------------------------------------------------------------------
# GPU path against the oracle

def _pair(oracle, gpu, name, iters):
    args_f, _, extra = CASES[name]
    p = dict(iteration_limit=iters, eps_optimal_relative=1e-14, eps_feasible_relative=1e-14, **extra)
    return solve(oracle, args_f(), **p), solve(gpu, args_f(), **p)


@pytest.mark.gpu
@pytest.mark.parametrize("iters", [50, 150])
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_iterates_track_oracle_to_first_restart(oracle_pc, gpu_pc, name, iters):
    print(f"[case] planted cp_al {name} iters={iters}", flush=True)
    o, g = _pair(oracle_pc, gpu_pc, name, iters)
    assert o.iterations == g.iterations == iters
    assert rel_err(g.x, o.x) <= 1e-10 and rel_err(g.y, o.y) <= 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("iters", [450, 1200])
@pytest.mark.parametrize("name", NO_INNER)
def test_gpu_tracks_oracle_through_restarts(oracle_pc, gpu_pc, name, iters):
    print(f"[case] planted cp_al {name} iters={iters}", flush=True)
    o, g = _pair(oracle_pc, gpu_pc, name, iters)
    assert rel_err(g.x, o.x) <= 1e-10 and rel_err(g.y, o.y) <= 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("iters", [600, 1200])
@pytest.mark.parametrize("name", BB)
def test_gpu_inner_solve_within_summation_noise(oracle_pc, oracle_pc_rev, gpu_pc, name, iters):
    print(f"[case] planted cp_al {name} iters={iters}", flush=True)
    o, g = _pair(oracle_pc, gpu_pc, name, iters)
    r = solve(oracle_pc_rev, CASES[name][0](), iteration_limit=iters, eps_optimal_relative=1e-14,
              eps_feasible_relative=1e-14, **CASES[name][2])
    floor, dist = rel_err(r.x, o.x), rel_err(g.x, o.x)
    print(f"[noise] {name} iters={iters} oracle-vs-reversed={floor:.3e} gpu-vs-oracle={dist:.3e}")
    assert dist <= 10.0 * floor + 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_solves_match_oracle_and_planted(oracle_pc, gpu_pc, name):
    args_f, opt_f, extra = CASES[name]
    print(f"[case] planted cp_al solve {name}", flush=True)
    o = solve(oracle_pc, args_f(), **TOL, **extra)
    g = solve(gpu_pc, args_f(), **TOL, **extra)
    opt = opt_f()
    assert g.status == o.status == "OPTIMAL"
    assert abs(g.primal_objective - o.primal_objective) <= 1e-6 * (1 + abs(o.primal_objective))
    assert abs(g.primal_objective - opt) <= 1e-6 * (1 + abs(opt))
    if name in NO_INNER:  # no inner solve: the same trajectory
        assert g.iterations == o.iterations


# ---------------------------------------------------------------------------
# The AL branch's sigma controller at restarts (perform_restart,
# QP/PLANTED_CP_AL/src/cp_al_core_op.cu:655-803) read off the trajectory CSV
# (rows written at checkpoints before a planned restart is performed: a
# restart row's fixed-point residual and sigma are the controller's inputs,
# the next row's sigma its output), and its first update recomputed from
# solves cut at the restart with every rescaling off.

def _csv_rows(path):
    lines = Path(path).read_text().splitlines()
    head = lines[0].split(",")
    return [dict(zip(head, (float(v) for v in line.split(",")))) for line in lines[1:]]


def test_sigma_controller_steps_and_rollbacks(oracle_pc, tmp_path):
    """Each restart moves sigma by a factor in [0.6, 1.7], unless the restart's
    fixed-point residual is above 1.05 times the best one so far: then sigma
    returns to the sigma of that best restart.  The primal step follows
    sigma (1 / (sigma lambda), or the nonconvexity floor when larger)."""
    csv = tmp_path / "t.csv"
    oracle_pc.reset_rng()
    p = oracle_pc.default_params(trajectory_csv=str(csv).encode(), termination_evaluation_frequency=40,
                                 iteration_limit=4000, eps_optimal_relative=1e-10, eps_feasible_relative=1e-10)
    oracle_pc.solve(**qp_inst().qp_args("explicit", True), params=p)
    rows = _csv_rows(csv)
    restarts = [k for k, r in enumerate(rows[:-1]) if r["restart"] == 1.0]
    assert len(restarts) >= 5
    lam = rows[0]["lambda_est"]
    base = 1.0 / np.sqrt(lam)
    best_fp, best_sigma, rolled, stepped = np.inf, None, 0, 0
    for k in restarts:
        old, fp, new = rows[k]["sigma"], rows[k]["fixed_point_residual"], rows[k + 1]["sigma"]
        if fp < best_fp:
            best_fp, best_sigma = fp, old
        if fp > 1.05 * best_fp:
            assert new == best_sigma, rows[k]["iteration"]
            rolled += 1
        else:
            r = new / old
            clamped = new in (pytest.approx(1e-4 * base, rel=1e-12), pytest.approx(100.0 * base, rel=1e-12))
            assert 0.6 - 1e-12 <= r <= 1.7 + 1e-12 or clamped, (rows[k]["iteration"], r)
            stepped += abs(r - 1.0) > 1e-9
    assert stepped >= 1
    for r in rows:  # pcal_tau and the step columns
        assert r["dual_step"] == r["sigma"]
        assert r["primal_step"] >= 1.0 / (r["sigma"] * r["lambda_est"]) / 2 * (1 - 1e-12)


def test_sigma_controller_first_update(oracle_pc, tmp_path):
    """sigma_1 = sigma_0 * clip(sqrt(est / sigma_0), 0.6, 1.7) with est =
    ||dy|| / sqrt(lambda ||dx||^2 - ||A dx||^2) over the first restart's
    displacement (x_T, y_T at iteration K minus the start), every rescaling
    off so the scaled space is the problem's own."""
    inst = qp_inst()
    args = inst.qp_args("explicit", True)
    K = 40
    out = []
    for limit in (0, K, K + 10):
        csv = tmp_path / f"t{limit}.csv"
        oracle_pc.reset_rng()
        p = oracle_pc.default_params(l_inf_ruiz_iterations=0, has_pock_chambolle_alpha=False,
                                     bound_objective_rescaling=False, trajectory_csv=str(csv).encode(),
                                     termination_evaluation_frequency=K, iteration_limit=limit,
                                     eps_optimal_relative=1e-14, eps_feasible_relative=1e-14)
        out.append((oracle_pc.solve(**args, params=p), _csv_rows(csv)))
    (r0, _), (rk, rows_k), (_, rows_after) = out
    restart_row = next(r for r in rows_after if r["iteration"] == K)
    assert restart_row["restart"] == 1.0
    sigma0, lam = restart_row["sigma"], restart_row["lambda_est"]
    sigma1 = next(r for r in rows_after if r["iteration"] > K)["sigma"]
    dx, dy = rk.x - r0.x, rk.y - r0.y
    metric = lam * dx @ dx - np.sum((inst.A @ dx) ** 2)
    est = np.linalg.norm(dy) / np.sqrt(metric)
    want = sigma0 * min(max(np.sqrt(est / sigma0), 0.6), 1.7)
    base = 1.0 / np.sqrt(lam)
    want = min(max(want, 1e-4 * base), 100.0 * base)
    assert sigma1 == pytest.approx(want, rel=1e-9)
    assert abs(sigma1 / sigma0 - 1.0) > 1e-3
