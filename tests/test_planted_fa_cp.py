"""RHR-FA-CP, the planted-KKT FA_CP QP solver (QP/PLANTED_FA_CP):
librhr_fa_cp.so / planted_fa_cp.h.  FA_CP's iteration with equality AL
kernels (right-hand side = the constraint lower bound), the AL branch only
when al_sigma != 0, its own inner-loop rule, a fixed-point error without the
clamp at zero, and its AL sigma rule at restarts (adaptive_al_sigma; off by
default) in place of FA_CP's sigma rules.

Pins of the oracle (CPU), independent of its code: solves reach the planted
optimum of an equality-constrained convex QP (explicit and low-rank Q, AL
branch, with and without the adaptive sigma rule), of the inequality form
with the AL branch off (al_sigma = 0, or use_al_qp = false), and of an
equality LP in the plain branch.  The AL branch on an LP is left out: the
reference runs its inner solve there with a Q product it never allocates.
The GPU path is held to the oracle: iterates within 1e-10 up to the first
restart (and through restarts for the LP, which has no inner solve); after
restarts the inner solve's stopping tests make the trajectory rounding-
sensitive, so it is held to 10x the distance between the oracle and the
oracle summing in reverse; full solves to the oracle's status and objective.
"""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from pdal_amd import qp as Q
from pdal_amd.generators import planted_lp, planted_qp

ROOT = Path(__file__).resolve().parent.parent
INCLUDE = ROOT / "include"
TOL = dict(eps_optimal_relative=1e-6, eps_feasible_relative=1e-6, iteration_limit=300000)


@pytest.fixture(scope="module")
def oracle_pl():
    return Q.PlantedFaCp(ROOT / "oracle" / "_build" / "libplanted_fa_cp_oracle.so")


@pytest.fixture(scope="module")
def oracle_pl_rev():
    return Q.PlantedFaCp(ROOT / "oracle" / "_build" / "libplanted_fa_cp_oracle_rev.so")


@pytest.fixture(scope="module")
def gpu_pl():
    """The MI355X build; the test fails (no fallback) if it is missing."""
    return Q.load_planted_fa_cp()


def qp_inst():
    return planted_qp(120, 60, 30, seed=5)


def lp_args():
    lp = planted_lp(80, 160, 6, seed=3)
    return lp, dict(c=lp.c, A=lp.A, con_lb=lp.b, con_ub=lp.b, var_lb=np.zeros(lp.n), var_ub=np.full(lp.n, np.inf))


CASES = {  # name: (args factory, optimum, parameter overrides)
    "eq_explicit_al": (lambda: qp_inst().qp_args("explicit", True), lambda: qp_inst().optimal_value, {}),
    "eq_lowrank_al": (lambda: qp_inst().qp_args("lowrank", True), lambda: qp_inst().optimal_value, {}),
    "eq_explicit_adaptive": (lambda: qp_inst().qp_args("explicit", True), lambda: qp_inst().optimal_value,
                             {"adaptive_al_sigma": True}),
    "ineq_explicit_sigma0": (lambda: qp_inst().qp_args("explicit"), lambda: qp_inst().optimal_value,
                             {"al_sigma": 0.0}),
    "lp_plain": (lambda: lp_args()[1], lambda: lp_args()[0].optimal_value, {"use_al_qp": False}),
}


def solve(lib, args, **params):
    lib.reset_rng()
    return lib.solve(**args, params=lib.default_params(**params))


def rel_err(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_reaches_planted_optimum(oracle_pl, name):
    args_f, opt_f, extra = CASES[name]
    r = solve(oracle_pl, args_f(), **TOL, **extra)
    opt = opt_f()
    assert r.status == "OPTIMAL"
    assert abs(r.primal_objective - opt) <= 1e-6 * (1 + abs(opt))


def test_oracle_defaults_and_al_switch(oracle_pl):
    """Defaults of QP/PLANTED_FA_CP/src/utils.cu:471-512; al_sigma = 0 turns
    the AL branch off, which is then the plain branch exactly."""
    p = oracle_pl.default_params()
    assert p.al_sigma == 0.1 and p.use_al_qp and not p.adaptive_al_sigma
    assert (p.al_sigma_min, p.al_sigma_max) == (5e-3, 3e-1) and p.reflection_coefficient == 1.0
    args = qp_inst().qp_args("explicit")
    a = solve(oracle_pl, args, iteration_limit=400, al_sigma=0.0)
    b = solve(oracle_pl, args, iteration_limit=400, use_al_qp=False)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)


def test_oracle_adaptive_sigma_moves_the_trajectory(oracle_pl):
    args = qp_inst().qp_args("explicit", True)
    fixed = solve(oracle_pl, args, iteration_limit=1000)
    adapt = solve(oracle_pl, args, iteration_limit=1000, adaptive_al_sigma=True, adaptive_al_sigma_rule=2,
                  al_sigma_primal_trigger=1e-3, al_sigma_gap_trigger=1e-3)
    assert not np.allclose(fixed.x, adapt.x, rtol=0, atol=1e-9)


# ---------------------------------------------------------------------------
# Boundary of librhr_fa_cp.so (no GPU call)

def _exported(lib: Path) -> set:
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib)], check=True, capture_output=True, text=True)
    return {ln.split()[-1] for ln in out.stdout.splitlines() if " T " in ln}


def test_library_exports_every_declared_symbol(gpu_pl):
    text = re.sub(r"/\*.*?\*/", "", (INCLUDE / "planted_fa_cp.h").read_text(), flags=re.S)
    declared = [n for n in re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text)
                if n not in ("if", "while", "return", "sizeof")]
    exported = _exported(gpu_pl.path)
    assert "solve_qp_problem" in declared and not set(declared) - exported
    assert not [n for n in exported if "pdal" in n and not n.startswith("pdal_planted_fa_cp_")]


def test_struct_layout_matches_header(tmp_path):
    structs = {"fa_cp_parameters_t": Q.PlantedFaCpParameters, "fa_cp_result_t": Q.LinCpAlResult}
    lines = []
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    src = tmp_path / "layout.c"
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"planted_fa_cp_types.h\"\nint main(void){\n"
                   + "\n".join(lines) + "\nreturn 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{INCLUDE}", str(src), "-o", str(exe)], check=True)
    got = dict(ln.rsplit(" ", 1) for ln in subprocess.run([str(exe)], check=True, capture_output=True,
                                                          text=True).stdout.splitlines())
    for cname, cls in structs.items():
        assert int(got[f"{cname} size"]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(cls, fname).offset, (cname, fname)


def test_default_parameters_match_oracle(gpu_pl, oracle_pl):
    assert bytes(gpu_pl.default_params()) == bytes(oracle_pl.default_params())


# ---------------------------------------------------------------------------
# GPU path against the oracle

@pytest.mark.gpu
@pytest.mark.parametrize("iters", [50, 150])
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_iterates_track_oracle_to_first_restart(oracle_pl, gpu_pl, name, iters):
    args_f, _, extra = CASES[name]
    p = dict(iteration_limit=iters, eps_optimal_relative=1e-14, eps_feasible_relative=1e-14, **extra)
    print(f"[case] planted fa_cp {name} iters={iters}", flush=True)
    o = solve(oracle_pl, args_f(), **p)
    g = solve(gpu_pl, args_f(), **p)
    assert o.iterations == g.iterations == iters
    assert rel_err(g.x, o.x) <= 1e-10 and rel_err(g.y, o.y) <= 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("iters", [450, 1200])
def test_gpu_lp_tracks_oracle_through_restarts(oracle_pl, gpu_pl, iters):
    args_f, _, extra = CASES["lp_plain"]
    p = dict(iteration_limit=iters, eps_optimal_relative=1e-14, eps_feasible_relative=1e-14, **extra)
    o = solve(oracle_pl, args_f(), **p)
    g = solve(gpu_pl, args_f(), **p)
    assert rel_err(g.x, o.x) <= 1e-10 and rel_err(g.y, o.y) <= 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("iters", [600, 1200])
@pytest.mark.parametrize("name", ["eq_explicit_al", "eq_lowrank_al"])
def test_gpu_inner_solve_within_summation_noise(oracle_pl, oracle_pl_rev, gpu_pl, name, iters):
    args_f, _, extra = CASES[name]
    p = dict(iteration_limit=iters, eps_optimal_relative=1e-14, eps_feasible_relative=1e-14, **extra)
    o = solve(oracle_pl, args_f(), **p)
    r = solve(oracle_pl_rev, args_f(), **p)
    g = solve(gpu_pl, args_f(), **p)
    floor = rel_err(r.x, o.x)
    dist = rel_err(g.x, o.x)
    print(f"[noise] {name} iters={iters} oracle-vs-reversed={floor:.3e} gpu-vs-oracle={dist:.3e}")
    assert dist <= 10.0 * floor + 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_solves_match_oracle_and_planted(oracle_pl, gpu_pl, name):
    args_f, opt_f, extra = CASES[name]
    print(f"[case] planted fa_cp solve {name}", flush=True)
    o = solve(oracle_pl, args_f(), **TOL, **extra)
    g = solve(gpu_pl, args_f(), **TOL, **extra)
    opt = opt_f()
    assert g.status == o.status == "OPTIMAL"
    assert abs(g.primal_objective - o.primal_objective) <= 1e-6 * (1 + abs(o.primal_objective))
    assert abs(g.primal_objective - opt) <= 1e-6 * (1 + abs(opt))
    if name == "lp_plain":  # no inner solve: the same trajectory
        assert g.iterations == o.iterations


# ---------------------------------------------------------------------------
# The adaptive AL sigma rule at restarts (maybe_update_al_sigma_on_restart,
# QP/PLANTED_FA_CP/src/solver.cu:670-747) read off the trajectory CSV: every
# row is written at a checkpoint before a planned restart is performed, so a
# restart row's residuals and al_sigma are the rule's inputs and the next row's
# al_sigma its output.

def _csv_rows(path):
    lines = Path(path).read_text().splitlines()
    head = lines[0].split(",")
    return [dict(zip(head, (float(v) for v in line.split(",")))) for line in lines[1:]]


def _rule_output(row, rule, mode, P):
    rp, rd, rg, s = abs(row["rel_primal"]), abs(row["rel_dual"]), abs(row["rel_gap"]), row["al_sigma"]
    dg = max(rd, rg)
    up = rp > P["trig_p"] * max(dg, 1e-12) and mode != "down"
    down = dg > P["trig_g"] * max(rp, 1e-12) and mode != "up"
    ratio = rp / max(dg, 1e-12)
    if rule == "trigger":
        return min(P["max"], s * P["inc"]) if up else max(P["min"], s * P["dec"]) if down else s
    if rule == "sqrt":
        return (min(P["max"], s * min(P["inc"], np.sqrt(ratio))) if up
                else max(P["min"], s * max(P["dec"], np.sqrt(ratio))) if down else s)
    return P["max"] if up else P["min"] if down else s  # bang-bang


@pytest.mark.parametrize("rule, mode", [("trigger", "both"), ("sqrt", "both"), ("bangbang", "both"),
                                        ("trigger", "up"), ("trigger", "down")])
def test_adaptive_sigma_rule_from_the_trajectory(oracle_pl, tmp_path, rule, mode):
    csv = tmp_path / "traj.csv"
    P = dict(trig_p=2.0, trig_g=1.5, inc=1.8, dec=0.6, min=0.01, max=5.0)
    params = oracle_pl.default_params(
        adaptive_al_sigma=True, adaptive_al_sigma_rule={"trigger": 0, "sqrt": 1, "bangbang": 2}[rule],
        adaptive_al_sigma_mode={"both": 0, "up": 1, "down": 2}[mode], al_sigma=0.2, al_sigma_min=P["min"],
        al_sigma_max=P["max"], al_sigma_increase_factor=P["inc"], al_sigma_decrease_factor=P["dec"],
        al_sigma_primal_trigger=P["trig_p"], al_sigma_gap_trigger=P["trig_g"], trajectory_csv=str(csv).encode(),
        termination_evaluation_frequency=50, iteration_limit=3000, eps_optimal_relative=1e-9,
        eps_feasible_relative=1e-9)
    oracle_pl.reset_rng()
    oracle_pl.solve(**qp_inst().qp_args("explicit", True), params=params)
    rows = _csv_rows(csv)
    restarts = [k for k, r in enumerate(rows[:-1]) if r["restart"] == 1.0]
    assert len(restarts) >= 3
    moved = 0
    for k in restarts:
        want = _rule_output(rows[k], rule, mode, P)
        got = rows[k + 1]["al_sigma"]
        assert got == pytest.approx(want, rel=1e-15), (rows[k]["iteration"], rows[k])
        moved += got != rows[k]["al_sigma"]
        # between restarts the rule does not act
        assert all(r["al_sigma"] == got for r in rows[k + 1:(restarts + [len(rows)])[restarts.index(k) + 1] + 1])
    assert moved >= 1
