"""Feasibility polishing of the three QP variants: the linearized CP-AL
(QP/LIN_CP_AL/src/cp_al_core_op.cu:839-1375, solver.cu:102-106), the
planted-KKT CP-AL (QP/PLANTED_CP_AL/src/cp_al_core_op.cu:1003-1539, the
linearized solver's code around its own iteration) and the planted-KKT FA_CP
(QP/PLANTED_FA_CP/src/fa_cp_core_op.cu:1009-1502, the same code as FA_CP's;
solver.cu:187-190).  All run FA_CP's two phases (primal: c = 0; dual: finite
bounds 0) with their own iteration.  The linearized
solver's phases never restart: its restart test (utils.cu:451-463) needs a
finite reference fixed-point error, which a phase starts without and only a
restart sets, so a phase is plain Halpern from its start point.

Pins of the oracle (CPU), in numpy on the unscaled problem, on the LP-type
instance of the FA_CP polishing tests after a 1e-3 main solve that does not
meet them: the returned x's primal residual, its bounds and the returned y's
reduced-cost residual, for the linearized solver's AL branch, both branches
of the planted CP-AL and the planted FA_CP's plain branch.  Without the AL branch the linearized solver's
unrestarted phases do not reach 1e-8 within the limit: both run to it and the
main solution comes back unchanged.  The GPU path is held to the oracle.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from pdal_amd import qp as Q

from test_qp_polishing import POLISH, lp_case, primal_rel, solve

ROOT = Path(__file__).resolve().parent.parent
ORACLE = ROOT / "oracle" / "_build"
CASES = {  # name: (oracle loader, GPU loader, parameter overrides)
    "lin_al": (lambda: Q.LinCpAl(ORACLE / "liblin_cp_al_oracle.so"), Q.load_lin_cp_al, {}),
    "planted_plain": (lambda: Q.PlantedFaCp(ORACLE / "libplanted_fa_cp_oracle.so"), Q.load_planted_fa_cp,
                      {"use_al_qp": False}),
    "planted_cp_al_al": (lambda: Q.PlantedCpAl(ORACLE / "libplanted_cp_al_oracle.so"), Q.load_planted_cp_al, {}),
    "planted_cp_al_plain": (lambda: Q.PlantedCpAl(ORACLE / "libplanted_cp_al_oracle.so"), Q.load_planted_cp_al,
                            {"use_al_qp": False}),
}
LIMIT = 3000


def dual_rel(args, y):
    rc = args["c"] - args["A"].T @ y
    return float(np.linalg.norm(np.minimum(rc, 0.0)) / (1.0 + np.linalg.norm(args["c"])))


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_variant_polishing_meets_tolerance(name, capfd):
    lib = CASES[name][0]()
    args = lp_case()
    p = {**POLISH, **CASES[name][2]}
    plain = solve(lib, args, **{**p, "feasibility_polishing": False})
    r = solve(lib, args, **p)
    assert "Feasibility Polishing Summary" in capfd.readouterr().out
    assert r.status == plain.status == "OPTIMAL" and r.iterations == plain.iterations
    assert r.feasibility_iteration > 0 and r.feasibility_polishing_time > 0.0
    assert primal_rel(args, plain.x) > 1e-7 and dual_rel(args, plain.y) > 1e-7  # the loose solve needed it
    assert primal_rel(args, r.x) <= 1e-8 * 1.0001
    assert np.all(r.x >= args["var_lb"]) and np.all(r.x <= args["var_ub"])
    assert dual_rel(args, r.y) <= 1e-8 * 1.0001


def test_oracle_lin_plain_phases_run_to_the_limit():
    lib = CASES["lin_al"][0]()
    args = lp_case()
    p = {**POLISH, "use_al_qp": False, "iteration_limit": LIMIT}
    plain = solve(lib, args, **{**p, "feasibility_polishing": False})
    r = solve(lib, args, **p)
    assert r.status == plain.status == "OPTIMAL"
    assert r.feasibility_iteration == 2 * LIMIT
    assert np.array_equal(r.x, plain.x) and np.array_equal(r.y, plain.y)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_variant_polishing_matches_oracle(name):
    oracle, gpu = CASES[name][0](), CASES[name][1]()
    args = lp_case()
    p = {**POLISH, **CASES[name][2]}
    print(f"[case] polishing {name}, LP-type instance", flush=True)
    o = solve(oracle, args, **p)
    g = solve(gpu, args, **p)
    assert g.status == o.status == "OPTIMAL" and g.iterations == o.iterations
    assert g.feasibility_iteration == o.feasibility_iteration > 0
    assert np.linalg.norm(g.x - o.x) <= 1e-7 * np.linalg.norm(o.x)
    assert np.linalg.norm(g.y - o.y) <= 1e-7 * np.linalg.norm(o.y)
    assert primal_rel(args, g.x) <= 1e-8 * 1.0001


@pytest.mark.gpu
def test_gpu_lin_plain_phases_run_to_the_limit():
    oracle, gpu = CASES["lin_al"][0](), CASES["lin_al"][1]()
    args = lp_case()
    p = {**POLISH, "use_al_qp": False, "iteration_limit": LIMIT}
    print("[case] polishing lin plain to the limit", flush=True)
    o = solve(oracle, args, **p)
    g = solve(gpu, args, **p)
    plain = solve(gpu, args, **{**p, "feasibility_polishing": False})
    assert g.feasibility_iteration == o.feasibility_iteration == 2 * LIMIT
    assert np.array_equal(g.x, plain.x) and np.array_equal(g.y, plain.y)
    assert np.linalg.norm(g.x - o.x) <= 1e-9 * np.linalg.norm(o.x)
