"""Benchmark of the MI355X CP_AL plain path on the large planted LP.

Workload (one GPU): m = 2e6 equality rows, n = 4e6 variables x >= 0, A with
Zipf(1.5)-clipped-to-[2,64] nonzeros per column at uniform rows, N(0,1)
values, planted optimum (pdal_amd.generators.config1, seed 1), FP64, solved to
relative KKT 1e-6 with the reflected-Halpern CP_AL iteration (alpha = 0).

A step is one full solve from the start point with the problem already
resident in HBM (device session: upload, transpose, rescaling and power
iteration happen once before the timed region and are reported separately).
`value` = solves per second over the K timed steps.

Printed line (one JSON object) also carries:
  * roofline of the dominant kernel (At.y SpMV fused with the primal update),
    timed with hipEvents on the solver's stream: algorithmic bytes per launch /
    average launch time, against 8 TB/s;
  * the CPU baseline: the FP64 oracle timed on the host on a bounded sample
    of the same instance (40 iterations), scaled to solves/s with the GPU
    solve's iteration count.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale S]
       python bench.py --kernels-only      (short run for rocprofv3 --pmc)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def fused_primal_bytes(nnz: int, n: int, m: int) -> int:
    # At.y SpMV over n rows / m columns (12 B per nonzero, 4 B per row pointer,
    # 8 B per gathered y entry counted once) + epilogue: read c, lb, ub, x0, x,
    # write xR, x  (7 x 8 B per row)
    return 12 * nnz + 4 * (n + 1) + 8 * m + 56 * n


def fused_dual_bytes(nnz: int, n: int, m: int) -> int:
    # A.xbar SpMV over m rows / n columns + epilogue: read l, u, y0, y, write y
    return 12 * nnz + 4 * (m + 1) + 8 * n + 40 * m


def cpu_baseline(inst, gpu_iterations: int, sample_iters: int = 40) -> dict:
    """Oracle (single-threaded FP64 port) on `sample_iters` iterations of the same instance."""
    from pdal_amd.lp import CpAl
    oracle_path = ROOT / "oracle" / "_build" / "libcp_al_oracle.so"
    oracle = CpAl(oracle_path)
    oracle.reset_rng()
    p = oracle.default_params(iteration_limit=sample_iters, termination_evaluation_frequency=sample_iters // 2,
                              sv_max_iter=2, eps_optimal_relative=1e-6, eps_feasible_relative=1e-6)
    t0 = time.perf_counter()
    r = oracle.solve(**inst.lp_args(), params=p)
    wall = time.perf_counter() - t0
    it_s = r.iterations / r.solve_time_sec
    return {
        "value": it_s / max(gpu_iterations, 1),
        "unit": "solves/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"{r.iterations} oracle iterations (main loop {r.solve_time_sec:.2f} s, {it_s:.3f} it/s; "
                   f"set-up + loop {wall:.1f} s) on the same instance, scaled by the GPU solve's "
                   f"{gpu_iterations} iterations"),
        "iterations_per_s": it_s,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0, help="shrink m and n of the instance (1.0 = full size)")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--kernels-only", action="store_true", help="set up and time kernels only (profiling)")
    ap.add_argument("--kernel-iters", type=int, default=20)
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        return 2

    os.environ.setdefault("PDAL_QUIET", "1")
    os.environ["CP_AL_LAMBDA"] = "0"  # plain CP_AL: no AL predictor (the benchmark protocol's alpha)

    from pdal_amd.generators import config1
    from pdal_amd.lp import LpSession, load_cp_al

    t0 = time.perf_counter()
    inst = config1(seed=1, scale=args.scale)
    m, n, nnz = inst.m, inst.n, int(inst.A.nnz)
    log(f"instance m={m} n={n} nnz={nnz} generated in {time.perf_counter() - t0:.1f} s")

    lib = load_cp_al()
    params = lib.default_params(eps_optimal_relative=args.tol, eps_feasible_relative=args.tol)
    t0 = time.perf_counter()
    sess = LpSession(lib, **{k: v for k, v in inst.lp_args().items()}, params=params)
    setup_s = sess.setup_seconds
    log(f"session set-up {setup_s:.2f} s (upload, transpose, rescale, power iteration); "
        f"wall {time.perf_counter() - t0:.2f} s")

    kt = sess.time_kernels(args.kernel_iters)
    bp = fused_primal_bytes(nnz, n, m)
    bd = fused_dual_bytes(nnz, n, m)
    gbs_p = bp / (kt["primal_ms"] * 1e-3) / 1e9
    gbs_d = bd / (kt["dual_ms"] * 1e-3) / 1e9
    log(f"kernels: primal {kt['primal_ms']:.4f} ms ({gbs_p:.0f} GB/s), dual {kt['dual_ms']:.4f} ms "
        f"({gbs_d:.0f} GB/s), eval {kt['eval_ms']:.4f} ms, lanes At={kt['lanes_at']} A={kt['lanes_a']}")
    if args.kernels_only:
        try:  # a plain copy of known size, to calibrate FETCH_SIZE / WRITE_SIZE
            import torch
            if torch.cuda.is_available():
                a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
                b = torch.empty_like(a)
                for _ in range(3):
                    b.copy_(a)
                torch.cuda.synchronize()
                log("copy calibration: 3 x 1 GiB device copies")
        except Exception as e:  # noqa: BLE001
            log(f"copy calibration skipped: {e}")
        sess.close()
        return 0

    for i in range(args.warmup):
        r = sess.solve()
        log(f"warmup solve {i}: {r.status} it={r.iterations} {r.solve_time_sec:.2f} s kkt={r.kkt:.2e}")
    times, iters, results = [], [], []
    t_all = time.perf_counter()
    for i in range(args.steps):
        t1 = time.perf_counter()
        r = sess.solve()
        dt = time.perf_counter() - t1
        times.append(dt)
        iters.append(r.iterations)
        results.append(r)
        log(f"solve {i}: {r.status} it={r.iterations} {dt:.2f} s kkt={r.kkt:.2e} "
            f"obj={r.primal_objective:.10g} planted={inst.optimal_value:.10g}")
    total = time.perf_counter() - t_all
    sess.close()

    r = results[-1]
    obj_rel = abs(r.primal_objective - inst.optimal_value) / max(abs(inst.optimal_value), 1e-300)
    mean_iters = sum(iters) / len(iters)
    it_s = sum(iters) / total
    solves_s = args.steps / total
    roof = {
        "bound": "hbm",
        "achieved": round(gbs_p, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(gbs_p / HBM_PEAK_GBS, 4),
        "traffic": None,
        "kernel": "primal_kernel (At.y SpMV + fused primal update)",
        "bytes_per_launch": bp,
        "avg_launch_ms": kt["primal_ms"],
    }
    cpu = None
    if not args.no_cpu:
        log("CPU baseline (oracle, 1 thread) ...")
        cpu = cpu_baseline(inst, int(round(mean_iters)))
        log(f"CPU baseline {cpu['iterations_per_s']:.3f} it/s")

    out = {
        "metric": "solves/sec to rel-KKT 1e-6 (plain CP_AL, single instance, config 1)",
        "value": round(solves_s, 6),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded planted LP)",
        "config": {
            "workload": "plain CP_AL, large planted LP (m=2e6, n=4e6, Zipf(1.5) column counts in [2,64])",
            "m": m, "n": n, "nnz": nnz, "tolerance": args.tol, "alpha": 0.0, "scale": args.scale,
            "eval_freq": 200,
        },
        "iterations_per_solve": mean_iters,
        "iterations_per_s": round(it_s, 1),
        "status": r.status,
        "kkt": r.kkt,
        "objective_rel_err_vs_planted": obj_rel,
        "setup_s": round(setup_s, 3),
        "spmv_gbs": {"primal_At": round(gbs_p, 1), "dual_A": round(gbs_d, 1)},
        "kernel_ms": {k: kt[k] for k in ("primal_ms", "dual_ms", "eval_ms", "iteration_ms")},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
