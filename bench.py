"""Benchmark of the MI355X CP_AL solvers (one GPU).

Default workload: the batched family of 4096 planted LPs (m_i log-uniform on
[200, 20000] with seed 12345, n_i = 2 m_i, 8 nonzeros per column at uniform
rows, N(0,1) values, instance i seeded with i; pdal_amd.generators
.config3_sizes / config3_instance), FP64, every instance solved to relative
KKT 1e-6 (alpha = 0, blocks of 64 iterations) in one batched device-resident
run.  A step is one solve of all 4096 instances with the batch already
resident in HBM (upload, transpose, rescaling and the batched power iteration
happen once before the timed region and are reported as setup_s).
`value` = instances solved per second over the K timed steps.

The same JSON line carries:
  * plain_solves_per_s: the single-instance GPU path (device session per
    instance, set-up untimed) on a random sample of the same instances;
  * roofline of the dominant kernel (the batched At.y SpMV fused with the
    primal update), timed live with hipEvents on the batch's stream:
    algorithmic bytes per launch / average launch time, against 8 TB/s;
  * spmv: the plain path's fused SpMV kernels on the large planted LP
    (m = 2e6, n = 4e6, Zipf(1.5) column counts; config1) with their HBM GB/s;
  * cpu_baseline: the FP64 oracle (single thread) on a bounded sample of the
    same instances, scaled to solves/s with the GPU iteration counts.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       python bench.py --workload config1   (plain single-instance solve of the large LP)
       python bench.py --kernels-only       (set-up and kernel timing only, for rocprofv3)
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
K_BATCH = 64


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def fused_primal_bytes(nnz: int, n: int, m: int, segments: int = 1) -> int:
    # At.y SpMV over n rows / m columns (12 B per nonzero, 4 B per row pointer,
    # 8 B per gathered y entry counted once) + epilogue: read c, lb, ub, x0, x,
    # write xR, x (7 x 8 B per row).  `segments` row-pointer arrays (one per instance).
    return 12 * nnz + 4 * (n + segments) + 8 * m + 56 * n


def fused_dual_bytes(nnz: int, n: int, m: int, segments: int = 1) -> int:
    # A.xbar SpMV over m rows / n columns + epilogue: read l, u, y0, y, write y
    return 12 * nnz + 4 * (m + segments) + 8 * n + 40 * m


def measured_traffic(tag: str):
    """HBM bytes per launch of a kernel from a committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE scaled by a copy kernel of known size); the
    summary states the commands that produced it.  None when absent."""
    f = ROOT / "profiles" / "batch_staged_spmv_4096_instances.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith(f"traffic_bytes_per_launch {tag}:"):
            return int(line.split(":")[1])
    return None


def roofline(bytes_per_launch: int, ms: float, kernel: str, traffic=None) -> dict:
    gbs = bytes_per_launch / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel,
            "bytes_per_launch": bytes_per_launch, "avg_launch_ms": ms}


def spmv_leg(lib, kernel_iters: int) -> dict:
    """Plain-path fused SpMV kernels on the large planted LP (kernel timing only)."""
    from pdal_amd.generators import config1
    from pdal_amd.lp import LpSession
    t0 = time.perf_counter()
    inst = config1(seed=1)
    m, n, nnz = inst.m, inst.n, int(inst.A.nnz)
    log(f"large LP m={m} n={n} nnz={nnz} generated in {time.perf_counter() - t0:.1f} s")
    with LpSession(lib, **inst.lp_args(), params=lib.default_params()) as sess:
        kt = sess.time_kernels(kernel_iters)
        setup = sess.setup_seconds
    bp, bd = fused_primal_bytes(nnz, n, m), fused_dual_bytes(nnz, n, m)
    out = {"workload": "large planted LP m=2e6 n=4e6 (Zipf(1.5) column counts in [2,64])", "nnz": nnz,
           "primal_At_gbs": round(bp / (kt["primal_ms"] * 1e-3) / 1e9, 1),
           "dual_A_gbs": round(bd / (kt["dual_ms"] * 1e-3) / 1e9, 1),
           "primal_ms": kt["primal_ms"], "dual_ms": kt["dual_ms"], "lanes_at": kt["lanes_at"],
           "lanes_a": kt["lanes_a"], "setup_s": round(setup, 2)}
    out["frac_of_peak"] = round(max(out["primal_At_gbs"], out["dual_A_gbs"]) / HBM_PEAK_GBS, 4)
    log(f"large LP kernels: {out}")
    return out


def cpu_baseline_batch(insts, gpu_iters: dict, sample: int, iters_cap: int) -> dict:
    """Oracle (single-threaded FP64 port): `iters_cap` iterations on each of a
    few instances, scaled to full solves with the GPU iteration counts."""
    from pdal_amd.lp import CpAl
    oracle = CpAl(ROOT / "oracle" / "_build" / "libcp_al_oracle.so")
    est = 0.0
    ids = sorted(gpu_iters)[:sample]
    done_its = 0
    t_all = time.perf_counter()
    for k in ids:
        oracle.reset_rng()
        p = oracle.default_params(iteration_limit=iters_cap, termination_evaluation_frequency=K_BATCH,
                                  eps_optimal_relative=1e-6, eps_feasible_relative=1e-6)
        r = oracle.solve(**insts[k].lp_args(), params=p)
        it_s = r.iterations / max(r.solve_time_sec, 1e-9)
        est += gpu_iters[k] / it_s
        done_its += r.iterations
    wall = time.perf_counter() - t_all
    return {"value": round(len(ids) / est, 4), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": (f"{len(ids)} instances of the family, {iters_cap} oracle iterations each ({done_its} "
                       f"iterations, {wall:.1f} s wall incl. set-up); each instance's full-solve time = its GPU "
                       f"iteration count / the oracle's iterations per second on it")}


def run_config3(args, lib) -> dict:
    import numpy as np
    from pdal_amd.generators import config3_instance, config3_sizes
    from pdal_amd.lp import LpBatch, LpSession
    count = args.instances
    t0 = time.perf_counter()
    ms = config3_sizes()[:count]
    insts = [config3_instance(i, int(ms[i])) for i in range(count)]
    log(f"{count} instances sum_m={int(ms.sum())} generated in {time.perf_counter() - t0:.1f} s")
    params = lib.default_params(termination_evaluation_frequency=K_BATCH, eps_optimal_relative=args.tol,
                                eps_feasible_relative=args.tol, iteration_limit=1000000, time_sec_limit=900.0)
    t0 = time.perf_counter()
    batch = LpBatch(lib, [i.lp_args() for i in insts], params)
    setup_s = batch.stats()["setup_s"]
    log(f"batch set-up {setup_s:.2f} s (wall {time.perf_counter() - t0:.2f} s)")
    kt = batch.time_kernels(args.kernel_iters)
    nnz, n, m = kt["nnz"], kt["n"], kt["m"]
    bp, bd = fused_primal_bytes(nnz, n, m, count), fused_dual_bytes(nnz, n, m, count)
    traffic = measured_traffic("batch_at_pass") if count == 4096 else None
    roof = roofline(bp, kt["primal_ms"], "batch At.y pass: primal_staged_kernel (LDS-staged At.y + fused primal "
                    "update), one launch per LDS class, over all instances", traffic)
    log(f"batch kernels: primal {kt['primal_ms']:.3f} ms ({roof['achieved']:.0f} GB/s), dual {kt['dual_ms']:.3f} ms "
        f"({bd / (kt['dual_ms'] * 1e-3) / 1e9:.0f} GB/s), lanes At={kt['lanes_at']} A={kt['lanes_a']}")
    result = {"roofline": roof, "setup_s": round(setup_s, 3),
              "batch_kernel_ms": {"primal": kt["primal_ms"], "dual": kt["dual_ms"]},
              "batch_spmv_gbs": {"primal_At": roof["achieved"],
                                 "dual_A": round(bd / (kt["dual_ms"] * 1e-3) / 1e9, 1)}}
    if args.kernels_only:
        batch.close()
        return result
    for w in range(args.warmup):
        batch.solve()
        log(f"warmup {w}: {batch.stats()}")
    times = []
    for s in range(args.steps):
        t1 = time.perf_counter()
        n_opt = batch.solve()
        times.append(time.perf_counter() - t1)
        st = batch.stats()
        log(f"step {s}: {n_opt}/{count} OPTIMAL in {times[-1]:.2f} s, replays={st['replays']} "
            f"compactions={st['compactions']}")
    st = batch.stats()
    its = np.array([batch.result(k).iterations for k in range(count)])
    res = [batch.result(k) for k in (0, count // 2, count - 1)]
    batch.close()
    total = sum(times)
    value = count * args.steps / total
    obj_err = max(abs(r.primal_objective - insts[k].optimal_value) / (1.0 + abs(insts[k].optimal_value))
                  for r, k in zip(res, (0, count // 2, count - 1)))
    # plain single-instance GPU path on a random sample of the same instances
    rng = np.random.default_rng(0)
    sample = rng.choice(count, size=min(args.plain_sample, count), replace=False)
    plain_t = 0.0
    for k in sample:
        with LpSession(lib, **insts[k].lp_args(), params=params) as sess:
            lib.reset_rng()
            t1 = time.perf_counter()
            r = sess.solve()
            plain_t += time.perf_counter() - t1
    plain = len(sample) / plain_t
    log(f"plain single-instance GPU: {len(sample)} solves in {plain_t:.2f} s -> {plain:.2f} solves/s")
    result.update({
        "value": value, "ms_per_step": 1e3 * total / args.steps, "optimal": st["optimal"],
        "iterations": {"mean": float(its.mean()), "p50": float(np.median(its)), "p99": float(np.percentile(its, 99)),
                       "max": int(its.max())},
        "replays": st["replays"], "compactions": st["compactions"],
        "seconds_until_half_stopped": round(st["half_stopped_s"], 2),
        "seconds_until_99pct_stopped": round(st["p99_stopped_s"], 2),
        "plain_solves_per_s": round(plain, 3), "speedup_vs_plain": round(value / plain, 2),
        "plain_sample": int(len(sample)), "objective_rel_err_vs_planted_max3": obj_err,
    })
    if not args.no_cpu:
        log("CPU baseline (oracle, 1 thread) ...")
        result["cpu_baseline"] = cpu_baseline_batch(insts, {int(k): int(its[k]) for k in range(count)},
                                                    args.cpu_sample, args.cpu_iters)
    return result


def run_config1(args, lib) -> dict:
    from pdal_amd.generators import config1
    from pdal_amd.lp import CpAl, LpSession
    inst = config1(seed=1, scale=args.scale)
    m, n, nnz = inst.m, inst.n, int(inst.A.nnz)
    params = lib.default_params(eps_optimal_relative=args.tol, eps_feasible_relative=args.tol)
    sess = LpSession(lib, **inst.lp_args(), params=params)
    kt = sess.time_kernels(args.kernel_iters)
    bp = fused_primal_bytes(nnz, n, m)
    result = {"roofline": roofline(bp, kt["primal_ms"], "primal_kernel (At.y SpMV + fused primal update)"),
              "setup_s": round(sess.setup_seconds, 3)}
    if args.kernels_only:
        sess.close()
        return result
    for _ in range(args.warmup):
        sess.solve()
    times, r = [], None
    for _ in range(args.steps):
        t1 = time.perf_counter()
        r = sess.solve()
        times.append(time.perf_counter() - t1)
        log(f"solve: {r.status} it={r.iterations} {times[-1]:.2f} s kkt={r.kkt:.2e}")
    sess.close()
    result.update({"value": args.steps / sum(times), "ms_per_step": 1e3 * sum(times) / args.steps,
                   "iterations": r.iterations, "status": r.status,
                   "objective_rel_err_vs_planted": abs(r.primal_objective - inst.optimal_value)
                   / max(abs(inst.optimal_value), 1e-300)})
    if not args.no_cpu:
        oracle = CpAl(ROOT / "oracle" / "_build" / "libcp_al_oracle.so")
        oracle.reset_rng()
        p = oracle.default_params(iteration_limit=40, termination_evaluation_frequency=20, sv_max_iter=2)
        ro = oracle.solve(**inst.lp_args(), params=p)
        it_s = ro.iterations / ro.solve_time_sec
        result["cpu_baseline"] = {"value": it_s / r.iterations, "unit": "solves/s", "cores": 1, "kind": "port",
                                  "sample": f"40 oracle iterations on the same instance, scaled by the GPU "
                                            f"solve's {r.iterations} iterations"}
    return result


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--workload", choices=["config3", "config1"], default="config3")
    ap.add_argument("--instances", type=int, default=4096)
    ap.add_argument("--scale", type=float, default=1.0, help="config1 only: shrink m and n")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--plain-sample", type=int, default=12)
    ap.add_argument("--cpu-sample", type=int, default=8)
    ap.add_argument("--cpu-iters", type=int, default=6400)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-spmv", action="store_true", help="skip the large-LP SpMV leg")
    ap.add_argument("--kernels-only", action="store_true", help="set up and time kernels only (profiling)")
    ap.add_argument("--kernel-iters", type=int, default=20)
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1", "n_gpus": args.gpus}),
              flush=True)
        return 2
    os.environ.setdefault("PDAL_QUIET", "1")
    os.environ["CP_AL_LAMBDA"] = "0"  # plain CP_AL: no AL predictor (the benchmark protocol's alpha)

    from pdal_amd.lp import load_cp_al
    lib = load_cp_al()
    if args.workload == "config3":
        res = run_config3(args, lib)
        metric = "solves/sec to rel-KKT 1e-6 (batched CP_AL, 4096 planted LPs, one GPU)"
        config = {"workload": "batched CP_AL: 4096 planted LPs, m log-uniform [200,20000], n=2m, 8 nnz/col",
                  "instances": args.instances, "tolerance": args.tol, "alpha": 0.0, "eval_freq": K_BATCH}
    else:
        res = run_config1(args, lib)
        metric = "solves/sec to rel-KKT 1e-6 (plain CP_AL, large planted LP)"
        config = {"workload": "plain CP_AL: planted LP m=2e6 n=4e6, Zipf(1.5) column counts in [2,64]",
                  "scale": args.scale, "tolerance": args.tol, "alpha": 0.0, "eval_freq": 200}
    if args.kernels_only:
        ms = lib.lib.pdal_copy_calibration(1 << 30, 3)  # 1 GiB read + 1 GiB written per launch
        log(f"kernels only: {res}; copy calibration kernel 1 GiB: {ms:.3f} ms per launch "
            f"({2 * (1 << 30) / (ms * 1e-3) / 1e9:.0f} GB/s read+write)")
        return 0
    out = {"metric": metric, "value": round(res.pop("value"), 4), "unit": "solves/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(res.pop("ms_per_step"), 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded planted LPs)", "config": config}
    roof = res.pop("roofline")
    cpu = res.pop("cpu_baseline", None)
    out.update(res)
    if args.workload == "config3" and not args.no_spmv:
        out["spmv"] = spmv_leg(lib, args.kernel_iters)
    out["roofline"] = roof
    out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
