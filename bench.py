"""Benchmark of the MI355X CP_AL solvers (one GPU).

Default workload: batched + learned-warm-start solves of the fixed-A LP family
(pdal_amd.generators.config4_family: one A with m = 5e4, n = 1e5, 8 nonzeros
per column, N(0,1); b = b0 + U theta_b, c = c0 + V theta_c with 32 + 32
parameters around a planted optimum).  Before the timed region: 512 training
members solved in one batch to 1e-8 (iteration cap 40,000), the affine head +
MLP (64 -> 512 -> 512 -> 512 -> 256, BF16 MFMA, trained on the device) fitted
on their POD coefficients and final primal weights, and the batch of 256
held-out members set up in HBM (shared-A engine).  A step = predict the starts
and primal weights of the 256 held-out members (MLP forward + affine head,
POD expansion written into the batch's start buffers on the device), solve all
256 to relative KKT 1e-6 (alpha = 0, blocks of 16 iterations), cold re-solve
of any member the warm pass leaves unsolved.
`value` = held-out solves per second over the K timed steps.

The same JSON line carries:
  * plain_solves_per_s: the single-instance GPU path (device session per
    instance, set-up untimed, cold) on held-out members; speedup_vs_plain;
  * cold_batched_solves_per_s: the same batch solved from zero;
  * roofline of the dominant kernel (the shared-A At.y pass fused with the
    primal update over the 256 members), timed live with hipEvents on the
    batch's stream: algorithmic bytes per launch / average launch time,
    against 8 TB/s, with the HBM traffic measured by rocprofv3;
  * batch_4096: the batched family of 4096 uneven planted LPs (m_i
    log-uniform on [200, 20000], n_i = 2 m_i, 8 nonzeros per column; cold),
    solves/s and its own roofline;
  * spmv: the plain path's fused SpMV kernels on the large planted LP
    (m = 2e6, n = 4e6, Zipf(1.5) column counts) with their HBM GB/s, and one
    full plain solve of it to rel KKT 1e-6 (iterations/s, solves/s, the
    planted objective against the bar the KKT test implies);
  * cpu_baseline: the FP64 oracle (-O3, OpenMP over the host's threads, cold)
    on a bounded sample of held-out members, scaled to solves/s with the GPU
    cold iteration counts, with the one-thread figure, the thread count and
    the CPU model; beside it, full oracle and GPU solves of the small planted
    LP at 1e-4 and 1e-8 (cpu_config0_full_solves);
  * qp_config2: FA_CP on the planted QP (explicit and low-rank Q), one solve
    each to 1e-6 within a time cap, time and bytes per inner BB step; and a
    whole explicit-Q solve of the same family at a fifth of the size.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       python bench.py --workload config3   (the 4096-instance batch alone)
       python bench.py --workload config1   (plain single-instance solve of the large LP)
       python bench.py --kernels-only       (set-up and kernel timing only, for rocprofv3)
       python bench.py --kernels-only --solve-blocks 10   (+ ten whole cold blocks of the held-out batch)
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
# evaluation frequency of the warm-started held-out solves: a warm member needs about 80 iterations, so
# a block of 64 makes most of them run to 128; measured steps at 256 members: K = 64 / 32 / 16 / 12 / 8
# give 1364 / 1797 / 1995 / 1918 / 1690 solves/s (the evaluation pass costs about one iteration)
K_WARM = 16


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


def shared_primal_bytes(nnz: int, n: int, m: int, count: int, shared_bounds: bool) -> int:
    # shared-A At.y pass: the one At streamed once (12 B per nonzero, 4 B per
    # row pointer), every instance's y gathered once (8 m per instance), and the
    # epilogue per instance: read c, lb, ub, x0, x, write xR, x (7 x 8 B per row);
    # with the bounds held once for all instances, lb and ub are 2 x 8 B per row
    if shared_bounds:
        return 12 * nnz + 4 * (n + 1) + 8 * m * count + 40 * n * count + 16 * n
    return 12 * nnz + 4 * (n + 1) + 8 * m * count + 56 * n * count


def measured_traffic(tag: str, profile: str = "batch_staged_spmv_4096_instances.md"):
    """HBM bytes per launch of a kernel from a committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE scaled by a copy kernel of known size); the
    summary states the commands that produced it.  None when absent."""
    f = ROOT / "profiles" / profile
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith(f"traffic_bytes_per_launch {tag}:"):
            return int(line.split(":")[1])
    return None


def planted_bar_ratio(r, fam, theta, tol: float) -> float:
    """|p - planted optimum| / the bar the KKT test at tol implies for member theta."""
    from pdal_amd.generators import kkt_objective_bar
    xs, ys, _ = fam.solution(theta)
    inst = fam.instance(theta)
    bar = kkt_objective_bar(r.primal_objective, r.dual_objective, xs, ys, inst["con_lb"], inst["c"], tol)
    return abs(r.primal_objective - fam.optimal_value(theta)) / bar


def roofline(bytes_per_launch: int, ms: float, kernel: str, traffic=None) -> dict:
    gbs = bytes_per_launch / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel,
            "bytes_per_launch": bytes_per_launch, "avg_launch_ms": ms}


def spmv_leg(lib, kernel_iters: int, tol: float = 1e-6, full_solve: bool = True) -> dict:
    """Plain-path fused SpMV kernels on the large planted LP, and (full_solve)
    one plain single-instance solve of it to rel KKT tol on the same device
    session: iterations/s and solves/s of the plain path at full size."""
    from pdal_amd.generators import config1, kkt_objective_bar
    from pdal_amd.lp import LpSession
    t0 = time.perf_counter()
    inst = config1(seed=1)
    m, n, nnz = inst.m, inst.n, int(inst.A.nnz)
    log(f"large LP m={m} n={n} nnz={nnz} generated in {time.perf_counter() - t0:.1f} s")
    solve = None
    with LpSession(lib, **inst.lp_args(), params=lib.default_params(eps_optimal_relative=tol,
                                                                    eps_feasible_relative=tol)) as sess:
        kt = sess.time_kernels(kernel_iters)
        setup = sess.setup_seconds
        if full_solve:
            t1 = time.perf_counter()
            r = sess.solve()
            wall = time.perf_counter() - t1
            opt = inst.optimal_value
            bar = kkt_objective_bar(r.primal_objective, r.dual_objective, inst.x_star, inst.y_star, inst.b, inst.c,
                                    tol)
            solve = {"status": r.status, "iterations": r.iterations, "solve_s": round(r.solve_time_sec, 3),
                     "wall_s": round(wall, 3), "iterations_per_s": round(r.iterations / r.solve_time_sec, 1),
                     "solves_per_s": round(1.0 / wall, 5), "kkt": r.kkt, "tolerance": tol,
                     "objective_rel_err_vs_planted": abs(r.primal_objective - opt) / abs(opt),
                     "planted_objective_over_kkt_bar": round(abs(r.primal_objective - opt) / bar, 4)}
            log(f"large LP full solve: {solve}")
    bp, bd = fused_primal_bytes(nnz, n, m), fused_dual_bytes(nnz, n, m)
    out = {"workload": "large planted LP m=2e6 n=4e6 (Zipf(1.5) column counts in [2,64])", "nnz": nnz,
           "primal_At_gbs": round(bp / (kt["primal_ms"] * 1e-3) / 1e9, 1),
           "dual_A_gbs": round(bd / (kt["dual_ms"] * 1e-3) / 1e9, 1),
           "primal_ms": kt["primal_ms"], "dual_ms": kt["dual_ms"], "lanes_at": kt["lanes_at"],
           "lanes_a": kt["lanes_a"], "setup_s": round(setup, 2)}
    out["frac_of_peak"] = round(max(out["primal_At_gbs"], out["dual_A_gbs"]) / HBM_PEAK_GBS, 4)
    # the floor of a pass with uniform column indices: the same count of random FP64 gathers alone
    g_at = lib.lib.pdal_gather_calibration(nnz, m, 5)  # At.y gathers y (m entries)
    g_a = lib.lib.pdal_gather_calibration(nnz, n, 5)   # A.x_bar gathers x_bar (n entries)
    if not (g_at > 0 and g_a > 0):
        raise RuntimeError(f"pdal_gather_calibration failed ({g_at}, {g_a})")
    out["gather_floor_ms"] = {"primal_At": round(g_at, 4), "dual_A": round(g_a, 4)}
    out["spmv_over_gather_floor"] = {"primal_At": round(kt["primal_ms"] / g_at, 3), "dual_A": round(kt["dual_ms"] / g_a, 3)}
    log(f"large LP kernels: {out}")
    if solve is not None:
        out["full_solve"] = solve
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
    from pdal_amd.generators import config3_instance, config3_sizes, kkt_objective_bar
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
    bar_ratio = max(abs(r.primal_objective - insts[k].optimal_value)
                    / kkt_objective_bar(r.primal_objective, r.dual_objective, insts[k].x_star, insts[k].y_star,
                                        insts[k].b, insts[k].c, args.tol)
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
        "planted_objective_over_kkt_bar_max3": round(bar_ratio, 4),
    })
    if not args.no_cpu:
        log("CPU baseline (oracle, 1 thread) ...")
        result["cpu_baseline"] = cpu_baseline_batch(insts, {int(k): int(its[k]) for k in range(count)},
                                                    args.cpu_sample, args.cpu_iters)
    return result


def host_cpu() -> dict:
    """Threads this process may use and the CPU model of the host."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"threads": cores, "model": model}


def cpu_baseline_family_parallel(insts, gpu_iters: list, iters_cap: int, threads: int, k_eval: int = K_BATCH) -> dict:
    """Oracle (FP64 port, -O3, OpenMP) on the host: the first held-out member
    on one thread, and `threads` members at once on `threads` threads, each
    capped at `iters_cap` iterations; each member's full-solve time is its GPU
    cold iteration count over the measured iterations per second on it."""
    from pdal_amd.lp import CpAl
    oracle = CpAl(ROOT / "oracle" / "_build" / "libcp_al_oracle_omp.so")
    p = oracle.default_params(iteration_limit=iters_cap, termination_evaluation_frequency=k_eval,
                              eps_optimal_relative=1e-6, eps_feasible_relative=1e-6)
    # iterations per second from each solve's own clock (started after its set-up)
    t0 = time.perf_counter()
    r1 = oracle.solve_many([insts[0]], p, 1)[0]
    one_s = time.perf_counter() - t0
    single = 1.0 / (gpu_iters[0] * r1.solve_time_sec / r1.iterations)
    k = min(threads, len(insts))
    t0 = time.perf_counter()
    rs = oracle.solve_many(insts[:k], p, k)
    wall = time.perf_counter() - t0
    # k members run side by side: the k-thread rate is k over a member's mean full-solve time
    full = sum(gpu_iters[i] * r.solve_time_sec / r.iterations for i, r in enumerate(rs)) / k
    host = host_cpu()
    log(f"CPU baseline: 1 thread {r1.iterations} it in {one_s:.1f} s; {k} threads {sum(r.iterations for r in rs)} it "
        f"in {wall:.1f} s wall ({host})")
    return {"value": round(k / full, 5), "unit": "solves/s", "cores": k, "kind": "port",
            "single_thread_value": round(single, 5), "cpu_model": host["model"], "host_threads": host["threads"],
            "sample": (f"oracle -O3 with OpenMP: {k} held-out members on {k} threads at once, {iters_cap} cold "
                       f"iterations each ({wall:.1f} s wall incl. set-up), and member 0 alone on one thread "
                       f"({one_s:.1f} s); each member's full-solve time = its GPU cold iteration count x the "
                       f"oracle's measured time per iteration on it (set-up excluded)")}


def cpu_config0_full_solves(gpu_lib, threads: int, seeds: int = 8) -> dict:
    """Full solves of the small planted LP (m = 1000, n = 2000, 10 nonzeros
    per column) at 1e-4 and 1e-8: the oracle on one thread and on `threads`
    threads, and the plain single-instance GPU path, on the same instances."""
    import numpy as np
    from pdal_amd.generators import config0
    from pdal_amd.lp import CpAl
    oracle = CpAl(ROOT / "oracle" / "_build" / "libcp_al_oracle_omp.so")
    insts = [config0(seed=s).lp_args() for s in range(seeds)]
    out = {"workload": f"planted LP m=1000 n=2000, 10 nnz/col, seeds 0..{seeds - 1}, full solves"}
    for tol in (1e-4, 1e-8):
        p = oracle.default_params(eps_optimal_relative=tol, eps_feasible_relative=tol)
        t0 = time.perf_counter()
        one = oracle.solve_many(insts[:2], p, 1)
        t_one = time.perf_counter() - t0
        k = min(threads, len(insts))
        t0 = time.perf_counter()
        many = oracle.solve_many(insts[:k], p, k)
        t_many = time.perf_counter() - t0
        gp = gpu_lib.default_params(eps_optimal_relative=tol, eps_feasible_relative=tol)
        t0 = time.perf_counter()
        g = []
        for kw in insts[:k]:
            gpu_lib.reset_rng()
            g.append(gpu_lib.solve(**kw, params=gp))
        t_gpu = time.perf_counter() - t0
        same = all(a.iterations == b.iterations and a.status == b.status == "OPTIMAL" for a, b in zip(many, g))
        out[f"tol_{tol:g}"] = {"cpu_1thread_solves_per_s": round(len(one) / t_one, 3),
                               f"cpu_{k}threads_solves_per_s": round(k / t_many, 3),
                               "gpu_plain_solves_per_s": round(k / t_gpu, 3),
                               "iterations_mean": float(np.mean([r.iterations for r in many])),
                               "gpu_equals_oracle_iterations": same}
        log(f"config0 tol {tol:g}: {out[f'tol_{tol:g}']}")
    return out


def cpu_baseline_family(insts, gpu_iters: list, sample: int, iters_cap: int) -> dict:
    """Oracle (single-threaded FP64 port of the reference's CP_AL, cold start):
    `iters_cap` iterations on each of a few held-out instances, scaled to full
    solves with the GPU's cold iteration counts of the same instances."""
    from pdal_amd.lp import CpAl
    oracle = CpAl(ROOT / "oracle" / "_build" / "libcp_al_oracle.so")
    est, done_its = 0.0, 0
    t_all = time.perf_counter()
    for k in range(sample):
        oracle.reset_rng()
        p = oracle.default_params(iteration_limit=iters_cap, termination_evaluation_frequency=K_BATCH,
                                  eps_optimal_relative=1e-6, eps_feasible_relative=1e-6)
        r = oracle.solve(**insts[k], params=p)
        est += gpu_iters[k] / (r.iterations / max(r.solve_time_sec, 1e-9))
        done_its += r.iterations
        log(f"CPU baseline instance {k}: {r.iterations} iterations in {r.solve_time_sec:.1f} s")
    wall = time.perf_counter() - t_all
    return {"value": round(sample / est, 5), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": (f"{sample} held-out instances, {iters_cap} cold oracle iterations each ({done_its} iterations, "
                       f"{wall:.1f} s wall incl. set-up); each instance's full-solve time = its GPU cold iteration "
                       f"count / the oracle's iterations per second on it")}


def run_config4(args, lib) -> dict:
    """Batched + learned-warm-start solves of held-out instances of the fixed-A
    family (m=5e4, n=1e5) against the plain single-instance GPU path."""
    import numpy as np
    from pdal_amd.generators import config4_family, config4_thetas
    from pdal_amd.lp import LpBatch, LpSession
    from pdal_amd.warmstart import LearnedWarmStart, load_ws
    count = args.heldout
    t0 = time.perf_counter()
    fam = config4_family(seed=4)
    te = config4_thetas(count, 2)
    test_inst = [fam.instance(t) for t in te]
    log(f"family m={fam.m} n={fam.n} nnz={fam.base.A.nnz}; {count} held-out instances in {time.perf_counter() - t0:.1f} s")
    params = lib.default_params(termination_evaluation_frequency=args.eval_freq, eps_optimal_relative=args.tol,
                                eps_feasible_relative=args.tol, iteration_limit=1000000, time_sec_limit=300.0)
    t0 = time.perf_counter()
    batch = LpBatch(lib, test_inst, params)
    st = batch.stats()
    if not st["shared_a"]:
        raise RuntimeError("the held-out batch did not take the shared-A engine")
    setup_s = st["setup_s"]
    log(f"held-out batch set-up {setup_s:.2f} s (wall {time.perf_counter() - t0:.2f} s)")
    kt = batch.time_kernels(args.kernel_iters)
    nnz1, n1, m1 = kt["nnz"] // count, kt["n"] // count, kt["m"] // count
    bp = shared_primal_bytes(nnz1, n1, m1, count, st["shared_bounds"])
    traffic = measured_traffic("shared_at_pass", "shared_spmv_fixed_a_family_256_instances.md") \
        if count == 256 and st["shared_bounds"] else None
    roof = roofline(bp, kt["primal_ms"], "shared-A At.y pass: sh_primal_kernel<4> (one At, 256 instances side by "
                    "side, bounds shared, fused primal update)", traffic)
    log(f"shared kernels: primal {kt['primal_ms']:.3f} ms ({roof['achieved']:.0f} GB/s), dual {kt['dual_ms']:.3f} ms")
    result = {"roofline": roof, "setup_s": round(setup_s, 3), "shared_bounds": st["shared_bounds"],
              "shared_kernel_ms": {"primal": kt["primal_ms"], "dual": kt["dual_ms"]}}
    if args.kernels_only:
        if args.solve_blocks > 0:  # a fixed number of whole cold blocks of the held-out batch (profiling)
            with LpBatch(lib, test_inst, lib.default_params(
                    termination_evaluation_frequency=args.eval_freq, eps_optimal_relative=args.tol,
                    eps_feasible_relative=args.tol, iteration_limit=args.solve_blocks * args.eval_freq)) as blk:
                blk.solve()
                log(f"{args.solve_blocks} cold blocks of {args.eval_freq} iterations, {count} members: "
                    f"{blk.stats()['solve_s'] * 1e3:.1f} ms")
        batch.close()
        return result
    # training: the solver's own solutions of other family members (outside the timed region)
    tr = config4_thetas(args.train, 1)
    p_train = lib.default_params(termination_evaluation_frequency=K_BATCH, eps_optimal_relative=1e-8,
                                 eps_feasible_relative=1e-8, iteration_limit=args.train_iters, time_sec_limit=300.0)
    t0 = time.perf_counter()
    with LpBatch(lib, [fam.instance(t) for t in tr], p_train) as b:
        b.solve()
        tst = b.stats()
        sols = [b.result(k) for k in range(len(tr))]
        pw = b.primal_weights()
    X, Y = np.stack([r.x for r in sols]), np.stack([r.y for r in sols])
    del sols
    log(f"training: {len(tr)} solves to 1e-8 (cap {args.train_iters} iterations) in {tst['solve_s']:.1f} s "
        f"(wall {time.perf_counter() - t0:.1f} s incl. set-up)")
    ws = load_ws()
    lws = LearnedWarmStart(ws, kx=128, seed=3)
    t0 = time.perf_counter()
    info = lws.fit(tr, X, Y, pw=pw, lb=np.zeros(fam.n), epochs=args.epochs)
    fit_s = time.perf_counter() - t0
    log(f"fit {fit_s:.1f} s: {info}")
    # cold batched solve of the held-out set (reference point, same batch)
    t1 = time.perf_counter()
    batch.solve()
    cold_s = time.perf_counter() - t1
    cold_its = np.array([batch.result(k).iterations for k in range(count)])
    log(f"cold batched: {count} in {cold_s:.2f} s, iterations mean {cold_its.mean():.0f} max {cold_its.max()}")

    # Every timed step solves a fresh batch of held-out members: the next
    # `count` parameter vectors of the pool.  Their b(theta), c(theta) are
    # input data, made on the host and put in device memory before the timed
    # region; the step replaces the batch's vectors from there
    # (pdal_cp_al_batch_set_vectors_device: the scaling passes replayed, A's
    # scaling and norm estimate kept), predicts the starts and solves.
    from pdal_amd.devbuf import DeviceArray
    n_batches = args.warmup + args.steps
    pool = config4_thetas(count * (1 + n_batches), 2)[count:]  # the first `count` rows are te
    p = fam.p
    batches = []
    t0 = time.perf_counter()
    for s in range(n_batches):
        th = pool[s * count:(s + 1) * count]
        for t in th:
            fam.solution(t)  # the planted-region check of every member
        bvec = DeviceArray.from_host((fam.base.b[:, None] + fam.U @ th[:, :p].T).T)
        cvec = DeviceArray.from_host((fam.base.c[:, None] + fam.V @ th[:, p:].T).T)
        batches.append((th, bvec, cvec))
    log(f"{n_batches} fresh batches of {count} members in device memory ({time.perf_counter() - t0:.1f} s)")

    def warm_step(th, bvec, cvec):
        t_up = time.perf_counter()
        batch.set_vectors_device(c=cvec, con_lb=bvec, con_ub=bvec)
        t_up = time.perf_counter() - t_up
        lws.predict_into(batch, th)  # MLP + head, POD expansion into the batch's start buffers on the device
        n_opt = batch.solve()
        if n_opt < count:  # the cold fallback for what the warm pass left unsolved
            bad = [k for k in range(count) if batch.result(k).status != "OPTIMAL"]
            with LpBatch(lib, [fam.instance(th[k]) for k in bad], params) as fb:
                fb.solve()
            return len(bad), t_up
        return 0, t_up

    for w in range(args.warmup):
        warm_step(*batches[w])
    times, fallbacks, splits, updates, obj_err, warm_its, bar_ratio = [], 0, [], [], 0.0, [], 0.0
    for s in range(args.steps):
        th, bvec, cvec = batches[args.warmup + s]
        t1 = time.perf_counter()
        fb, t_up = warm_step(th, bvec, cvec)
        times.append(time.perf_counter() - t1)
        fallbacks += fb
        updates.append(t_up)
        st = batch.stats()
        splits.append((st["solve_s"], st["half_stopped_s"], st["p99_stopped_s"], st["handoffs"]))
        res = [batch.result(k) for k in range(count)]
        warm_its += [r.iterations for r in res]
        obj_err = max(obj_err, max(abs(r.primal_objective - fam.optimal_value(t)) / (1.0 + abs(fam.optimal_value(t)))
                                   for r, t in zip(res, th) if r.status == "OPTIMAL"))
        bar_ratio = max(bar_ratio, max(planted_bar_ratio(r, fam, t, args.tol) for r, t in zip(res, th)
                                       if r.status == "OPTIMAL"))
        log(f"warm step {s}: {count} fresh held-out solves in {times[-1]:.3f} s (vector update {t_up * 1e3:.1f} ms, "
            f"solve {st['solve_s']:.3f} s, half stopped {st['half_stopped_s']:.3f} s, "
            f"99 % {st['p99_stopped_s']:.3f} s, fallbacks {fb})")
    warm_its = np.array(warm_its)
    for _, bvec, cvec in batches:
        bvec.close()
        cvec.close()
    batch.close()
    lws.close()
    total = sum(times)
    value = count * args.steps / total
    # plain single-instance GPU path, cold, on held-out instances (set-up untimed)
    plain_t, plain_its = 0.0, []
    for k in range(args.plain_sample):
        with LpSession(lib, **test_inst[k], params=params) as sess:
            lib.reset_rng()
            t1 = time.perf_counter()
            r = sess.solve()
            plain_t += time.perf_counter() - t1
            plain_its.append(r.iterations)
    plain = args.plain_sample / plain_t
    log(f"plain single-instance GPU: {args.plain_sample} solves in {plain_t:.2f} s -> {plain:.2f} solves/s")
    plain_k64 = None
    if args.eval_freq != K_BATCH:  # the plain path at the evaluation frequency that suits it best
        p64 = lib.default_params(termination_evaluation_frequency=K_BATCH, eps_optimal_relative=args.tol,
                                 eps_feasible_relative=args.tol, iteration_limit=1000000, time_sec_limit=300.0)
        t64 = 0.0
        for k in range(args.plain_sample):
            with LpSession(lib, **test_inst[k], params=p64) as sess:
                lib.reset_rng()
                t1 = time.perf_counter()
                sess.solve()
                t64 += time.perf_counter() - t1
        plain_k64 = args.plain_sample / t64
        log(f"plain single-instance GPU at K={K_BATCH}: {plain_k64:.2f} solves/s")
    result.update({
        "value": value, "ms_per_step": 1e3 * total / args.steps, "optimal": int(sum(r.status == "OPTIMAL" for r in res)),
        "fallbacks": fallbacks, "warm_iterations": {"mean": float(warm_its.mean()), "max": int(warm_its.max())},
        "fresh_batches": {"members": count * args.steps, "vector_update_ms": round(1e3 * float(np.mean(updates)), 2),
                          "full_setup_s": round(setup_s, 3)},
        "warm_solve_split_s": {k: round(float(np.mean([sp[i] for sp in splits])), 4)
                               for i, k in enumerate(("solve", "half_stopped", "p99_stopped", "handoffs"))},
        "cold_batched_solves_per_s": round(count / cold_s, 3),
        "cold_iterations": {"mean": float(cold_its.mean()), "max": int(cold_its.max())},
        "plain_solves_per_s": round(plain, 3), "plain_sample": args.plain_sample,
        "plain_iterations_mean": float(np.mean(plain_its)),
        "speedup_vs_plain": round(value / plain, 2), "speedup_cold_batched_vs_plain": round(count / cold_s / plain, 2),
        "plain_solves_per_s_eval64": None if plain_k64 is None else round(plain_k64, 3),
        "speedup_vs_plain_eval64": None if plain_k64 is None else round(value / plain_k64, 2),
        "objective_rel_err_vs_planted_max": obj_err,
        # |p - planted optimum| over the bar the KKT test at tol implies (generators.kkt_objective_bar); <= 1 passes
        "planted_objective_over_kkt_bar_max": round(bar_ratio, 4),
        "training": {"instances": args.train, "tolerance": 1e-8, "iteration_cap": args.train_iters,
                     "solve_s": round(tst["solve_s"], 2), "fit_s": round(fit_s, 2), "epochs": args.epochs,
                     "head_rel_resid": info.get("head_rel_resid")},
    })
    if not args.no_cpu:
        threads = min(host_cpu()["threads"], args.cpu_threads)
        log(f"CPU baseline (oracle, 1 and {threads} threads) ...")
        result["cpu_baseline"] = cpu_baseline_family_parallel(test_inst, [int(i) for i in cold_its],
                                                              args.cpu_family_iters, threads, args.eval_freq)
        result["cpu_config0_full_solves"] = cpu_config0_full_solves(lib, threads)
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
    label = ("primal_sliced_kernel (column-sliced At.y + fused primal update)" if kt["lanes_at"] < 0
             else f"primal_kernel<{kt['lanes_at']}> (At.y SpMV + fused primal update)")
    traffic = measured_traffic("primal_sliced_kernel", "cp_al_config1_fused_kernels.md") if kt["lanes_at"] < 0 \
        else None
    result = {"roofline": roofline(bp, kt["primal_ms"], label, traffic), "setup_s": round(sess.setup_seconds, 3)}
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


def bb_step_bytes(inst, form: str) -> int:
    """Algorithmic bytes of one projected-BB inner step of FA_CP on the
    config-2 QP: the Q product (explicit Q: 12 B per nonzero, 4 B per row
    pointer, the gathered x once; low-rank form: the diagonal part, R.x and
    R^T.(R x)), the gradient refresh (reads c, A^T u, x_T, x, grad, dir, writes
    grad: 56 n) and the direction update (reads x_T, grad, lb, ub, writes dir,
    x_T: 48 n)."""
    n = inst.n
    vec = 56 * n + 48 * n
    if form == "explicit":
        nnz = int(inst.q_explicit().nnz)
        return 12 * nnz + 4 * (n + 1) + 8 * n + vec
    r, nnz_r = inst.M.shape[0], int(inst.M.nnz)
    diag = 12 * n + 4 * (n + 1) + 8 * n + 8 * n         # Q_diag . x_T into qx
    rx = 12 * nnz_r + 4 * (r + 1) + 8 * n + 8 * r       # R . x_T
    rtrx = 12 * nnz_r + 4 * (n + 1) + 8 * r + 8 * n     # R^T . (R x_T), + qx read
    return diag + rx + rtrx + vec


def run_config2_leg(args) -> dict:
    """FA_CP on the planted QP (n = 1e5, m = 5e4, Q = M^T M + 1e-3 I with M
    2e4 x n, 5 nonzeros per column; A 8 per column; box [0, 10]), in the
    explicit-Q and the low-rank (R = M) form: one solve each to rel KKT tol,
    capped at qp_time seconds; per inner BB step, wall time and algorithmic
    bytes (a whole-step roofline: every kernel of the step is in the time)."""
    from pdal_amd.generators import config2
    from pdal_amd.qp import load_fa_cp
    qlib = load_fa_cp()
    inst = config2(seed=2)
    out = {"workload": "planted QP n=1e5 m=5e4, Q=M^T M+1e-3 I (M 2e4 x n, 5 nnz/col), A 8 nnz/col, box [0,10]",
           "tolerance": args.tol, "time_cap_s": args.qp_time}
    objs = {}
    for form in ("explicit", "lowrank"):
        qargs = inst.qp_args(form)
        p = qlib.default_params(eps_optimal_relative=args.tol, eps_feasible_relative=args.tol, verbose=0,
                                time_sec_limit=args.qp_time, iteration_limit=args.qp_iters)
        qlib.reset_rng()
        t0 = time.perf_counter()
        r = qlib.solve(**qargs, params=p)
        wall = time.perf_counter() - t0
        us_step = 1e6 * r.solve_time_sec / max(r.inner_iterations, 1)
        bts = bb_step_bytes(inst, form)
        gbs = bts / (us_step * 1e-6) / 1e9
        objs[form] = r.primal_objective
        out[form] = {"status": r.status, "iterations": r.iterations, "inner_steps": r.inner_iterations,
                     "solve_s": round(r.solve_time_sec, 3), "wall_s": round(wall, 3), "kkt": r.kkt,
                     "us_per_inner_step": round(us_step, 3), "bytes_per_inner_step": bts,
                     "step_gbs": round(gbs, 1), "step_frac_of_peak": round(gbs / HBM_PEAK_GBS, 4),
                     "objective_rel_err_vs_planted": abs(r.primal_objective - inst.optimal_value)
                     / abs(inst.optimal_value)}
        log(f"config2 {form}: {out[form]}")
    if all(k in objs for k in ("explicit", "lowrank")):
        out["explicit_vs_lowrank_objective_rel"] = abs(objs["explicit"] - objs["lowrank"]) / abs(objs["explicit"])
    # a whole solve within the cap: the same family at a fifth of the size
    # (n = 2e4, m = 1e4, M 4e3 x n, so Q keeps about 126 nonzeros per row)
    from pdal_amd.generators import planted_qp
    small = planted_qp(20000, 10000, 4000, seed=7)
    p = qlib.default_params(eps_optimal_relative=args.tol, eps_feasible_relative=args.tol, verbose=0,
                            time_sec_limit=args.qp_time, iteration_limit=args.qp_iters)
    qlib.reset_rng()
    t0 = time.perf_counter()
    r = qlib.solve(**small.qp_args("explicit"), params=p)
    wall = time.perf_counter() - t0
    out["explicit_fifth_size"] = {
        "workload": "planted QP n=2e4 m=1e4, Q=M^T M+1e-3 I (M 4e3 x n, 5 nnz/col), explicit Q",
        "status": r.status, "iterations": r.iterations, "inner_steps": r.inner_iterations,
        "solve_s": round(r.solve_time_sec, 3), "wall_s": round(wall, 3), "solves_per_s": round(1.0 / wall, 4),
        "kkt": r.kkt, "us_per_inner_step": round(1e6 * r.solve_time_sec / max(r.inner_iterations, 1), 3),
        "objective_rel_err_vs_planted": abs(r.primal_objective - small.optimal_value) / abs(small.optimal_value)}
    log(f"config2 fifth size: {out['explicit_fifth_size']}")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["config4", "config3", "config1", "config2"], default="config4")
    ap.add_argument("--heldout", type=int, default=256, help="config4: held-out instances per batch")
    ap.add_argument("--train", type=int, default=512, help="config4: training solves")
    ap.add_argument("--train-iters", type=int, default=40000, help="config4: iteration cap of the training solves")
    ap.add_argument("--epochs", type=int, default=300, help="config4: MLP training epochs")
    ap.add_argument("--instances", type=int, default=4096, help="config3: batch size")
    ap.add_argument("--batch-steps", type=int, default=1, help="config3 leg of the default run: timed solves")
    ap.add_argument("--scale", type=float, default=1.0, help="config1 only: shrink m and n")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--plain-sample", type=int, default=12)
    ap.add_argument("--cpu-sample", type=int, default=8)
    ap.add_argument("--cpu-iters", type=int, default=6400)
    ap.add_argument("--cpu-family-sample", type=int, default=2)
    ap.add_argument("--cpu-family-iters", type=int, default=640)
    ap.add_argument("--cpu-threads", type=int, default=16, help="CPU baseline: at most this many host threads")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-spmv", action="store_true", help="skip the large-LP SpMV leg")
    ap.add_argument("--no-c1-solve", action="store_true", help="large-LP leg: time the kernels only, no full solve")
    ap.add_argument("--no-batch-leg", action="store_true", help="config4: skip the 4096-instance batch leg")
    ap.add_argument("--kernels-only", action="store_true", help="set up and time kernels only (profiling)")
    ap.add_argument("--no-qp", action="store_true", help="config4: skip the FA_CP config-2 leg")
    ap.add_argument("--qp-time", type=float, default=25.0, help="config-2 leg: time cap per FA_CP solve (s)")
    ap.add_argument("--qp-iters", type=int, default=10_000_000, help="config-2 leg: iteration cap per FA_CP solve")
    ap.add_argument("--kernel-iters", type=int, default=20)
    ap.add_argument("--solve-blocks", type=int, default=0,
                    help="with --kernels-only (config4): also solve this many whole cold blocks of the held-out batch")
    ap.add_argument("--eval-freq", type=int, default=K_WARM,
                    help="config4: evaluation frequency of the held-out solves (warm, cold and plain)")
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
    if args.workload == "config2":
        res = run_config2_leg(args)
        print(json.dumps({"metric": "FA_CP planted QP, inner BB steps (see fields)", "config2": res}), flush=True)
        return 0
    if args.workload == "config4":
        res = run_config4(args, lib)
        metric = ("solves/sec to rel-KKT 1e-6, batched + learned warm start (fixed-A LP family, 256 held-out "
                  "instances per batch, one GPU)")
        config = {"workload": "fixed-A LP family m=5e4 n=1e5 8 nnz/col, b=b0+U theta_b, c=c0+V theta_c (32+32 "
                              "parameters, planted optimum); 256 held-out instances per batch, learned warm start",
                  "heldout": args.heldout, "train": args.train, "tolerance": args.tol, "alpha": 0.0,
                  "eval_freq": args.eval_freq}
    elif args.workload == "config3":
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
    if args.workload == "config4" and not args.no_batch_leg:
        leg_args = argparse.Namespace(**{**vars(args), "steps": args.batch_steps, "warmup": 0, "no_cpu": True})
        leg = run_config3(leg_args, lib)
        leg.pop("ms_per_step")
        leg["solves_per_s"] = round(leg.pop("value"), 4)
        leg["workload"] = "4096 planted LPs, m log-uniform [200,20000], n=2m, 8 nnz/col (uneven sizes, cold)"
        out["batch_4096"] = leg
    if args.workload in ("config4", "config3") and not args.no_spmv:
        out["spmv"] = spmv_leg(lib, args.kernel_iters, args.tol, not args.no_c1_solve)
    if args.workload == "config4" and not args.no_qp:
        out["qp_config2"] = run_config2_leg(args)
    out["roofline"] = roof
    out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
