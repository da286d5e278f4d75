#!/usr/bin/env python3
"""Throughput benchmark: converged power-flow solves per second at 1e-8 pu.

Workload (one GPU): a batch of 4,096 independent grids, sizes log-uniform in
[50, 5000] buses (random tree + 40 extra edges; per-bus loads drawn as for the
200-bus grid and scaled by min(1, 200/n)^1.5, per grid by U[0.7, 1.3] and by a
batch load level of 0.5), solved together in one plan with FP32 factorisation
and FP64 residuals.  A step is one full batched solve from flat start (Y-bus
assembly + damped Newton-Raphson to max|F| < 1e-8 pu on every grid); inputs
are resident in HBM before timing.  The value counts converged solves only;
grids that do not converge are reported in `not_converged`.

Also reported:
  * plain_fp64: the same grids solved one at a time with the FP64 path (a
    stratified sample of the batch, plans built before timing) -> solves/s and
    the batched/plain ratio;
  * roofline of the dominant kernel family: algorithmic bytes per launch over
    its mean launch time (HIP events around every launch, on the plan's stream);
  * cpu_baseline: the reference's own CPU Newton-Raphson (compiled from its
    sources, 1 thread) on the smallest grids of the batch, about 15 s of work;
    cpu_port_baseline: the FP64 C restatement (pforacle.c, 1 thread) on a
    stratified sample of the batch;
  * max error of the measured (mixed) solution against the FP64 CPU oracle on
    that stratified sample (or, with --no-cpu, on the committed oracle fixture
    of the batch), and the number of grids whose iteration count differs.

Before timing, gap_amd.tuning.autotune_plan builds the batch's plan for each
plan-time candidate (supernode amalgamation thresholds, the grid tasks'
smallest lane class and packing), times its launch-shape candidates, and keeps
the fastest plan (`autotune` in the output; --no-tune keeps the defaults).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--grids 4096] [--no-tune]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "converged PF solves/sec @1e-8 pu: mixed/batched vs plain FP64 GPU NR"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stratified(sizes, count):
    order = np.argsort(sizes, kind="stable")
    pick = np.linspace(0, len(order) - 1, count).round().astype(int)
    return order[pick]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grids", type=int, default=4096)
    ap.add_argument("--plain-sample", type=int, default=64)
    ap.add_argument("--cpu-sample", type=int, default=1536)
    ap.add_argument("--cpu-budget", type=float, default=15.0,
                    help="seconds of reference CPU solves for cpu_baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-solves", type=int, default=5,
                    help="solves timed launch by launch (HIP events) for the roofline")
    ap.add_argument("--no-tune", action="store_true", help="use the default plan options and launch shape")
    ap.add_argument("--probe-copy", action="store_true",
                    help="after the run, copy a 512 MiB device tensor 3 times (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the batch shards by "
              "independent grids with no exchange, so N GPUs = N independent runs.", file=sys.stderr)
        sys.exit(2)

    import torch
    from gap_amd import capi, gridgen, tuning

    if not torch.cuda.is_available() or not capi.device_available():
        raise SystemExit("bench.py needs a gfx950 GPU and the built HIP library")
    torch.cuda.set_device(0)

    tol, max_it, base = 1e-8, 100, 100e6
    t0 = time.time()
    nets = gridgen.batch_grids(args.grids)
    sizes = np.array([n.n_bus for n in nets])
    log(f"[bench] generated {len(nets)} grids, {sizes.sum()} buses in {time.time() - t0:.1f}s")
    t0 = time.time()
    if args.no_tune:
        plan = capi.Plan(nets, "mixed")
        tuned = None
    else:
        # plan-time options (amalgamation, grid task layout) and the launch
        # shape chosen by timing solves of this batch; outside the timed region
        tuned = tuning.autotune_plan(nets, "mixed", tolerance=tol, max_iterations=max_it, base_power=base,
                                     repeats=2)
        plan = tuned["plan"]
    st = plan.stats()
    log(f"[bench] plan: {time.time() - t0:.1f}s (analysis {st['analysis_ms']:.0f} ms), fronts {st['n_fronts']}, "
        f"levels {st['n_levels']}, max front {st['max_front']}; options {plan.options}, tuning {plan.tuning()}")

    for _ in range(args.warmup):
        r = plan.solve(tol, max_it, base, fetch_v=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        r = plan.solve(tol, max_it, base, fetch_v=False)
        log(f"[bench] step {k + 1}/{args.steps} done")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    its = r["iterations"]
    n_conv = int(r["converged"].sum())
    if n_conv < len(nets):
        log(f"[bench] {len(nets) - n_conv} grids did not converge; only converged solves are counted")
    value = n_conv * args.steps / dt
    ms_per_step = 1e3 * dt / args.steps
    log(f"[bench] batched mixed: {value:.0f} solves/s ({ms_per_step:.2f} ms/step), iterations p50 "
        f"{np.percentile(its, 50):.0f} p99 {np.percentile(its, 99):.0f}")

    # the same plan on the level-synchronous schedule (one launch per phase and tree level)
    levels_value = None
    level_kernels = {}
    if plan.schedule == "grid":
        plan.set_schedule("levels")
        plan.solve(tol, max_it, base, fetch_v=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(max(1, args.steps // 2)):
            plan.solve(tol, max_it, base, fetch_v=False)
        torch.cuda.synchronize()
        levels_value = len(nets) * max(1, args.steps // 2) / (time.perf_counter() - t0)
        # its kernels one by one: algorithmic bytes over HIP-event time, per solve
        plan.profile(True)
        plan.solve(tol, max_it, base, fetch_v=False)
        lrep = plan.profile_report()
        plan.profile(False)
        for name, v in lrep.items():
            if not v["launches"]:
                continue
            gbs = v["bytes"] / (v["ms"] * 1e-3) / 1e9 if v["ms"] > 0 and v["bytes"] > 0 else None
            level_kernels[name] = {"ms_per_solve": round(v["ms"], 3), "launches": v["launches"],
                                   "achieved_GBps": None if gbs is None else round(gbs, 1),
                                   "frac_hbm_peak": None if gbs is None else round(gbs / HBM_PEAK_GBS, 4)}
        plan.set_schedule("auto")
        log(f"[bench] batched mixed, level schedule: {levels_value:.0f} solves/s; "
            + ", ".join(f"{k} {v['achieved_GBps']} GB/s" for k, v in level_kernels.items()
                        if k in ("k_ybus", "k_mismatch", "k_jacobian", "k_update") and v["achieved_GBps"]))

    # accuracy of the mixed path against the FP64 path on the same batch
    r_mixed = plan.solve(tol, max_it, base, fetch_v=True)
    plan64 = capi.Plan(nets, "fp64")
    r64 = plan64.solve(tol, max_it, base, fetch_v=True)
    dv = max(float(np.abs(a - b).max()) for a, b in zip(r_mixed["v"], r64["v"]))
    it_diff = int(np.abs(r_mixed["iterations"].astype(int) - r64["iterations"].astype(int)).max())
    mixed_v, mixed_its = r_mixed["v"], r_mixed["iterations"].copy()
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps // 2)):
        plan64.solve(tol, max_it, base, fetch_v=False)
    torch.cuda.synchronize()
    batched_fp64 = len(nets) * max(1, args.steps // 2) / (time.perf_counter() - t0)
    del plan64, r64

    # plain FP64 single-instance GPU path on a stratified sample of the same grids,
    # on both schedules (level-synchronous, and one workgroup for the grid)
    idx = stratified(sizes, min(args.plain_sample, len(nets)))
    plans = [capi.Plan([nets[i]], "fp64") for i in idx]
    single = {}
    for sched in ("levels", "auto"):
        for p in plans:
            p.set_schedule(sched)
        for p in plans[:4]:
            p.solve(tol, max_it, base, fetch_v=False)
        times = []
        for p in plans:
            t1 = time.perf_counter()
            p.solve(tol, max_it, base, fetch_v=False)
            times.append(time.perf_counter() - t1)
        single[sched] = len(times) / sum(times)
    del plans
    plain = max(single.values())
    log(f"[bench] plain FP64 single-instance: levels {single['levels']:.1f}, auto {single['auto']:.1f} solves/s "
        f"(sample of {len(idx)}); batched FP64 {batched_fp64:.0f} solves/s")

    # live per-kernel timing (HIP events around every launch on the plan's stream)
    # (summed over a few solves: one launch of the grid kernel is one solve)
    rep = {}
    plan.profile(True)
    for _ in range(args.profile_solves):
        plan.solve(tol, max_it, base, fetch_v=False)
        for k, v in plan.profile_report().items():
            acc = rep.setdefault(k, {"ms": 0.0, "launches": 0, "bytes": 0.0})
            for f in acc:
                acc[f] += v[f]
    plan.profile(False)
    fam = max(rep, key=lambda k: rep[k]["ms"])  # dominant kernel by device time
    f = rep[fam]
    avg_ms = f["ms"] / max(f["launches"], 1)
    bytes_per_launch = f["bytes"] / max(f["launches"], 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": fam, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(fam),
            "avg_launch_us": round(avg_ms * 1e3, 3), "launches_per_step": f["launches"] // args.profile_solves,
            "algorithmic_bytes_per_launch": round(bytes_per_launch),
            "kernels_ms_per_step": {k: round(v["ms"] / args.profile_solves, 3) for k, v in rep.items() if v["launches"]},
            "profiled_solves": args.profile_solves}

    cpu, port, oracle_cmp = None, None, None
    if not args.no_cpu:
        cpu = cpu_reference(nets, sizes, args.cpu_budget, tol, max_it, base)
        port, oracle_sol = cpu_port(nets, sizes, args.cpu_sample, tol, max_it, base)
        if oracle_sol:
            oracle_cmp = compare_to_oracle(oracle_sol, mixed_v, mixed_its,
                                           f"{len(oracle_sol)} grids stratified by size (the cpu_port sample)")
    if oracle_cmp is None and args.grids == 4096:
        oracle_cmp = compare_to_oracle(fixture_solutions(), mixed_v, mixed_its,
                                       "the committed oracle fixture of this batch "
                                       "(tests/golden/batch4096_oracle_sample.npz)")
    if oracle_cmp:
        log(f"[bench] vs FP64 CPU oracle: max|dV| {oracle_cmp['max_abs_dv']:.3e} on {oracle_cmp['sample']}, "
            f"{oracle_cmp['iteration_mismatches']} iteration-count mismatches")

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 factor+solve / fp64 residual, update, convergence", "data": "synthetic",
        "config": {"workload": f"batch of {len(nets)} grids, 50-5000 buses log-uniform (random tree + 40 extra "
                               "edges); per-bus loads as the 200-bus grid's x min(1,200/n)^1.5, per grid "
                               "x U[0.7,1.3], batch load level 0.5; one plan, mixed precision",
                   "grids": len(nets), "buses": int(sizes.sum()), "tolerance_pu": tol,
                   "base_power_va": base, "max_iterations": max_it},
        "schedule": plan.schedule,
        "autotune": None if tuned is None else {
            "options": tuned["options"], "tuning": tuned["tuning"], "ms_per_solve": round(tuned["ms"], 3),
            "candidates": [{"options": o, "best_tuning": {k: t[k] for k in ("grid_waves", "grid_big_waves",
                                                                              "grid_big_work")},
                            "ms_per_solve": round(ms, 3)} for o, t, ms in tuned["table"]]},
        "plain_fp64_solves_per_s": round(plain, 2),
        "plain_fp64_by_schedule": {k: round(v, 2) for k, v in single.items()},
        "speedup_vs_plain_fp64": round(value / plain, 2),
        "batched_mixed_level_schedule_solves_per_s": None if levels_value is None else round(levels_value, 1),
        "level_schedule_kernels": level_kernels or None,
        "batched_fp64_solves_per_s": round(batched_fp64, 1),
        "converged": n_conv, "not_converged": len(nets) - n_conv,
        "iterations_p50": float(np.percentile(its, 50)), "iterations_p99": float(np.percentile(its, 99)),
        "max_abs_dv_vs_fp64_oracle": None if oracle_cmp is None else oracle_cmp["max_abs_dv"],
        "iteration_mismatches_vs_oracle": None if oracle_cmp is None else oracle_cmp["iteration_mismatches"],
        "oracle_comparison_sample": None if oracle_cmp is None else oracle_cmp["sample"],
        "mixed_vs_fp64_max_abs_dv": dv, "mixed_vs_fp64_max_iteration_diff": it_diff,
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_port_baseline": port,
    }
    print(json.dumps(out), flush=True)
    if args.probe_copy:
        n = 512 * 1024 * 1024 // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: 3 x {4 * n} bytes read and written")


def cpu_port(nets, sizes, count, tol, max_it, base):
    """FP64 C restatement of the reference path (oracle), one thread, on a
    stratified sample of the batch."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    if not pyoracle.ORACLE_SO.exists():
        return None
    o = pyoracle.Oracle()
    idx = stratified(sizes, min(count, len(nets)))
    sols = {}
    t0 = time.perf_counter()
    for i in idx:
        r = o.solve(nets[i], tol=tol, max_it=max_it, base_power=base, j_call=-1, traj=False)
        sols[int(i)] = (r["iterations"], r["v"])
    dt = time.perf_counter() - t0
    return ({"value": round(len(sols) / dt, 3), "unit": "solves/s", "cores": 1, "kind": "port",
             "sample": f"{len(sols)} grids stratified by size from the same batch "
                       f"({int(sizes[list(sols)].sum())} buses), FP64 C oracle (pforacle.c), {dt:.1f} s",
             **host_cpu()}, sols)


def host_cpu():
    """CPU model and core counts of the host the CPU legs ran on."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"cpu_model": model, "host_cpus": os.cpu_count(), "usable_cpus": usable}


def compare_to_oracle(sols, mixed_v, mixed_its, sample):
    """max |V - V_oracle| and iteration-count mismatches over the oracle's grids."""
    if not sols:
        return None
    dv, bad = 0.0, 0
    for i, (its, v) in sols.items():
        dv = max(dv, float(np.abs(mixed_v[i] - v).max()))
        bad += int(int(mixed_its[i]) != int(its))
    return {"max_abs_dv": dv, "iteration_mismatches": bad, "sample": sample}


def fixture_solutions():
    """The committed FP64 oracle solutions of a sample of the 4,096-grid batch."""
    p = ROOT / "tests" / "golden" / "batch4096_oracle_sample.npz"
    if not p.exists():
        return None
    d = np.load(p, allow_pickle=False)
    return {int(i): (int(d["iterations"][k]), d["v"][d["v_off"][k]:d["v_off"][k + 1]])
            for k, i in enumerate(d["index"])}


def cpu_reference(nets, sizes, budget_s, tol, max_it, base):
    """The reference's own CPU Newton-Raphson (its sources compiled under
    oracle/_ref), one thread, on the smallest grids of the batch until about
    budget_s seconds are spent.  Its dense Jacobian and unordered LU grow
    steeply with size (about 3.7 s for 200 buses), so only the small end of the
    batch is timed; the figure is an upper bound for the whole batch."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    if not pyoracle.reference_available():
        return None
    ref = pyoracle.Reference()
    order = np.argsort(sizes, kind="stable")
    done, t0 = [], time.perf_counter()
    for i in order:
        ref.solve(nets[i], tol=tol, max_it=max_it, base_power=base, j_call=-1, traj=False)
        done.append(int(i))
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    sz = sizes[done]
    return {"value": round(len(done) / dt, 3), "unit": "solves/s", "cores": 1, "kind": "reference",
            "sample": f"the {len(done)} smallest grids of the batch ({int(sz.min())}-{int(sz.max())} buses, "
                      f"{int(sz.sum())} buses), reference CPU NR compiled from its sources, {dt:.1f} s; "
                      f"an upper bound for the batch (the batch's grids reach 5000 buses)",
            **host_cpu()}


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` on this workload from the committed
    rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE passes, calibrated), or None."""
    p = ROOT / "profiles" / "batch4096_traffic.json"
    if not p.exists():
        return None
    d = json.loads(p.read_text())
    if d.get("kernel") != kernel:
        return None
    return d.get("hbm_bytes_per_launch")


if __name__ == "__main__":
    main()
