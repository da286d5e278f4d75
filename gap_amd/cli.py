"""Command line: Newton-Raphson power flow on the MI355X backend.

    python -m gap_amd.cli input.json [-o result.json] [--precision fp64|mixed]
    python -m gap_amd.cli --synthetic 200|10k|100k|batch4096 [--precision ...] [--repeat K]

An input file is a Power Grid Model JSON dataset (node / line / sym_load /
source), converted as the reference's validated path converts it (the voltage
base S_base is 1 MVA there, as in the reference harness).  --synthetic builds
one of the seeded workloads: the 200-bus grid (tree + 40 extra edges), the
10k-bus meshed grid with transformers, shunts and PV buses, the 100k-bus
meshed grid, or the batch of 4,096 grids of 50-5000 buses.  The output lists,
per network, the iteration count, convergence and final mismatch, and for a
single network per bus |V| in pu and the angle in rad and per branch the
flows (p, q, s, current magnitude and angle at both ends).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def _synthetic(name: str):
    from . import gridgen
    if name == "200":
        return [gridgen.small_mesh(200, seed=0, n_extra=40)]
    if name == "10k":
        return [gridgen.large_mesh(10_000, seed=1)]
    if name == "100k":
        return [gridgen.c2_grid()]
    if name == "batch4096":
        return gridgen.batch_grids(4096)
    raise SystemExit(f"unknown synthetic workload {name!r}")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m gap_amd.cli", description=__doc__.split("\n\n")[0])
    ap.add_argument("input", nargs="?", help="PGM JSON input dataset")
    ap.add_argument("--synthetic", choices=["200", "10k", "100k", "batch4096"])
    ap.add_argument("-o", "--output", help="write the result as JSON")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--schedule", choices=["auto", "levels", "grid"], default="auto")
    ap.add_argument("-t", "--tolerance", type=float, default=1e-8)
    ap.add_argument("-m", "--max-iter", type=int, default=100)
    ap.add_argument("--base-power", type=float, default=None, help="VA (default 1e6 for PGM input, 1e8 synthetic)")
    ap.add_argument("--repeat", type=int, default=1, help="time this many solves after the first")
    ap.add_argument("--graph", choices=["default", "on", "off"], default="default",
                    help="level schedule: replay the Newton step from a HIP graph, or launch its kernels one by one")
    ap.add_argument("--grid-shape", metavar="W,B,WORK",
                    help="grid schedule launch shape (the autotuner's knobs): wavefronts per workgroup, wavefronts "
                         "per workgroup of big networks, front slots + factor entries from which a network is big")
    ap.add_argument("--probe-copy", action="store_true",
                    help="after the solves, copy a 512 MiB device buffer 3 times (calibrates HBM byte counters)")
    args = ap.parse_args(argv)
    if bool(args.input) == bool(args.synthetic):
        ap.error("give an input file or --synthetic")

    from . import capi
    from .network import Network
    if args.input:
        from .pgm import read_pgm_input
        base = args.base_power or 1e6
        bus, br, _ = read_pgm_input(args.input, base)
        nets = [Network.from_rows(bus, br)]
    else:
        base = args.base_power or 100e6
        nets = _synthetic(args.synthetic)
    if not capi.device_available():
        raise SystemExit("no gfx950 GPU visible (or libgap_hip.so not built)")

    t0 = time.perf_counter()
    plan = capi.Plan(nets, args.precision, schedule=args.schedule)
    if args.graph != "default":
        plan.set_graph(args.graph == "on")
    if args.grid_shape:
        w, b, work = (int(x) for x in args.grid_shape.split(","))
        plan.set_tuning(grid_waves=w, grid_big_waves=b, grid_big_work=work)
    t_plan = time.perf_counter() - t0
    t0 = time.perf_counter()
    res = plan.solve(args.tolerance, args.max_iter, base)
    t_first = time.perf_counter() - t0
    times = []
    for _ in range(args.repeat - 1 if args.repeat > 1 else 0):
        t0 = time.perf_counter()
        plan.solve(args.tolerance, args.max_iter, base, fetch_v=False)
        times.append(time.perf_counter() - t0)
    st = plan.stats()
    summary = {
        "networks": len(nets), "buses": int(st["n_bus"]), "precision": args.precision,
        "schedule": plan.schedule, "max_front": int(st["max_front"]), "plan_s": round(t_plan, 3),
        "first_solve_ms": round(1e3 * t_first, 3),
        "solve_ms": round(1e3 * float(np.median(times)), 3) if times else None,
        "converged": int(res["converged"].sum()), "iterations_max": int(res["iterations"].max()),
        "final_mismatch_max": float(res["final_mismatch"].max()),
    }
    print(json.dumps(summary), flush=True)
    if args.probe_copy:
        import torch
        src = torch.ones(512 * 1024 * 1024 // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
    if args.output:
        out = {"summary": summary, "networks": []}
        for k, net in enumerate(nets):
            v = res["v"][k]
            entry = {"iterations": int(res["iterations"][k]), "converged": bool(res["converged"][k]),
                     "final_mismatch": float(res["final_mismatch"][k])}
            if len(nets) == 1:
                entry["node"] = [{"id": int(i), "u_pu": float(abs(x)), "u_angle": float(np.angle(x))}
                                 for i, x in zip(net.bus_id, v)]
                fl = plan.branch_flows()
                entry["line"] = [{"id": int(bid), **{f: float(fl[f][k]) for f in capi.FLOW_FIELDS}}
                                 for k, bid in enumerate(net.br_id)]
            out["networks"].append(entry)
        with open(args.output, "w") as f:
            json.dump(out, f)
    return 0 if res["converged"].all() else 1


if __name__ == "__main__":
    sys.exit(main())
