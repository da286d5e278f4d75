"""Writes the golden fixtures under tests/golden/ (run in the container that has
the reference tree; the fixtures, not the reference, travel with the repo).

Sources of truth:
  * oracle/_ref/libgapref.so — the reference's own CPU power-flow sources
    (cpu_admittance_matrix.cpp, cpu_newton_raphson.cpp, cpu_lu_solver.cpp),
    compiled from the reference tree by oracle/Makefile;
  * the reference's PGM fixtures tests/pgm_validation/test_data/*/{input,output}.json,
    converted to solver rows with the reference harness's conversion
    (gap_amd.pgm) — the PGM |V| outputs are kept as expected values.

Each case is one .npz (no pickles) holding the network arrays, the solve
settings, the reference Y-bus (CSR + values), the reference Jacobian of the
first Newton iteration (pattern + values), the per-iteration max|F|, the
iteration count, the convergence flag and the final voltages.

Usage:  python tests/golden/make_golden.py [--failure-only]
"""
from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

from gap_amd import gridgen  # noqa: E402
from gap_amd.network import PQ, PV, SHUNT, SLACK, LOADGEN, Network  # noqa: E402
from gap_amd.pgm import read_pgm_input, read_pgm_output_voltages  # noqa: E402
import pyoracle  # noqa: E402

OUT = Path(__file__).resolve().parent
PGM_DIR = Path("/root/reference/tests/pgm_validation/test_data")
PGM_CASES = ["radial_1feeder_2nodepf", "radial_1feeder_4nodepf", "radial_1feeder_8nodepf",
             "radial_2feeder_2nodepf", "radial_2feeder_4nodepf", "radial_2feeder_8nodepf",
             "radial_3feeder_2nodepf", "radial_3feeder_4nodepf", "radial_3feeder_8nodepf",
             "radial_10feeder_10nodepf"]


def net_arrays(net: Network) -> dict:
    return {f"net_{k}": v for k, v in net.__dict__.items()}


def special_cases():
    """Hand-built networks covering edge cases of the reference semantics."""
    cases = {}
    # PV buses, a shunt appliance, a LOADGEN appliance and a parallel branch pair
    net = gridgen.small_mesh(40, seed=11)
    net = net.copy()
    bt = net.bus_type.copy()
    bt[[5, 17, 29]] = PV
    net.bus_type = bt
    u = net.u_pu.copy()
    u[[5, 17, 29]] = [1.01, 0.995, 1.0]
    net.u_pu = u
    p = net.p_w.copy()
    p[[5, 17, 29]] += [0.3e6, 0.2e6, 0.25e6]
    net.p_w = p
    q = net.q_var.copy()
    q[[5, 17, 29]] = 0.0
    net.q_var = q
    # duplicate of branch 3 (parallel line) and an out-of-service branch
    for k in ("br_id", "br_from", "br_to", "br_status", "br_r", "br_x", "br_g", "br_b"):
        a = getattr(net, k)
        extra = a[[3, 7]].copy()
        setattr(net, k, np.concatenate([a, extra]))
    net.br_id[-2:] = [900, 901]
    net.br_status[-1] = 0
    net.br_x[-2] *= 1.7
    net.app_id = np.array([1000, 1001, 1002], np.int32)
    net.app_node = np.array([9, 9, 21], np.int32)
    net.app_status = np.array([1, 1, 1], np.int32)
    net.app_type = np.array([SHUNT, LOADGEN, SHUNT], np.int32)
    net.app_p = np.array([0.0, -0.05e6, 0.0])
    net.app_q = np.array([0.0, -0.01e6, 0.0])
    net.app_g = np.array([0.0, 0.0, 1e-4])
    net.app_b = np.array([3e-3, 0.0, -2e-3])
    cases["special_pv_shunt_parallel"] = (Network(**net.__dict__), 100e6, 1e-8, 100)
    # 1-based bus ids and a slack that is not bus 0
    n2 = gridgen.small_mesh(25, seed=5)
    n2 = n2.copy()
    n2.bus_id = n2.bus_id + 1
    n2.br_from = n2.br_from + 1
    n2.br_to = n2.br_to + 1
    cases["one_based_ids"] = (n2, 100e6, 1e-8, 100)
    # tiny grid that does not converge within the iteration cap
    n3 = gridgen.small_mesh(12, seed=3, load_factor=40.0)
    cases["nonconvergent_small"] = (n3, 100e6, 1e-8, 7)
    return cases


def failure_cases():
    """The reference's failure semantics (cpu_newton_raphson.cpp:115-157, 382-433):
    a NaN load (in a later mismatch row, which max_element never takes, and in
    the first row, which it always keeps), and a PQ bus with a load whose every
    branch is out of service (its Jacobian rows and columns are exactly zero and
    dropped by the 1e-12 compaction, the factorisation fails and every
    correction is zero, so the iteration runs to the cap)."""
    cases = {}
    base = gridgen.small_mesh(20, seed=4)
    for name, bus in (("fail_nan_load_later_row", 7), ("fail_nan_load_first_row", 1)):
        n = base.copy()
        p = n.p_w.copy()
        p[bus] = np.nan
        n.p_w = p
        cases[name] = (n, 100e6, 1e-8, 6)
    n = base.copy()
    st = n.br_status.copy()
    iso = int(n.bus_id[13])
    st[(n.br_from == iso) | (n.br_to == iso)] = 0
    n.br_status = st
    cases["fail_isolated_pq_bus"] = (n, 100e6, 1e-8, 6)
    return cases


def run_case(R, name, net, base, tol, max_it, extra=None):
    t = time.time()
    r = R.solve(net, tol=tol, max_it=max_it, base_power=base, j_call=0)
    y = R.build_ybus(net)
    d = dict(net_arrays(net))
    d.update(base_power=base, tolerance=tol, max_iterations=max_it,
             y_rp=y[0], y_ci=y[1], y_val=y[2],
             iterations=r["iterations"], converged=r["converged"], final_mismatch=r["final_mismatch"],
             traj=r["traj"], v=r["v"])
    if "J" in r:
        d.update(j_rp=r["J"][0], j_ci=r["J"][1], j_val=r["J"][2])
    if extra:
        d.update(extra)
    np.savez_compressed(OUT / f"{name}.npz", **d)
    print(f"{name}: n={net.n_bus} it={r['iterations']} conv={r['converged']} ({time.time() - t:.1f}s)",
          flush=True)


def main():
    pyoracle.build(with_ref=True)
    R = pyoracle.Reference()
    if "--failure-only" in sys.argv:
        manifest = json.load(open(OUT / "manifest.json"))
        for name, (net, base, tol, mi) in failure_cases().items():
            run_case(R, name, net, base, tol, mi)
            manifest[name] = {"generator": "make_golden.failure_cases", "base_power": base}
        with open(OUT / "manifest.json", "w") as f:
            json.dump(manifest, f, indent=1)
        return
    manifest = {}
    for case in PGM_CASES:
        d = PGM_DIR / case
        meta = json.load(open(d / "metadata.json"))
        base = float(meta.get("base_power", 1e6))  # the harness default (test_pgm_validation.py:139)
        bus, br, _ = read_pgm_input(d / "input.json", base)
        net = Network.from_rows(bus, br)
        u_pu, u_ang = read_pgm_output_voltages(d / "output.json")
        run_case(R, f"pgm_{case}", net, base, 1e-8, 20, {"pgm_u_pu": u_pu, "pgm_u_angle": u_ang})
        manifest[f"pgm_{case}"] = {"source": f"tests/pgm_validation/test_data/{case}", "base_power": base}
    c0 = gridgen.small_mesh(200, seed=0)
    run_case(R, "mesh200_seed0", c0, 100e6, 1e-8, 100)
    manifest["mesh200_seed0"] = {"generator": "gridgen.small_mesh(200, seed=0)", "base_power": 100e6}
    for n, seed in [(60, 3), (120, 7)]:
        run_case(R, f"mesh{n}_seed{seed}", gridgen.small_mesh(n, seed=seed), 100e6, 1e-8, 100)
        manifest[f"mesh{n}_seed{seed}"] = {"generator": f"gridgen.small_mesh({n}, seed={seed})"}
    for name, (net, base, tol, mi) in special_cases().items():
        run_case(R, name, net, base, tol, mi)
        manifest[name] = {"generator": "make_golden.special_cases", "base_power": base}
    for name, (net, base, tol, mi) in failure_cases().items():
        run_case(R, name, net, base, tol, mi)
        manifest[name] = {"generator": "make_golden.failure_cases", "base_power": base}
    # LU known answers on seeded diagonally dominant complex matrices (the shape of
    # tests/validation/lu_solver_validation.cpp:23-105), solved by the reference LU
    rng = np.random.default_rng(12345)
    for n in (10, 60):
        dense = np.zeros((n, n), complex)
        for i in range(n):
            cols = rng.choice(n, size=min(4, n), replace=False)
            dense[i, cols] = rng.uniform(-1, 1, len(cols)) + 1j * rng.uniform(-1, 1, len(cols))
        dense = dense + dense.T * (rng.uniform(0, 1) > 2)  # keep unsymmetric values
        for i in range(n):
            dense[i, i] = np.abs(dense[i]).sum() + 1.0 + 0.5j
        rp = [0]
        ci, vals = [], []
        for i in range(n):
            nz = np.nonzero(dense[i])[0]
            ci.extend(nz.tolist())
            vals.extend(dense[i, nz].tolist())
            rp.append(len(ci))
        b = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
        rc, x = R.lu_solve(n, np.array(rp), np.array(ci), np.array(vals), b)
        assert rc == 0
        np.savez_compressed(OUT / f"lu_dd{n}.npz", rp=np.array(rp, np.int32), ci=np.array(ci, np.int32),
                            val=np.array(vals), b=b, x=x)
        manifest[f"lu_dd{n}"] = {"generator": "seeded diagonally dominant complex, mt-like seed 12345"}
    with open(OUT / "manifest.json", "w") as f:
        json.dump(manifest, f, indent=1)


if __name__ == "__main__":
    main()
