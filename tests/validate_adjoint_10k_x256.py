"""Full-size adjoint validation: the 10k-bus meshed grid with 256 load
scenarios (+-30 %) in one plan, L = total active loss + sum over PQ buses of
(|V| - 1)^2, gradients with respect to every load p, q and every line r, x.

  * FP64 adjoint of all 256 scenarios in one call (J^T lambda = dL/dx on the
    FP64 factors at the solution);
  * against the sparse FP64 reference gradient (tests/adjoint_ref.py: SuperLU
    on the oracle's Jacobian, partials by fourth- and eighth-order
    differences) on two
    scenarios, every parameter;
  * against central differences of full re-solves on 500 random parameters:
    second order at h = 1e-6 relative and eighth order at h = 3e-2, every
    perturbed network an instance of a batched plan solved to the round-off
    floor, each parameter judged against the noise floor of L measured on
    near-identical copies (tests/fd_check.py); and eighth order on the 25
    parameters of each kind with the largest gradients, where the difference
    quotient resolves 1e-6.

Not collected by pytest (it takes minutes); run on a GPU box as
    python tests/validate_adjoint_10k_x256.py
Prints one JSON line with the error statistics.
"""
from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent / "oracle"))

from adjoint_ref import reference_gradient  # noqa: E402
from fd_check import GRAD_KEY, STENCILS, assess, fd_gradients, noise_floor  # noqa: E402
from gap_amd import capi, gridgen  # noqa: E402
from gap_amd.network import PQ  # noqa: E402


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fd_report(oracle, scen, out, params, order, h_rel, sigma, base, nb, nr):
    log(f"  order {order}, h = {h_rel:g}: {len(params) * len(STENCILS[order][0])} perturbed solves")
    t0 = time.time()
    g_fd, hs = fd_gradients(oracle, scen, params, order, h_rel, base)
    log(f"  done in {time.time() - t0:.1f} s")
    g_adj = np.array([out[GRAD_KEY[a]][s * (nb if a in ("p_w", "q_var") else nr) + k] for s, a, k in params])
    kinds = np.array([a for _, a, _ in params])
    rep = {}
    for kind in ("all", "p_w", "q_var", "br_r", "br_x"):
        m = np.ones(len(params), bool) if kind == "all" else kinds == kind
        r = assess(g_adj[m], g_fd[m], hs[m], sigma, order)
        rep[kind] = {k: v for k, v in r.items() if k not in ("ok", "resolvable_mask")}
    return rep


def main():
    import pyoracle
    oracle = pyoracle.Oracle()
    base = 100e6
    net = gridgen.large_mesh(10_000, seed=1)
    scen = gridgen.load_scenarios(net, 256, seed=2000)
    nb, nr = net.n_bus, net.n_branch
    t0 = time.time()
    plan = capi.Plan(scen)
    log(f"plan of 256 scenarios: {time.time() - t0:.1f} s")
    t0 = time.time()
    r = plan.solve(1e-12, 60, base)
    t_solve = time.time() - t0
    assert r["converged"].all()
    t0 = time.time()
    out = plan.adjoint(1.0, 1.0)
    t_adj = time.time() - t0
    log(f"solve {t_solve:.2f} s, adjoint {t_adj:.2f} s")
    report = {"scenarios": 256, "buses": nb, "branches": nr, "solve_s": round(t_solve, 3),
              "adjoint_s": round(t_adj, 3), "iterations_max": int(r["iterations"].max())}
    ref = {}
    for s in (0, 137):
        t0 = time.time()
        gp, gq, gr, gx, L = reference_gradient(oracle, scen[s], r["v"][s], base)
        log(f"reference gradient of scenario {s}: {time.time() - t0:.1f} s")

        def rel(a, b):
            return float(np.abs(a - b).max() / np.abs(b).max())
        ref[s] = {"p": rel(out["grad_p"][s * nb:(s + 1) * nb], gp), "q": rel(out["grad_q"][s * nb:(s + 1) * nb], gq),
                  "r": rel(out["grad_r"][s * nr:(s + 1) * nr], gr), "x": rel(out["grad_x"][s * nr:(s + 1) * nr], gx),
                  "loss": float(abs(out["loss"][s] - L) / abs(L))}
    report["vs_sparse_reference"] = ref
    sig = [noise_floor(oracle, scen[s], base, seed=s) for s in (0, 63, 137, 255)]
    sigma = max(sig)
    log(f"noise floor of L: {sig}")
    report["loss_noise_floor"] = sigma
    rng = np.random.default_rng(7)
    pq_buses = np.nonzero(net.bus_type == PQ)[0]
    params = []
    for _ in range(500):
        s = int(rng.integers(256))
        kind = ["p_w", "q_var", "br_r", "br_x"][int(rng.integers(4))]
        k = int(rng.choice(pq_buses)) if kind in ("p_w", "q_var") else int(rng.integers(nr))
        params.append((s, kind, k))
    report["fd_order2_h1e-6"] = fd_report(oracle, scen, out, params, 2, 1e-6, sigma, base, nb, nr)
    report["fd_order8_h3e-2"] = fd_report(oracle, scen, out, params, 8, 3e-2, sigma, base, nb, nr)
    # the 25 parameters of each kind with the largest |gradient| over all
    # scenarios: the ones whose L changes enough for a difference quotient to
    # resolve 1e-6 of the gradient
    top = []
    for arr in ("p_w", "q_var", "br_r", "br_x"):
        g = np.abs(out[GRAD_KEY[arr]])
        width = nb if arr in ("p_w", "q_var") else nr
        for flat in np.argsort(-g)[:25]:
            top.append((int(flat // width), arr, int(flat % width)))
    report["fd_order8_h3e-2_top25"] = fd_report(oracle, scen, out, top, 8, 3e-2, sigma, base, nb, nr)
    print(json.dumps(report), flush=True)


if __name__ == "__main__":
    main()
