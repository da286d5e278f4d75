"""Central finite differences of full re-solves, with the noise floor they
can resolve, for checking adjoint gradients on large grids.

A perturbed network is solved to the round-off floor, but that floor is not
zero: with per-unit admittances up to ~1e4 (short lines on a 110 kV base) the
mismatch stops near 1e-12 and the loss L evaluated at the solution carries an
error of the same order that is not smooth in the parameter.  A difference
quotient over a step h therefore has a noise term sigma_L * |w|_2 / (den * h)
next to its truncation error.  `noise_floor` measures sigma_L from copies of
the network whose loads differ in the 13th digit (so L differs by ~1e-18 and
the spread is the solver's own); `assess` passes a parameter when

    |g_adjoint - g_fd| <= 1e-6 |g_fd| + 3 * noise bound,

and reports separately the parameters whose 3-sigma noise bound is at most
half of 1e-6 |g_fd| (the ones a finite difference can check to 1e-6).

Test infrastructure: imports the oracle only as the checker.
"""
from __future__ import annotations

import sys

import numpy as np
import scipy.sparse as sp

from gap_amd import capi
from gap_amd.network import PQ

STENCILS = {
    2: ((1, -1), (1.0, -1.0), 2.0),
    4: ((2, 1, -1, -2), (-1.0, 8.0, -8.0, 1.0), 12.0),
    8: ((4, 3, 2, 1, -1, -2, -3, -4), (-3.0, 32.0, -168.0, 672.0, -672.0, 168.0, -32.0, 3.0), 840.0),
}
GRAD_KEY = {"p_w": "grad_p", "q_var": "grad_q", "br_r": "grad_r", "br_x": "grad_x"}


def loss(oracle, net, v, base, w_loss=1.0, w_dev=1.0):
    """L = w_loss * sum_i Re(S_i) [pu] + w_dev * sum_PQ (|V_i| - 1)^2 with the
    oracle's Y-bus."""
    rp, ci, val = oracle.build_ybus(net)
    Y = sp.csr_matrix((val * (net.u_rated[0] ** 2 / base), ci, rp), shape=(net.n_bus, net.n_bus))
    s = v * np.conj(Y @ v)
    pq = net.bus_type == PQ
    return w_loss * s.real.sum() + w_dev * ((np.abs(v[pq]) - 1.0) ** 2).sum()


def _solve_losses(oracle, nets, base, chunk):
    out = np.zeros(len(nets))
    for c0 in range(0, len(nets), chunk):
        part = nets[c0:c0 + chunk]
        res = capi.Plan(part).solve(1e-13, 60, base)
        assert res["converged"].all()
        for u, n2 in enumerate(part):
            out[c0 + u] = loss(oracle, n2, res["v"][u], base)
        if len(nets) > chunk:
            print(f"    {c0 + len(part)}/{len(nets)} perturbed networks solved", file=sys.stderr, flush=True)
    return out


def fd_gradients(oracle, scen, params, order, h_rel, base, chunk=256):
    """params: (scenario, array name, index) triples.  Returns (g_fd, h)."""
    steps, weights, den = STENCILS[order]
    nets, hs = [], []
    for s, arr, k in params:
        h = h_rel * abs(getattr(scen[s], arr)[k])
        hs.append(h)
        for st in steps:
            n2 = scen[s].copy()
            getattr(n2, arr)[k] += st * h
            nets.append(n2)
    Ls = _solve_losses(oracle, nets, base, chunk).reshape(len(params), len(steps))
    hs = np.array(hs)
    return (Ls @ np.array(weights)) / (den * hs), hs


def noise_floor(oracle, net, base, copies=8, seed=0):
    """Standard deviation of L over copies of `net` whose loads are scaled by
    1 + 1e-13 at one random PQ bus each."""
    rng = np.random.default_rng(seed)
    loads = np.nonzero((net.bus_type == PQ) & (net.p_w != 0))[0]
    nets = []
    for _ in range(copies):
        n2 = net.copy()
        n2.p_w[int(rng.choice(loads))] *= 1.0 + 1e-13
        nets.append(n2)
    return float(np.std(_solve_losses(oracle, nets, base, copies), ddof=1))


def assess(g_adj, g_fd, hs, sigma, order, rel=1e-6):
    _, weights, den = STENCILS[order]
    bound = sigma * np.linalg.norm(weights) / (den * hs)
    err = np.abs(g_adj - g_fd)
    ok = err <= rel * np.abs(g_fd) + 3.0 * bound
    resolvable = 3.0 * bound <= 0.5 * rel * np.abs(g_fd)
    relerr = err / np.abs(g_fd)
    return {
        "params": int(len(g_fd)), "pass": int(ok.sum()),
        "worst_ratio": float((err / (rel * np.abs(g_fd) + 3.0 * bound)).max()),
        "resolvable": int(resolvable.sum()),
        "resolvable_max_rel": float(relerr[resolvable].max()) if resolvable.any() else None,
        "median_rel": float(np.median(relerr)),
        "normwise_rel": float(np.linalg.norm(g_adj - g_fd) / np.linalg.norm(g_fd)),
        "ok": ok, "resolvable_mask": resolvable,
    }
