"""Seeded synthetic grids for the power-flow workloads.

All grids are single-voltage-level 10 kV networks built on a random recursive
spanning tree (bus i attaches to a uniform bus j < i) plus extra random edges,
rejecting self-loops and duplicate bus pairs, so no parallel branches occur
(the dense Jacobian of the reference CPU path keeps only the last duplicate Y
entry while its mismatch sums them, cpu_newton_raphson.cpp:585-597 vs 317-330).

Line parameters: r1, x1 ~ U[0.05, 0.5] ohm, c1 ~ U[1e-8, 1e-7] F, tan1 = 0,
giving b1 = 2*pi*50*c1 S and g1 = 0 (types.h:258-266).  One source at bus 0
(u_ref = 1.0 pu) becomes the single SLACK bus; every other bus carries a load
p ~ U[0.05, 0.5] MW, q = p * U[0.1, 0.4], entered in generator sign.

Workloads (sizes and seeds):
  small_mesh(200, seed=0)          200-bus grid, tree + 40 extra edges
  large_mesh(10_000, seed=1)       tree + 3,000 edges, 1,500 transformer branches,
                                   300 shunts, 15 % PV buses
  batch_grids(4096, seed=1000)     sizes log-uniform in [50, 5000], each a
                                   small_mesh-style grid (tree + 40 extra edges)
                                   with its own seed, loads scaled by U[0.7, 1.3]
                                   and by a batch load level of 0.5

Feasibility: with 0.05-0.5 ohm lines, a 10 kV feeder cannot carry hundreds of
0.275 MW loads, so the per-bus load is scaled by ``min(1, 200/n) ** 1.5``
(grids of up to 200 buses keep the stated distribution; a 5,000-bus grid
carries about 11 MW).  With the constant-total-load scaling ``min(1, 200/n)``,
about 3 % of the deep radial grids collapse under the damped Newton iteration
(|V| driven to the 0.5 pu clamp, never converging); the same scaling applies to
PV injections.
"""
from __future__ import annotations

import json
import math

import numpy as np

from .network import PQ, PV, SHUNT, SLACK, Network

U_RATED = 10_000.0
OMEGA = 2.0 * math.pi * 50.0


def _edges(rng: np.random.Generator, n: int, n_extra: int):
    """Random recursive tree plus ``n_extra`` distinct non-tree edges."""
    i = np.arange(1, n)
    parent = np.floor(rng.random(n - 1) * i).astype(np.int64)
    frm = list(parent)
    to = list(i)
    seen = set(zip(np.minimum(parent, i).tolist(), np.maximum(parent, i).tolist()))
    max_extra = n * (n - 1) // 2 - (n - 1)
    if n_extra > max_extra:
        raise ValueError(f"{n_extra} extra edges do not fit a {n}-bus simple graph")
    while n_extra > 0:
        a, b = rng.integers(0, n, size=2)
        if a == b:
            continue
        key = (min(a, b), max(a, b))
        if key in seen:
            continue
        seen.add(key)
        frm.append(int(a))
        to.append(int(b))
        n_extra -= 1
    return np.asarray(frm, np.int64), np.asarray(to, np.int64)


def load_scale(n: int) -> float:
    return min(1.0, 200.0 / n) ** 1.5


def small_mesh(n: int = 200, seed: int = 0, n_extra: int | None = None,
               load_factor: float = 1.0) -> Network:
    """Tree + extra-edge 10 kV grid; ``n_extra`` defaults to n/5 (40 at 200 buses)."""
    if n < 2:
        raise ValueError("a grid needs at least two buses")
    rng = np.random.default_rng(seed)
    if n_extra is None:
        n_extra = min(max(n // 5, 0), n * (n - 1) // 2 - (n - 1))
    frm, to = _edges(rng, n, n_extra)
    m = len(frm)
    r = rng.uniform(0.05, 0.5, m)
    x = rng.uniform(0.05, 0.5, m)
    c1 = rng.uniform(1e-8, 1e-7, m)
    s = load_scale(n) * load_factor
    p = rng.uniform(0.05e6, 0.5e6, n) * s
    q = p * rng.uniform(0.1, 0.4, n)
    p[0] = q[0] = 0.0
    bus_type = np.full(n, PQ, np.int32)
    bus_type[0] = SLACK
    return Network(
        bus_id=np.arange(n), u_rated=np.full(n, U_RATED), bus_type=bus_type,
        u_pu=np.ones(n), p_w=-p, q_var=-q,
        br_id=np.arange(n, n + m), br_from=frm, br_to=to, br_status=np.ones(m, np.int32),
        br_r=r, br_x=x, br_g=np.zeros(m), br_b=OMEGA * c1)


def transformer_rx(uk: float | np.ndarray, sn: float, u_ref: float, xr: float = 10.0):
    """r, x of a two-winding transformer from its short-circuit voltage.

    Same construction as the reference JSON reader (json_io.cpp:230-252):
    |z| = uk * U^2 / sn and X/R = 10.  The impedance is referred to ``u_ref``;
    the reference per-unit system uses a single voltage base (bus 0,
    cpu_newton_raphson.cpp:46-47), so transformers are referred to the 10 kV side.
    """
    z = np.asarray(uk) * (u_ref * u_ref) / sn
    rf = 1.0 / math.sqrt(1.0 + xr * xr)
    return z * rf, z * xr * rf


def large_mesh(n: int = 10_000, seed: int = 1, n_extra: int = 3000, n_trafo: int = 1500,
               n_shunt: int = 300, pv_frac: float = 0.15) -> Network:
    """Meshed grid with transformer branches, shunts and PV buses."""
    rng = np.random.default_rng(seed)
    frm, to = _edges(rng, n, n_extra)
    m = len(frm)
    r = rng.uniform(0.05, 0.5, m)
    x = rng.uniform(0.05, 0.5, m)
    b = OMEGA * rng.uniform(1e-8, 1e-7, m)
    trafo = rng.choice(m, size=min(n_trafo, m), replace=False)
    uk = rng.uniform(0.08, 0.12, len(trafo))
    rng.integers(-5, 6, len(trafo))  # tap +-5 x 1.25 %: drawn, unused (the Y-bus never reads k, cpu_admittance_matrix.cpp:121-139)
    tr, tx = transformer_rx(uk, 40e6, U_RATED)
    r[trafo], x[trafo], b[trafo] = tr, tx, 0.0
    s = load_scale(n)
    p = rng.uniform(0.05e6, 0.5e6, n) * s
    q = p * rng.uniform(0.1, 0.4, n)
    p[0] = q[0] = 0.0
    bus_type = np.full(n, PQ, np.int32)
    bus_type[0] = SLACK
    n_pv = int(round(pv_frac * n))
    pv = 1 + rng.choice(n - 1, size=n_pv, replace=False)
    bus_type[pv] = PV
    pg = rng.uniform(1e6, 20e6, n_pv) * s * 0.05
    u_pu = np.ones(n)
    u_pu[pv] = 1.0
    p_inj = -p
    q_inj = -q
    p_inj[pv] += pg
    q_inj[pv] = 0.0
    sh_bus = 1 + rng.choice(n - 1, size=n_shunt, replace=False)
    q_mvar = rng.uniform(-2.0, 2.0, n_shunt)
    sh_b = q_mvar * 1e6 / (U_RATED * U_RATED) * s
    k = np.arange(n_shunt)
    return Network(
        bus_id=np.arange(n), u_rated=np.full(n, U_RATED), bus_type=bus_type,
        u_pu=u_pu, p_w=p_inj, q_var=q_inj,
        br_id=np.arange(n, n + m), br_from=frm, br_to=to, br_status=np.ones(m, np.int32),
        br_r=r, br_x=x, br_g=np.zeros(m), br_b=b,
        app_id=n + m + k, app_node=sh_bus, app_status=np.ones(n_shunt, np.int32),
        app_type=np.full(n_shunt, SHUNT, np.int32), app_p=np.zeros(n_shunt),
        app_q=np.zeros(n_shunt), app_g=np.zeros(n_shunt), app_b=sh_b)


def c2_grid(n: int = 100_000, seed: int = 7) -> Network:
    """The 100k-bus meshed grid: tree plus 0.3 n extra edges (average degree
    2.6), 15 % of the branches transformers and 3 % of the buses shunts as in
    the 10k-bus grid, 15 % PV.  Its loads carry the size scaling of
    ``load_scale``; the flat start's 0.98 pu against the 1.0 pu slack and PV
    set-points dominates the first mismatch, so the damped iteration takes
    about 11 steps to 1e-8 pu whatever the load level."""
    return large_mesh(n, seed=seed, n_extra=int(0.3 * n), n_trafo=int(0.15 * n), n_shunt=int(0.03 * n))


def batch_sizes(count: int = 4096, seed: int = 1000, lo: int = 50, hi: int = 5000) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi + 1), count))).astype(np.int64).clip(lo, hi)


def batch_grids(count: int = 4096, seed: int = 1000, lo: int = 50, hi: int = 5000,
                n_extra: int = 40, load_level: float = 0.5) -> list[Network]:
    """Independent grids built like the 200-bus one: random tree plus ``n_extra``
    (40) random extra edges, loads scaled by a per-grid factor U[0.7, 1.3].

    ``load_level`` scales every grid's loads on top of that.  At 1.0, 12 of the
    4,096 grids (123-321 buses, unlucky deep trees) diverge under the damped
    Newton iteration of the reference -- in the FP64 oracle and on the GPU
    alike -- and 3 still do at 0.75; at 0.5 all converge in 8-10 iterations."""
    sizes = batch_sizes(count, seed, lo, hi)
    out = []
    for k, n in enumerate(sizes):
        f = np.random.default_rng(seed + 1 + k).uniform(0.7, 1.3)
        out.append(small_mesh(int(n), seed=seed + 1 + k, n_extra=n_extra, load_factor=float(f) * load_level))
    return out


def load_scenarios(net: Network, count: int, seed: int, spread: float = 0.3) -> list[Network]:
    """Copies of ``net`` with every PQ/PV load scaled by its own U[1-spread, 1+spread]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        f = rng.uniform(1.0 - spread, 1.0 + spread, net.n_bus)
        g = net.copy()
        g.p_w = g.p_w * f
        g.q_var = g.q_var * f
        out.append(g)
    return out


def to_pgm_json(net: Network, u_ref: float = 1.0) -> dict:
    """PGM input dataset (node / line / sym_load / source) for a grid made here."""
    nodes = [{"id": int(i), "u_rated": float(u)} for i, u in zip(net.bus_id, net.u_rated)]
    lines = []
    for k in range(net.n_branch):
        lines.append({"id": int(net.br_id[k]), "from_node": int(net.br_from[k]),
                      "to_node": int(net.br_to[k]), "from_status": int(net.br_status[k]),
                      "to_status": int(net.br_status[k]), "r1": float(net.br_r[k]),
                      "x1": float(net.br_x[k]), "c1": float(net.br_b[k] / OMEGA), "tan1": 0.0,
                      "i_n": 1000.0})
    base = int(max(net.br_id.max(initial=0), net.bus_id.max(initial=0))) + 1
    loads, sources = [], []
    for i in range(net.n_bus):
        if net.bus_type[i] == SLACK:
            sources.append({"id": base + net.n_bus + i, "node": int(net.bus_id[i]), "status": 1,
                            "u_ref": float(net.u_pu[i]), "sk": 1e20, "rx_ratio": 0.1})
        elif net.p_w[i] != 0.0 or net.q_var[i] != 0.0:
            loads.append({"id": base + i, "node": int(net.bus_id[i]), "status": 1,
                          "type": "const_power", "p_specified": float(-net.p_w[i]),
                          "q_specified": float(-net.q_var[i])})
    return {"version": "1.0", "type": "input", "is_batch": False, "attributes": {},
            "data": {"node": nodes, "line": lines, "sym_load": loads, "source": sources}}


def write_pgm_json(net: Network, path) -> None:
    with open(path, "w") as f:
        json.dump(to_pgm_json(net), f, indent=1)
