"""Power Grid Model (PGM) JSON input -> network rows, with the conversion the
reference uses on its validated path (the PGM harness feeding
``solve_simple_power_flow``):

  * nodes sorted by id (tests/pgm_validation/json_io/gap_json_parser.py:177-194);
  * line r1, x1 converted to per unit on Z_max = max(u_rated)^2 / S_base and
    b = 2*pi*50*c1 * Z_max (gap_json_parser.py:132-137, 322-337), then back to
    ohm / siemens on Z_0 = u_rated[0]^2 / S_base when building branch rows
    (tests/pgm_validation/test_pgm_validation.py:189-206), from/to given as node
    *indices*;
  * bus rows: type 2 with u_pu = u_ref for a node carrying a source, else type 0
    with u_pu = 1.0; loads summed with a negative sign, p_pu * S_base
    (test_pgm_validation.py:160-187).

Only node / line / sym_load / source are read, as in the reference harness.
The float operations are the same, in the same order, so the rows are
bit-identical to the ones the reference harness would build.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from .network import Network


def read_pgm_input(path, base_power: float = 1e6):
    with open(path) as f:
        doc = json.load(f)
    data = doc.get("data", doc)
    return rows_from_pgm(data, base_power)


def rows_from_pgm(data: dict, base_power: float = 1e6):
    """Returns (bus_rows, branch_rows, base_power) as the reference harness builds them."""
    nodes = data.get("node", [])
    if not nodes:
        raise ValueError("No nodes found in input data")
    lines = data.get("line", [])
    if not lines:
        raise ValueError("No lines found in input data")
    sources = data.get("source", [])
    if not sources:
        raise ValueError("No sources found in input data")
    loads = data.get("sym_load", [])

    node_ids = np.array([n["id"] for n in nodes], dtype=np.int32)
    order = np.argsort(node_ids)
    node_ids = node_ids[order]
    u_rated = np.array([n["u_rated"] for n in nodes], dtype=np.float64)[order]
    id_to_idx = {int(nid): k for k, nid in enumerate(node_ids)}

    from_idx = np.array([id_to_idx[l["from_node"]] for l in lines], dtype=np.int32)
    to_idx = np.array([id_to_idx[l["to_node"]] for l in lines], dtype=np.int32)
    status = np.array([int(l.get("from_status", 1)) * int(l.get("to_status", 1)) for l in lines],
                      dtype=np.int32)
    line_ids = np.array([l["id"] for l in lines], dtype=np.int32)
    r1 = np.array([l["r1"] for l in lines], dtype=np.float64)
    x1 = np.array([l["x1"] for l in lines], dtype=np.float64)
    c1 = np.array([l.get("c1", 0.0) for l in lines], dtype=np.float64)

    load_node = np.array([id_to_idx[l["node"]] for l in loads], dtype=np.int32)
    load_status = np.array([int(l.get("status", 1)) for l in loads], dtype=np.int32)
    p_spec = np.array([l["p_specified"] for l in loads], dtype=np.float64)
    q_spec = np.array([l["q_specified"] for l in loads], dtype=np.float64)

    src_node = np.array([id_to_idx[s["node"]] for s in sources], dtype=np.int32)
    u_ref = np.array([s["u_ref"] for s in sources], dtype=np.float64)

    base_voltage = u_rated.max()
    z_max = base_voltage ** 2 / base_power
    r_pu = r1 / z_max
    x_pu = x1 / z_max
    omega = 2 * np.pi * 50.0
    b_pu = omega * c1 * z_max
    p_pu = p_spec / base_power
    q_pu = q_spec / base_power

    bus_rows = []
    for i in range(len(node_ids)):
        if i in src_node:
            btype, vpu = 2, float(u_ref[src_node == i][0])
        else:
            btype, vpu = 0, 1.0
        p_load = 0.0
        q_load = 0.0
        if i in load_node:
            for k in np.nonzero(load_node == i)[0]:
                if load_status[k]:
                    p_load -= p_pu[k] * base_power
                    q_load -= q_pu[k] * base_power
        bus_rows.append([int(node_ids[i]), float(u_rated[i]), btype, vpu, 0.0, p_load, q_load])
    z0 = (u_rated[0] ** 2) / base_power
    branch_rows = []
    for k in range(len(lines)):
        if status[k]:
            branch_rows.append([int(line_ids[k]), int(from_idx[k]), int(to_idx[k]),
                                float(r_pu[k]) * z0, float(x_pu[k]) * z0, float(b_pu[k]) / z0, 1])
    return bus_rows, branch_rows, base_power


def read_pgm_output_voltages(path):
    """|V| (pu) and the angle of the PGM output dataset as stored, ordered by
    node id.  The reference's fixtures hold the angle in degrees (its
    tests/pgm_validation/run_validation.py:223-226 reads them so; a 2-bus check
    of the angle drop (P x - Q r) / U^2 confirms it)."""
    with open(path) as f:
        doc = json.load(f)
    nodes = sorted(doc["node"], key=lambda n: n["id"])
    return (np.array([n["u_pu"] for n in nodes]), np.array([n["u_angle"] for n in nodes]))


def network_from_pgm(path, base_power: float = 1e6) -> Network:
    bus, br, _ = read_pgm_input(path, base_power)
    return Network.from_rows(bus, br)


__all__ = ["read_pgm_input", "rows_from_pgm", "read_pgm_output_voltages", "network_from_pgm", "Path"]
