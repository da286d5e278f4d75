"""Flat (structure-of-arrays) power-flow network description.

The fields mirror the reference data model (include/gap/core/types.h):
``BusData`` (types.h:204-219), ``BranchData`` (types.h:224-274) and
``ApplianceData`` (types.h:284-310), restricted to what the Newton-Raphson
path reads.  Bus and branch endpoints are *ids*, mapped to positions through
a dense id table exactly like cpu_admittance_matrix.cpp:56-73.

Units follow the reference C++ model: volts, ohms, siemens, watts and VAr,
with powers in generator sign (a load is a negative injection,
tests/pgm_validation/test_pgm_validation.py:176-184).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

PQ, PV, SLACK = 0, 1, 2                 # types.h:30-34
SOURCE, LOADGEN, SHUNT = 0, 1, 2        # types.h:39-43


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


@dataclass
class Network:
    bus_id: np.ndarray
    u_rated: np.ndarray
    bus_type: np.ndarray
    u_pu: np.ndarray
    p_w: np.ndarray
    q_var: np.ndarray
    br_id: np.ndarray
    br_from: np.ndarray
    br_to: np.ndarray
    br_status: np.ndarray
    br_r: np.ndarray
    br_x: np.ndarray
    br_g: np.ndarray
    br_b: np.ndarray
    app_id: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    app_node: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    app_status: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    app_type: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    app_p: np.ndarray = field(default_factory=lambda: np.zeros(0))
    app_q: np.ndarray = field(default_factory=lambda: np.zeros(0))
    app_g: np.ndarray = field(default_factory=lambda: np.zeros(0))
    app_b: np.ndarray = field(default_factory=lambda: np.zeros(0))

    def __post_init__(self):
        for k in ("bus_id", "bus_type", "br_id", "br_from", "br_to", "br_status",
                  "app_id", "app_node", "app_status", "app_type"):
            setattr(self, k, _i32(getattr(self, k)))
        for k in ("u_rated", "u_pu", "p_w", "q_var", "br_r", "br_x", "br_g", "br_b",
                  "app_p", "app_q", "app_g", "app_b"):
            setattr(self, k, _f64(getattr(self, k)))
        nb, nr, na = len(self.bus_id), len(self.br_id), len(self.app_id)
        for k in ("u_rated", "bus_type", "u_pu", "p_w", "q_var"):
            if len(getattr(self, k)) != nb:
                raise ValueError(f"bus field {k} has length {len(getattr(self, k))}, expected {nb}")
        for k in ("br_from", "br_to", "br_status", "br_r", "br_x", "br_g", "br_b"):
            if len(getattr(self, k)) != nr:
                raise ValueError(f"branch field {k} has length {len(getattr(self, k))}, expected {nr}")
        for k in ("app_node", "app_status", "app_type", "app_p", "app_q", "app_g", "app_b"):
            if len(getattr(self, k)) != na:
                raise ValueError(f"appliance field {k} has length {len(getattr(self, k))}, expected {na}")

    @property
    def n_bus(self) -> int:
        return len(self.bus_id)

    @property
    def n_branch(self) -> int:
        return len(self.br_id)

    @property
    def n_app(self) -> int:
        return len(self.app_id)

    def copy(self) -> "Network":
        return Network(**{k: np.array(v, copy=True) for k, v in self.__dict__.items()})

    # ---- the Python row API of the reference (solver_bindings.cpp:57-111) ----
    @classmethod
    def from_rows(cls, bus_data, branch_data) -> "Network":
        """Build from ``solve_simple_power_flow`` rows.

        bus row:    [id, u_rated, type, u_pu, u_angle, p_inj, q_inj]  (>= 7 columns)
        branch row: [id, from_id, to_id, r_ohm, x_ohm, (b_S)]         (>= 5 columns)
        Raises ValueError with the reference's messages for short rows
        (solver_bindings.cpp:69-73, 91-95).
        """
        nb = len(bus_data)
        bus = np.zeros((nb, 7))
        for i, row in enumerate(bus_data):
            if len(row) < 7:
                raise ValueError("Bus data must have at least 7 elements: [id, u_rated, bus_type, "
                                 "u_pu, u_angle, p, q]")
            bus[i] = [float(v) for v in row[:7]]
        nr = len(branch_data)
        br = np.zeros((nr, 6))
        for i, row in enumerate(branch_data):
            if len(row) < 5:
                raise ValueError("Branch data must have at least 5 elements: [id, from_bus, to_bus, r, x]")
            vals = [float(v) for v in row[:6]]
            if len(vals) == 5:
                vals.append(0.0)
            br[i] = vals
        return cls(
            bus_id=bus[:, 0].astype(np.int64).astype(np.int32), u_rated=bus[:, 1],
            bus_type=bus[:, 2].astype(np.int64).astype(np.int32), u_pu=bus[:, 3],
            p_w=bus[:, 5], q_var=bus[:, 6],
            br_id=br[:, 0].astype(np.int64).astype(np.int32),
            br_from=br[:, 1].astype(np.int64).astype(np.int32),
            br_to=br[:, 2].astype(np.int64).astype(np.int32),
            br_status=np.ones(nr, np.int32), br_r=br[:, 3], br_x=br[:, 4],
            br_g=np.zeros(nr), br_b=br[:, 5])

    def to_rows(self):
        bus = [[int(self.bus_id[i]), float(self.u_rated[i]), int(self.bus_type[i]),
                float(self.u_pu[i]), 0.0, float(self.p_w[i]), float(self.q_var[i])]
               for i in range(self.n_bus)]
        br = [[int(self.br_id[k]), int(self.br_from[k]), int(self.br_to[k]),
               float(self.br_r[k]), float(self.br_x[k]), float(self.br_b[k])]
              for k in range(self.n_branch)]
        return bus, br
