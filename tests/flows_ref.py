"""Branch flows of a network at a given voltage vector, in numpy: the checker
for the device kernel behind gaphip_branch_flows.  The model is the Y-bus
builder's (reference cpu_admittance_matrix.cpp:121-135): series admittance
y = 1/(r + jx) (0 when |z|^2 <= 1e-12), y + g1 + j b1/2 on each end's
diagonal, -y between the ends, in the network's per-unit system
(impedance base u_rated[0]^2 / S_base); S = V conj(I) in VA, currents in A
(I_base = S_base / (sqrt(3) u_rated[0])).  Pinned against power-grid-model's
own line results (tests/golden/pgm_lines.json) and by the power balance
sum_branches (S_from + S_to) + sum_shunts = sum_buses S.

Test infrastructure only."""
from __future__ import annotations

import numpy as np

FIELDS = ("p_from", "p_to", "q_from", "q_to", "s_from", "s_to", "i_from", "i_from_angle", "i_to", "i_to_angle")


def branch_flows(net, v, base):
    idx = {int(b): i for i, b in enumerate(net.bus_id)}
    ub = float(net.u_rated[0])
    zb, ib = ub * ub / base, base / (np.sqrt(3.0) * ub)
    out = {f: np.zeros(net.n_branch) for f in FIELDS}
    for k in range(net.n_branch):
        f, t = idx.get(int(net.br_from[k]), -1), idx.get(int(net.br_to[k]), -1)
        if not net.br_status[k] or f < 0 or t < 0:
            continue
        r, x = float(net.br_r[k]), float(net.br_x[k])
        y = complex(r, -x) / (r * r + x * x) if r * r + x * x > 1e-12 else 0j
        yd = (y + float(net.br_g[k]) + 0.5j * float(net.br_b[k])) * zb
        yo = -y * zb
        i_f = yd * v[f] + yo * v[t]
        i_t = yd * v[t] + yo * v[f]
        s_f = v[f] * np.conj(i_f) * base
        s_t = v[t] * np.conj(i_t) * base
        vals = (s_f.real, s_t.real, s_f.imag, s_t.imag, abs(s_f), abs(s_t), abs(i_f) * ib, np.angle(i_f),
                abs(i_t) * ib, np.angle(i_t))
        for name, val in zip(FIELDS, vals):
            out[name][k] = val
    return out
