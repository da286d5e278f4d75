"""Branch flows (p, q, s, i at both ends of every branch), the BranchData
result fields the reference declares (include/gap/core/types.h:233-243) but
never fills.

CPU: the numpy checker (tests/flows_ref.py) fed power-grid-model's own
voltages reproduces power-grid-model's own line results (its output fixtures,
tests/golden/pgm_lines.json), and flows satisfy the power balance
sum_branches (S_from + S_to) + sum_shunts = sum_buses V conj(Y V).
GPU: gaphip_branch_flows against the checker at the device solution."""
import json
from pathlib import Path

import numpy as np
import pytest

from conftest import finite_golden_cases, golden_cases, load_golden
from flows_ref import FIELDS, branch_flows
from gap_amd import capi, gridgen
from gap_amd.network import SHUNT

PGM_LINES = json.loads((Path(__file__).parent / "golden" / "pgm_lines.json").read_text())
PGM_FIELDS = ("p_from", "q_from", "i_from", "s_from", "p_to", "q_to", "i_to", "s_to")


def rel_max(a, b):
    return float(np.max(np.abs(a - b)) / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("case", sorted(PGM_LINES))
def test_flow_model_matches_pgm_line_results(case):
    """PGM's fixtures hold u_angle in degrees (a 2-bus hand check of the angle
    drop, (P x - Q r) / U^2, fixes the unit)."""
    net, g = load_golden(case)
    v = g["pgm_u_pu"] * np.exp(1j * np.deg2rad(g["pgm_u_angle"]))
    fl = branch_flows(net, v, float(g["base_power"]))
    pos = {int(b): k for k, b in enumerate(net.br_id)}
    ks = [pos[i] for i in PGM_LINES[case]["id"]]
    for f in PGM_FIELDS:
        assert rel_max(fl[f][ks], np.array(PGM_LINES[case][f])) < 1e-10, f


def _bus_power(oracle, net, v, base):
    rp, ci, val = oracle.build_ybus(net)
    zb = net.u_rated[0] ** 2 / base
    i = np.zeros(net.n_bus, complex)
    for r in range(net.n_bus):
        for q in range(rp[r], rp[r + 1]):
            i[r] += val[q] * zb * v[ci[q]]
    return v * np.conj(i) * base


@pytest.mark.parametrize("case", ["mesh200_seed0", "special_pv_shunt_parallel", "pgm_radial_10feeder_10nodepf"])
def test_flows_power_balance(oracle, case):
    net, g = load_golden(case)
    base = float(g["base_power"])
    v = g["v"]
    fl = branch_flows(net, v, base)
    s_br = (fl["p_from"] + fl["p_to"]).sum() + 1j * (fl["q_from"] + fl["q_to"]).sum()
    idx = {int(b): i for i, b in enumerate(net.bus_id)}
    zb = net.u_rated[0] ** 2 / base
    s_sh = 0j
    for a in range(net.n_app):
        if net.app_type[a] == SHUNT and int(net.app_node[a]) in idx:
            vi = v[idx[int(net.app_node[a])]]
            s_sh += vi * np.conj((net.app_g[a] + 1j * net.app_b[a]) * zb * vi) * base
    s_bus = _bus_power(oracle, net, v, base).sum()
    scale = np.abs(fl["s_from"]).sum()
    assert abs(s_br + s_sh - s_bus) < 1e-10 * scale


def test_out_of_service_branch_is_zero():
    net, g = load_golden("special_pv_shunt_parallel")
    off = np.nonzero(net.br_status == 0)[0]
    assert len(off) > 0
    fl = branch_flows(net, g["v"], float(g["base_power"]))
    for f in FIELDS:
        assert np.all(fl[f][off] == 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("name", finite_golden_cases())
def test_gpu_branch_flows_vs_checker(device, name):
    net, g = load_golden(name)
    base = float(g["base_power"])
    plan = capi.Plan([net])
    r = plan.solve(float(g["tolerance"]), int(g["max_iterations"]), base)
    fl = plan.branch_flows()
    ref = branch_flows(net, r["v"][0], base)
    for f in FIELDS:
        if f.endswith("angle"):
            # angles of near-zero currents are not meaningful; compare where |i| is not tiny
            mag = ref[f.replace("_angle", "")]
            m = mag > 1e-9 * max(mag.max(), 1e-300)
            d = np.angle(np.exp(1j * (fl[f][m] - ref[f][m])))
            assert np.all(np.abs(d) < 1e-12), (name, f)
        else:
            assert rel_max(fl[f], ref[f]) < 1e-12, (name, f)
    if name in PGM_LINES:
        pos = {int(b): k for k, b in enumerate(net.br_id)}
        ks = [pos[i] for i in PGM_LINES[name]["id"]]
        for f in PGM_FIELDS:
            # the device solution and power-grid-model's agree to the solvers'
            # tolerance (1e-8), which moves flows by ~1e-7 of their scale
            assert rel_max(fl[f][ks], np.array(PGM_LINES[name][f])) < 1e-6, (name, f)


@pytest.mark.gpu
def test_gpu_branch_flows_batch_order_and_10k(device):
    """Instance-concatenated order over a batch of uneven grids, and the 10k-bus
    grid with transformers and shunts."""
    base = 100e6
    nets = [gridgen.small_mesh(n, seed=s, n_extra=10) for n, s in ((60, 1), (150, 2), (90, 3))]
    nets.append(gridgen.large_mesh(10_000, seed=1))
    plan = capi.Plan(nets)
    r = plan.solve(1e-10, 60, base)
    assert r["converged"].all()
    fl = plan.branch_flows()
    off = 0
    for k, net in enumerate(nets):
        ref = branch_flows(net, r["v"][k], base)
        for f in ("p_from", "p_to", "q_from", "q_to", "s_from", "s_to", "i_from", "i_to"):
            assert rel_max(fl[f][off:off + net.n_branch], ref[f]) < 1e-12, (k, f)
        off += net.n_branch
    assert off == len(fl["p_from"])


@pytest.mark.gpu
def test_gpu_branch_flows_before_solve_is_state_error(device):
    plan = capi.Plan([gridgen.small_mesh(30, seed=0)])
    with pytest.raises(capi.GapHipError) as e:
        plan.branch_flows()
    assert e.value.code == capi.E_STATE
