"""Incremental admittance update on the device (gaphip_plan_update_branches)
and the plan's Y-bus read back (gaphip_plan_get_ybus): after r1 / x1 of some
branches change, the plan's Y-bus equals the oracle's assembly of the changed
network (times the impedance base) bit for bit, and equals a fresh plan's;
the next solve equals a fresh plan's solve bit for bit."""
import numpy as np
import pytest

from conftest import load_golden
from gap_amd import capi, gridgen

pytestmark = pytest.mark.gpu


def oracle_ypu(oracle, net, base):
    rp, ci, val = oracle.build_ybus(net)
    zb = net.u_rated[0] ** 2 / base
    return rp, ci, val * zb


@pytest.mark.parametrize("name", ["special_pv_shunt_parallel", "mesh200_seed0"])
def test_update_matches_fresh_assembly(device, oracle, name):
    net, g = load_golden(name)
    base = float(g["base_power"])
    plan = capi.Plan([net])
    plan.solve(1e-10, 60, base)
    rp, ci, y = plan.ybus(0)
    orp, oci, oy = oracle_ypu(oracle, net, base)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci) and np.array_equal(y, oy)
    rng = np.random.default_rng(3)
    changed = net.copy()
    idx = rng.choice(net.n_branch, min(6, net.n_branch), replace=False)
    r = net.br_r[idx] * rng.uniform(0.5, 1.5, len(idx))
    x = net.br_x[idx] * rng.uniform(0.5, 1.5, len(idx))
    changed.br_r[idx] = r
    changed.br_x[idx] = x
    plan.update_branches(0, idx, r, x)
    _, _, y2 = plan.ybus(0)
    _, _, oy2 = oracle_ypu(oracle, changed, base)
    assert np.array_equal(y2, oy2)
    assert not np.array_equal(y2, y)
    fresh = capi.Plan([changed])
    rf = fresh.solve(1e-10, 60, base)
    assert np.array_equal(fresh.ybus(0)[2], y2)
    r2 = plan.solve(1e-10, 60, base)
    assert np.array_equal(r2["v"][0], rf["v"][0]) and int(r2["iterations"][0]) == int(rf["iterations"][0])


def test_update_in_a_batch_touches_only_its_instance(device, oracle):
    base = 100e6
    nets = [gridgen.small_mesh(n, seed=s, n_extra=8) for n, s in ((40, 1), (70, 2), (55, 3))]
    plan = capi.Plan(nets)
    plan.solve(1e-10, 60, base)
    before = [plan.ybus(k)[2].copy() for k in range(3)]
    plan.update_branches(1, [0, 5], [0.3, 0.2], [0.4, 0.45])
    after = [plan.ybus(k)[2] for k in range(3)]
    assert np.array_equal(after[0], before[0]) and np.array_equal(after[2], before[2])
    changed = nets[1].copy()
    changed.br_r[[0, 5]] = [0.3, 0.2]
    changed.br_x[[0, 5]] = [0.4, 0.45]
    assert np.array_equal(after[1], oracle_ypu(oracle, changed, base)[2])


def test_errors(device):
    plan = capi.Plan([gridgen.small_mesh(30, seed=0)])
    with pytest.raises(capi.GapHipError) as e:
        plan.ybus(0)
    assert e.value.code == capi.E_STATE
    plan.solve(1e-8, 50, 100e6)
    with pytest.raises(capi.GapHipError) as e:
        plan.update_branches(0, [10_000], [0.1], [0.1])
    assert e.value.code == capi.E_ARG
    with pytest.raises(capi.GapHipError):
        plan.update_branches(3, [0], [0.1], [0.1])
