"""gaphip_power_mismatches_noalloc: the mismatch evaluation of a caller-given
admittance matrix with every byte of device memory in the library's static
per-GPU array, against gaphip_power_mismatches (the same evaluation with a
device region allocated per call), through the C ABI.  The two must agree bit
for bit on consecutive calls, the second with a smaller network (the array keeps
the first call's contents), and both must match the oracle's public mismatch."""
import ctypes as C

import numpy as np
import pytest

from conftest import load_golden
from gap_amd import capi, gridgen

pytestmark = pytest.mark.gpu


def call(name, net, v, y, base=100e6, rows=None):
    s, keep = capi.make_net(net)
    rp, ci, val = y
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    val = np.asarray(val, np.complex128)
    yri = np.ascontiguousarray(np.stack([val.real, val.imag], 1).ravel())
    v = np.asarray(v, np.complex128)
    vri = np.ascontiguousarray(np.stack([v.real, v.imag], 1).ravel())
    cap = 2 * net.n_bus + 1
    f = np.zeros(cap)
    m = C.c_int(-1)
    nnz = min(len(ci), len(val))
    rows = len(rp) - 1 if rows is None else rows
    rc = getattr(capi.lib(), name)(C.byref(s), vri.ctypes.data_as(capi._d), rows, rp.ctypes.data_as(capi._i),
                                   ci.ctypes.data_as(capi._i), yri.ctypes.data_as(capi._d), nnz, base,
                                   f.ctypes.data_as(capi._d), cap, C.byref(m))
    return rc, f[:max(m.value, 0)].copy()


def both(net, v, y, **kw):
    rc1, f1 = call("gaphip_power_mismatches", net, v, y, **kw)
    rc2, f2 = call("gaphip_power_mismatches_noalloc", net, v, y, **kw)
    assert rc1 == 0 and rc2 == 0, (rc1, rc2, capi.lib().gaphip_last_error())
    assert f1.shape == f2.shape
    assert np.array_equal(f1.view(np.uint64), f2.view(np.uint64))
    return f2


def test_noalloc_equals_allocating_entry_point(device, oracle):
    rng = np.random.default_rng(3)
    big = gridgen.large_mesh(10_000, seed=1)
    vb = (1 + 0.02 * rng.standard_normal(big.n_bus)) * np.exp(0.05j * rng.standard_normal(big.n_bus))
    yb = oracle.build_ybus(big)
    small, g = load_golden("special_pv_shunt_parallel")  # PV buses, shunts, a LOADGEN appliance
    ys = oracle.build_ybus(small)
    # two calls in a row, the second on a smaller network
    for net, v, y in ((big, vb, yb), (small, g["v"], ys)):
        f = both(net, v, y)
        fo = oracle.mismatches(net, v, y) / 100e6  # the public mismatch in W, per unit here
        assert f.shape == fo.shape
        assert np.allclose(f, fo, rtol=1e-12, atol=1e-12 * np.abs(fo).max())
    # the reference's bounds rules: rows past y_rows give zero current, columns
    # past n_bus are skipped, a row stops at the end of the arrays
    rp, ci, val = ys
    ci2 = ci.copy()
    ci2[::7] = small.n_bus + 3
    f_cut = both(small, g["v"], (rp, ci2, val), rows=small.n_bus // 2)
    assert np.isfinite(f_cut).all()
    # empty matrix: F = -S_spec
    f0 = both(small, g["v"], (np.zeros(small.n_bus + 1, np.int32), np.zeros(0, np.int32), np.zeros(0)))
    assert np.allclose(f0, oracle.mismatches(small, g["v"], (np.zeros(small.n_bus + 1, np.int32),
                                                              np.zeros(0, np.int32), np.zeros(0))) / 100e6)


def test_noalloc_limits_and_errors(device):
    net = gridgen.small_mesh(40_000, seed=2, n_extra=10)
    v = np.ones(net.n_bus, complex)
    y = (np.zeros(net.n_bus + 1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    rc, _ = call("gaphip_power_mismatches_noalloc", net, v, y)
    assert rc == capi.E_CAPACITY
    assert b"32768" in capi.lib().gaphip_last_error()
    rc, f = call("gaphip_power_mismatches", net, v, y)
    assert rc == 0 and len(f) == 2 * (net.n_bus - 1)
    small = gridgen.small_mesh(30, seed=1)
    rp = np.array([0, 2] + [2] * small.n_bus, np.int32)
    for name in ("gaphip_power_mismatches", "gaphip_power_mismatches_noalloc"):
        rc, _ = call(name, small, np.ones(small.n_bus, complex), (rp, np.array([0, -1], np.int32), np.ones(2)))
        assert rc == capi.E_ARG
    # after an error the next call is whole again
    both(small, np.ones(small.n_bus, complex), (rp, np.array([0, 1], np.int32), np.ones(2) * (1 - 2j)))
