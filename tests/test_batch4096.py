"""The 4,096-grid batch (the bench workload) against the FP64 C oracle.

Fixture: tests/golden/batch4096_oracle_sample.npz, written by
tests/golden/make_oracle_fixtures.py -- the oracle's solutions of 127 grids of
the batch (118 evenly spaced in size, the 10 largest; the smallest of them also
checked against the reference's own CPU sources when the fixture was made).

CPU: the fixture still describes gridgen.batch_grids(4096) and the oracle
reproduces it.  GPU: the whole batch in one plan, mixed precision (the bench's
path) and FP64: every grid converges; on the sample the iteration counts equal
the oracle's and |V| is within 1e-8 pu of its solution; mixed and FP64 give
the same iteration count on all 4,096 grids and voltages within 1e-8 pu.

The same batch at load level 1.0 (tests/golden/batch4096_full_load_oracle.npz:
the oracle's iteration count and converged flag for every grid, voltages for
59 of them): the FP32 factorisation must not change the outcome of a heavily
loaded or diverging grid.  12 grids diverge (the two smallest checked against
the reference's own CPU sources when the fixture was made), the others need
8-12 iterations.  On the GPU, in mixed precision and in FP64: converged flags
and iteration counts equal to the oracle's on all 4,096 grids, voltages within
1e-8 pu of the oracle on the sampled converged grids and between the two
precisions on every converged grid."""
import numpy as np
import pytest

from conftest import GOLDEN
from gap_amd import gridgen

FIX = GOLDEN / "batch4096_oracle_sample.npz"
FIX_FULL = GOLDEN / "batch4096_full_load_oracle.npz"
V_TOL = 1e-8


def _fixture():
    return np.load(FIX, allow_pickle=False)


def _checksum(net):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_oracle_fixtures", GOLDEN / "make_oracle_fixtures.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.grid_checksum(net)


@pytest.fixture(scope="module")
def batch():
    return gridgen.batch_grids(4096)


def test_fixture_matches_generator_and_oracle(batch, oracle):
    d = _fixture()
    sizes = np.array([n.n_bus for n in batch])
    assert np.array_equal(d["size"], sizes[d["index"]])
    for k, i in enumerate(d["index"]):
        assert _checksum(batch[i]) == d["checksum"][k], i
    # the oracle reproduces the stored solutions (a few, spread over the sizes)
    for k in np.linspace(0, len(d["index"]) - 1, 6).round().astype(int):
        r = oracle.solve(batch[int(d["index"][k])], tol=1e-8, max_it=100, j_call=-1, traj=False)
        v = d["v"][d["v_off"][k]:d["v_off"][k + 1]]
        assert r["iterations"] == d["iterations"][k]
        assert np.array_equal(r["v"], v)


@pytest.mark.gpu
def test_full_batch_mixed_and_fp64_vs_oracle(device, batch):
    from gap_amd import capi
    d = _fixture()
    idx = d["index"]
    results = {}
    for prec in ("mixed", "fp64"):
        plan = capi.Plan(batch, prec)
        assert plan.schedule == "grid"
        r = plan.solve(1e-8, 100, 100e6)
        results[prec] = r
        assert r["converged"].all(), f"{prec}: {(~r['converged']).sum()} grids did not converge"
        dv = 0.0
        for k, i in enumerate(idx):
            v = d["v"][d["v_off"][k]:d["v_off"][k + 1]]
            assert int(r["iterations"][i]) == int(d["iterations"][k]), (prec, int(i))
            dv = max(dv, float(np.abs(r["v"][i] - v).max()))
        print(f"{prec}: {len(idx)} sampled grids, max|dV| vs FP64 oracle {dv:.3e}, iterations p50 "
              f"{np.percentile(r['iterations'], 50):.0f} p99 {np.percentile(r['iterations'], 99):.0f}")
        assert dv <= V_TOL
        plan.close()
    a, b = results["mixed"], results["fp64"]
    assert np.array_equal(a["iterations"], b["iterations"])
    dv = max(float(np.abs(x - y).max()) for x, y in zip(a["v"], b["v"]))
    print(f"mixed vs fp64 on all {len(batch)} grids: max|dV| {dv:.3e}")
    assert dv <= V_TOL


@pytest.fixture(scope="module")
def batch_full_load():
    return gridgen.batch_grids(4096, load_level=1.0)


def test_full_load_fixture_matches_generator_and_oracle(batch_full_load, oracle):
    d = np.load(FIX_FULL, allow_pickle=False)
    assert np.array_equal(d["size"], [n.n_bus for n in batch_full_load])
    assert np.array_equal(d["checksum"], [_checksum(n) for n in batch_full_load])
    div = np.nonzero(~d["converged"])[0]
    assert len(div) == 12 and (d["iterations"][div] == 100).all()
    # the oracle reproduces a diverging grid, the slowest converging one and a sampled voltage
    slow = int(np.argmax(np.where(d["converged"], d["iterations"], 0)))
    for i in (int(div[0]), slow):
        r = oracle.solve(batch_full_load[i], tol=1e-8, max_it=100, j_call=-1, traj=False)
        assert r["iterations"] == d["iterations"][i] and r["converged"] == d["converged"][i]
        if r["converged"]:
            k = int(np.nonzero(d["sample"] == i)[0][0])
            assert np.array_equal(r["v"], d["v"][d["v_off"][k]:d["v_off"][k + 1]])


@pytest.mark.gpu
def test_full_load_batch_mixed_and_fp64_vs_oracle(device, batch_full_load):
    from gap_amd import capi
    d = np.load(FIX_FULL, allow_pickle=False)
    conv = d["converged"]
    results = {}
    for prec in ("mixed", "fp64"):
        plan = capi.Plan(batch_full_load, prec)
        r = plan.solve(1e-8, 100, 100e6)
        results[prec] = r
        plan.close()
        bad_conv = np.nonzero(r["converged"] != conv)[0]
        bad_it = np.nonzero(r["iterations"] != d["iterations"])[0]
        dv = 0.0
        for k, i in enumerate(d["sample"]):
            v = d["v"][d["v_off"][k]:d["v_off"][k + 1]]
            dv = max(dv, float(np.abs(r["v"][i] - v).max()))
        print(f"{prec}, load level 1.0: {int(r['converged'].sum())} converged, flags differing from the oracle "
              f"{bad_conv.tolist()}, iteration counts differing {bad_it.tolist()}, max|dV| vs oracle on "
              f"{len(d['sample'])} sampled grids {dv:.3e}")
        assert len(bad_conv) == 0 and len(bad_it) == 0
        assert dv <= V_TOL
    a, b = results["mixed"], results["fp64"]
    dv = max(float(np.abs(a["v"][i] - b["v"][i]).max()) for i in np.nonzero(conv)[0])
    print(f"mixed vs fp64 on the {int(conv.sum())} converged grids: max|dV| {dv:.3e}")
    assert dv <= V_TOL
