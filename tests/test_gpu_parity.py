"""GPU parity: the HIP path (through the C ABI) against the reference goldens
and the FP64 C oracle.

Bars: Y-bus values and the Jacobian pattern bit-exact; Jacobian values to
1e-12 relative (device sin/cos may differ from glibc by an ulp); converged
voltages within 1e-8 pu of the reference with the same iteration count."""
import numpy as np
import pytest

from conftest import GOLDEN, finite_golden_cases, golden_cases, load_golden, max_dv, same_nan
from gap_amd import capi, gridgen
from gap_amd.network import PQ, SLACK

pytestmark = pytest.mark.gpu
CASES = golden_cases()
V_TOL = 1e-8


@pytest.mark.parametrize("name", CASES)
def test_ybus_bit_exact(device, name):
    net, g = load_golden(name)
    rp, ci, val = capi.ybus(net)
    assert np.array_equal(rp, g["y_rp"]) and np.array_equal(ci, g["y_ci"])
    assert np.array_equal(val.view(np.uint64), g["y_val"].view(np.uint64))


def test_ybus_bit_exact_large(device, oracle):
    net = gridgen.large_mesh(10_000, seed=1)
    rp, ci, val = capi.ybus(net)
    orp, oci, oval = oracle.build_ybus(net)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
    assert np.array_equal(val.view(np.uint64), oval.view(np.uint64))


def flat_start(net):
    return np.where(net.bus_type == PQ, 0.98, net.u_pu).astype(complex)


@pytest.mark.parametrize("name", CASES)
def test_jacobian_first_iteration(device, name):
    net, g = load_golden(name)
    plan = capi.Plan([net])
    jrp, jci, jv = plan.jacobian(0, flat_start(net), float(g["base_power"]))
    assert np.array_equal(jrp, g["j_rp"]) and np.array_equal(jci, g["j_ci"])
    assert np.allclose(jv, g["j_val"], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("name", CASES)
def test_mismatch_at_reference_solution(device, oracle, name):
    net, g = load_golden(name)
    plan = capi.Plan([net])
    base = float(g["base_power"])
    f = plan.mismatches(0, g["v"], base)
    y = oracle.build_ybus(net)
    zb = net.u_rated[0] ** 2 / base
    fo = oracle.mismatches(net, g["v"], (y[0], y[1], y[2] * zb)) / 100e6 * (100e6 / base)
    # oracle.mismatches scales S_spec by 100 MVA; rebuild at this base from S
    assert f.shape == fo.shape
    if base == 100e6:
        assert np.allclose(f, fo, rtol=1e-12, atol=1e-14, equal_nan=True)
    if bool(g["converged"]):
        assert np.abs(f).max() < float(g["tolerance"])


@pytest.mark.parametrize("schedule", ["auto", "levels"])
@pytest.mark.parametrize("name", CASES)
def test_solve_fp64_matches_reference(device, name, schedule):
    net, g = load_golden(name)
    plan = capi.Plan([net], schedule=schedule)
    if schedule == "auto":
        assert plan.schedule == ("grid" if plan.stats()["max_front"] <= 64 else "levels")
    r = plan.solve(float(g["tolerance"]), int(g["max_iterations"]), float(g["base_power"]))
    print(name, "max|dV| vs reference:", max_dv(r["v"][0], g["v"]))
    assert int(r["iterations"][0]) == int(g["iterations"])
    assert bool(r["converged"][0]) == bool(g["converged"])
    assert max_dv(r["v"][0], g["v"]) <= V_TOL
    assert same_nan(r["final_mismatch"][0], g["final_mismatch"])
    tr = plan.trajectory(0)
    assert len(tr) == len(g["traj"]) and same_nan(tr, g["traj"])
    assert np.allclose(tr[:4], g["traj"][:4], rtol=1e-8, equal_nan=True)


@pytest.mark.parametrize("schedule", ["auto", "levels"])
@pytest.mark.parametrize("name", CASES)
def test_solve_mixed_matches_reference(device, name, schedule):
    """FP32 factors and solves, FP64 mismatch / update / convergence test: the
    same iteration count as the reference CPU path and |V| within 1e-8 pu of
    its solution on every golden case, the failure cases included."""
    net, g = load_golden(name)
    plan = capi.Plan([net], precision="mixed", schedule=schedule)
    r = plan.solve(float(g["tolerance"]), int(g["max_iterations"]), float(g["base_power"]))
    dv = max_dv(r["v"][0], g["v"])
    print(name, schedule, "mixed: iterations", int(r["iterations"][0]), "reference", int(g["iterations"]),
          "max|dV|", dv)
    assert int(r["iterations"][0]) == int(g["iterations"])
    assert bool(r["converged"][0]) == bool(g["converged"])
    assert same_nan(r["final_mismatch"][0], g["final_mismatch"])
    if bool(g["converged"]):
        assert dv <= V_TOL
    else:
        # not converged: the iterates of the inexact Newton steps drift from the
        # reference's; both end at the cap, NaN at the same buses
        assert same_nan(r["v"][0], g["v"])
    tr = plan.trajectory(0)
    assert len(tr) == len(g["traj"]) and same_nan(tr, g["traj"])
    assert np.allclose(tr[:1], g["traj"][:1], rtol=1e-12, equal_nan=True)


@pytest.mark.parametrize("name", golden_cases("pgm_"))
def test_solve_pgm_fixture_voltages(device, name):
    net, g = load_golden(name)
    r = capi.Plan([net]).solve(1e-8, 20, float(g["base_power"]))
    assert r["converged"][0]
    assert np.abs(np.abs(r["v"][0]) - g["pgm_u_pu"]).max() < 5e-6


@pytest.mark.parametrize("schedule", ["auto", "levels"])
def test_batch_equals_single(device, schedule):
    names = [n for n in CASES if not n.startswith(("nonconvergent", "fail_"))]
    nets = [load_golden(n)[0] for n in names]
    # one tolerance / base for the whole batch: use the mesh settings
    batch = capi.Plan(nets, schedule=schedule).solve(1e-8, 100, 100e6)
    for k, net in enumerate(nets):
        single = capi.Plan([net], schedule=schedule).solve(1e-8, 100, 100e6)
        assert batch["iterations"][k] == single["iterations"][0]
        assert np.array_equal(batch["v"][k], single["v"][0]), names[k]


def test_solve_vs_oracle_midsize(device, oracle):
    for net in (gridgen.small_mesh(1500, seed=21), gridgen.large_mesh(2000, seed=3, n_extra=600,
                                                                       n_trafo=300, n_shunt=60)):
        o = oracle.solve(net, tol=1e-8, max_it=100, j_call=-1)
        r = capi.Plan([net]).solve(1e-8, 100, 100e6)
        assert o["converged"] and r["converged"][0]
        assert int(r["iterations"][0]) == o["iterations"]
        assert np.abs(r["v"][0] - o["v"]).max() <= V_TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_10k_grid_vs_oracle_fixture(device, oracle, precision):
    """10k-bus meshed grid with transformers, shunts and PV buses against the
    FP64 oracle's solution (tests/golden/grid10k_oracle.npz): the same iteration
    count, |V| within 1e-8 pu, the same max|F| trajectory to 1e-6 relative
    (the first entry, at flat start, to 1e-12), and the solution a fixed point
    of the oracle's own mismatch."""
    d = np.load(GOLDEN / "grid10k_oracle.npz", allow_pickle=False)
    net = gridgen.large_mesh(10_000, seed=1)
    import importlib.util
    spec = importlib.util.spec_from_file_location("fx", GOLDEN / "make_oracle_fixtures.py")
    fx = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(fx)
    assert fx.grid_checksum(net) == float(d["checksum"])
    plan = capi.Plan([net], precision=precision)
    r = plan.solve(1e-8, 100, 100e6)
    dv = float(np.abs(r["v"][0] - d["v"]).max())
    print(f"10k grid {precision}: iterations {int(r['iterations'][0])} (oracle {int(d['iterations'])}), "
          f"max|dV| vs oracle {dv:.3e}")
    assert r["converged"][0] and int(r["iterations"][0]) == int(d["iterations"])
    assert dv <= V_TOL
    tr = plan.trajectory(0)
    assert len(tr) == len(d["traj"])
    assert np.isclose(tr[0], d["traj"][0], rtol=1e-12)
    assert np.allclose(tr, d["traj"], rtol=1e-6 if precision == "fp64" else 1e-3, atol=1e-12)
    y = oracle.build_ybus(net)
    f = oracle.mismatches(net, r["v"][0], (y[0], y[1], y[2] * (net.u_rated[0] ** 2 / 100e6))) / 100e6
    assert np.abs(f).max() < 1e-8


def test_100k_grid_mixed_vs_fp64(device, oracle):
    """100k-bus meshed grid (root front above 15k slots): FP32 factors with FP64
    residuals reach the same solution as the FP64 path, and both are fixed
    points of the oracle's own mismatch."""
    net = gridgen.c2_grid()
    a = capi.Plan([net], precision="fp64").solve(1e-8, 100, 100e6)
    b = capi.Plan([net], precision="mixed").solve(1e-8, 100, 100e6)
    assert a["converged"][0] and b["converged"][0]
    dv = np.abs(a["v"][0] - b["v"][0]).max()
    print("100k grid: iterations fp64", int(a["iterations"][0]), "mixed", int(b["iterations"][0]), "max|dV|", dv)
    assert dv <= V_TOL
    assert int(a["iterations"][0]) == int(b["iterations"][0])
    y = oracle.build_ybus(net)
    ypu = (y[0], y[1], y[2] * (net.u_rated[0] ** 2 / 100e6))
    for r in (a, b):
        f = oracle.mismatches(net, r["v"][0], ypu) / 100e6
        assert np.abs(f).max() < 1e-8


def test_network_above_grid_limits_runs_on_levels(device, oracle):
    """A network of more than 65,535 buses (small fronts: a tree with 40 extra
    edges) is outside the grid schedule's list formats: a batch holding it runs
    on the level schedule, asking for the grid schedule is an argument error,
    and every network still reaches a fixed point of the oracle's mismatch."""
    big = gridgen.small_mesh(70_000, seed=5, n_extra=40)
    nets = [gridgen.small_mesh(300, seed=6), big]
    plan = capi.Plan(nets, precision="mixed")
    assert plan.schedule == "levels"
    with pytest.raises(capi.GapHipError):
        plan.set_schedule("grid")
    r = plan.solve(1e-8, 100, 100e6)
    assert r["converged"].all()
    for k, net in enumerate(nets):
        y = oracle.build_ybus(net)
        ypu = (y[0], y[1], y[2] * (net.u_rated[0] ** 2 / 100e6))
        f = oracle.mismatches(net, r["v"][k], ypu) / 100e6
        assert np.abs(f).max() < 1e-8
    small = capi.Plan(nets[:1], precision="mixed")
    assert small.schedule == "grid"
    rs = small.solve(1e-8, 100, 100e6)
    assert int(rs["iterations"][0]) == int(r["iterations"][0])
    assert np.abs(rs["v"][0] - r["v"][0]).max() <= V_TOL


@pytest.mark.parametrize("schedule", ["auto", "levels"])
def test_mixed_precision_close_to_fp64(device, schedule):
    nets = [gridgen.small_mesh(n, seed=s) for n, s in [(200, 0), (800, 4), (3000, 9)]]
    a = capi.Plan(nets, precision="fp64", schedule=schedule).solve(1e-8, 100, 100e6)
    b = capi.Plan(nets, precision="mixed", schedule=schedule).solve(1e-8, 100, 100e6)
    assert b["converged"].all()
    for k in range(len(nets)):
        assert np.abs(a["v"][k] - b["v"][k]).max() <= V_TOL
        assert int(a["iterations"][k]) == int(b["iterations"][k])


def _batch(count, seed):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(50), np.log(3000), count)).astype(int)
    return [gridgen.small_mesh(int(n), seed=seed + k, n_extra=40,
                               load_factor=float(rng.uniform(0.35, 0.65))) for k, n in enumerate(sizes)]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_grid_schedule_matches_levels(device, precision):
    """One workgroup per network in one launch gives the same Newton iterates
    as the level-synchronous schedule: same iteration counts and trajectories,
    voltages equal to rounding (the grid schedule evaluates the Jacobian and
    the update in rectangular form, equal to the polar formulas up to rounding)."""
    nets = _batch(24, 500)
    g = capi.Plan(nets, precision=precision, schedule="grid")
    lv = capi.Plan(nets, precision=precision, schedule="levels")
    assert g.schedule == "grid" and lv.schedule == "levels"
    rg = g.solve(1e-8, 100, 100e6)
    rl = lv.solve(1e-8, 100, 100e6)
    assert rg["converged"].all() and rl["converged"].all()
    assert np.array_equal(rg["iterations"], rl["iterations"])
    for k in range(len(nets)):
        assert np.abs(rg["v"][k] - rl["v"][k]).max() <= (1e-12 if precision == "fp64" else 1e-10)
        tg, tl = g.trajectory(k), lv.trajectory(k)
        assert len(tg) == len(tl) == rg["iterations"][k]
        # the FP32 factors carry the schedules' last-bit Jacobian differences
        # into later iterates at FP32 level
        assert np.allclose(tg, tl, rtol=1e-6 if precision == "fp64" else 1e-3, atol=1e-13)
    # repeated solves on the same plan give the same answer
    rg2 = g.solve(1e-8, 100, 100e6)
    assert all(np.array_equal(a, b) for a, b in zip(rg["v"], rg2["v"]))


def test_grid_bus_order_is_internal(device):
    """The grid schedule works on its own bus order (per network a depth-first
    order of the Y-bus graph) inside the plan; what a caller reads back is in
    the plan's order: the Y-bus after a grid solve (copied from the schedule's
    copy) bit-exact against the standalone assembly and the level schedule's,
    branch updates applied to it, and the adjoint equal to the level
    schedule's to rounding."""
    nets = _batch(10, 777)
    g = capi.Plan(nets, precision="fp64", schedule="grid")
    lv = capi.Plan(nets, precision="fp64", schedule="levels")
    assert g.schedule == "grid" and lv.schedule == "levels"
    rg = g.solve(1e-8, 100, 100e6)
    rl = lv.solve(1e-8, 100, 100e6)
    assert rg["converged"].all() and np.array_equal(rg["iterations"], rl["iterations"])
    for k, net in enumerate(nets):
        assert np.abs(rg["v"][k] - rl["v"][k]).max() <= 1e-12
        rp, ci, yv = g.ybus(k)
        rp2, ci2, yv2 = lv.ybus(k)
        rp3, ci3, yv3 = capi.ybus(net)
        assert np.array_equal(rp, rp2) and np.array_equal(ci, ci2) and np.array_equal(yv, yv2)
        assert np.array_equal(rp, rp3) and np.array_equal(ci, ci3)
    ag, al = g.adjoint(1.0, 0.5), lv.adjoint(1.0, 0.5)
    for key in ag:
        assert np.allclose(ag[key], al[key], rtol=1e-9, atol=1e-12), key
    # a branch update after the grid solve lands in the plan's Y, and the
    # next grid solve assembles the schedule's copy from the new values
    g.update_branches(3, [0, 2], [0.3, 0.2], [0.4, 0.45])
    lv.update_branches(3, [0, 2], [0.3, 0.2], [0.4, 0.45])
    assert np.array_equal(g.ybus(3)[2], lv.ybus(3)[2])
    rg = g.solve(1e-8, 100, 100e6)
    rl = lv.solve(1e-8, 100, 100e6)
    assert np.array_equal(rg["iterations"], rl["iterations"])
    assert max(np.abs(a - b).max() for a, b in zip(rg["v"], rl["v"])) <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_grid_schedule_tiny_networks(device, precision):
    """Networks of two and three buses (one front, or a few) batched with
    larger ones: the grid schedule's own bus order and per-wavefront task
    dependencies hold at the smallest sizes, equal to the level schedule."""
    nets = [gridgen.small_mesh(2, seed=11, n_extra=0), gridgen.small_mesh(3, seed=12, n_extra=0),
            gridgen.small_mesh(3, seed=13, n_extra=1)] + _batch(5, 901)
    g = capi.Plan(nets, precision=precision, schedule="grid")
    lv = capi.Plan(nets, precision=precision, schedule="levels")
    assert g.schedule == "grid" and lv.schedule == "levels"
    rg = g.solve(1e-8, 100, 100e6)
    rl = lv.solve(1e-8, 100, 100e6)
    assert rg["converged"].all() and rl["converged"].all()
    assert np.array_equal(rg["iterations"], rl["iterations"])
    for k in range(len(nets)):
        assert np.abs(rg["v"][k] - rl["v"][k]).max() <= (1e-12 if precision == "fp64" else 1e-10)


def test_grid_tuning_does_not_change_results(device):
    from gap_amd import tuning
    nets = _batch(12, 321) + [gridgen.small_mesh(4000, seed=5, n_extra=40, load_factor=0.5)]
    p = capi.Plan(nets, precision="mixed", schedule="grid")
    ref = p.solve(1e-8, 100, 100e6)
    for gw, bw, work in [(1, 8, 20_000), (4, 4, -1), (8, 2, 0), (2, 8, 40_000)]:
        p.set_tuning(gw, bw, work)
        t = p.tuning()
        assert (t["grid_waves"], t["grid_big_waves"], t["grid_big_work"]) == (gw, bw, work)
        r = p.solve(1e-8, 100, 100e6)
        assert np.array_equal(r["iterations"], ref["iterations"])
        assert all(np.array_equal(a, b) for a, b in zip(r["v"], ref["v"]))
    with pytest.raises(capi.GapHipError):
        p.set_tuning(3, 8, 0)
    out = tuning.autotune(p, candidates={"grid_waves": (1, 2), "grid_big_waves": (8,), "grid_big_work": (-1, 40_000)},
                          repeats=1)
    assert p.tuning() == out["tuning"] and len(out["table"]) == 4
    r = p.solve(1e-8, 100, 100e6)
    assert all(np.array_equal(a, b) for a, b in zip(r["v"], ref["v"]))


def test_level_tuning_does_not_change_results(device):
    """The level schedule's knobs -- graph replay, the large fronts' outer panel
    width and two-level threshold -- change no result bit (a 3,000-bus grid
    with a root front above 128 slots; the threshold is lowered so that its
    fronts take the two-level path)."""
    from gap_amd import tuning
    big = gridgen.large_mesh(3000, seed=2, n_extra=900, n_trafo=300, n_shunt=60)
    p = capi.Plan([big], precision="mixed", schedule="levels")
    ref = p.solve(1e-8, 100, 100e6)
    for knobs in ({"level_graph": 0, "big_outer": 128, "big_two_level_min": 256},
                  {"level_graph": 1, "big_outer": 512, "big_two_level_min": 256},
                  {"level_graph": 1, "big_outer": 64, "big_two_level_min": 0},
                  {"level_graph": 0, "big_outer": 256, "big_two_level_min": 4096}):
        p.set_tuning(**knobs)
        r = p.solve(1e-8, 100, 100e6)
        assert np.array_equal(r["iterations"], ref["iterations"]), knobs
        assert np.abs(r["v"][0] - ref["v"][0]).max() <= 1e-10, knobs
    for bad in ({"big_outer": 100}, {"big_outer": 2048}, {"big_two_level_min": -1}, {"level_graph": 2}):
        with pytest.raises(capi.GapHipError):
            p.set_tuning(**bad)
    out = tuning.autotune(p, candidates={"level_graph": (0, 1), "big_outer": (256,), "big_two_level_min": (4096,)},
                          repeats=1)
    assert len(out["table"]) == 2 and p.tuning() == out["tuning"]


def test_plan_options_and_autotune_plan(device, oracle):
    """Plan-time choices (amalgamation thresholds, the grid tasks' smallest lane
    class and packing) give the same iteration counts and voltages within
    1e-10 of each other and of the oracle within 1e-8; the plan autotuner keeps
    the fastest plan with its options."""
    from gap_amd import tuning
    nets = _batch(8, 77)
    ref = capi.Plan(nets, precision="mixed").solve(1e-8, 100, 100e6)
    for opts in ({"grid_pack": 0}, {"grid_min_class": 1}, {"grid_min_class": 2, "grid_pack": 0},
                 {"relax_abs": 2, "relax_frac": 0.05}, {"relax_abs": 8, "relax_frac": 0.2}):
        p = capi.Plan(nets, precision="mixed", options=opts)
        assert p.options == dict(capi.PlanOptions.DEFAULTS, **opts)
        r = p.solve(1e-8, 100, 100e6)
        assert np.array_equal(r["iterations"], ref["iterations"]), opts
        assert max(np.abs(a - b).max() for a, b in zip(r["v"], ref["v"])) <= 1e-10, opts
    with pytest.raises(capi.GapHipError):
        capi.Plan(nets, options={"grid_min_class": 5})
    out = tuning.autotune_plan(nets, "mixed", plan_candidates=({}, {"grid_pack": 0}),
                               candidates={"grid_waves": (2,), "grid_big_waves": (4,), "grid_big_work": (-1,)},
                               repeats=1)
    assert len(out["table"]) == 2 and out["plan"].options == out["options"]
    r = out["plan"].solve(1e-8, 100, 100e6)
    for k, net in enumerate(nets[:3]):
        o = oracle.solve(net, tol=1e-8, max_it=100, j_call=-1)
        assert int(r["iterations"][k]) == o["iterations"] and np.abs(r["v"][k] - o["v"]).max() <= V_TOL


def test_grid_schedule_vs_oracle(device, oracle):
    nets = _batch(6, 900)
    r = capi.Plan(nets, schedule="grid").solve(1e-8, 100, 100e6)
    for k, net in enumerate(nets):
        o = oracle.solve(net, tol=1e-8, max_it=100, j_call=-1)
        assert o["converged"] and r["converged"][k]
        assert int(r["iterations"][k]) == o["iterations"]
        assert np.abs(r["v"][k] - o["v"]).max() <= V_TOL


def test_grid_schedule_limits(device):
    # iteration cap reached: same bookkeeping as the level schedule
    nets = _batch(4, 77)
    for sched in ("grid", "levels"):
        r = capi.Plan(nets, schedule=sched).solve(1e-8, 2, 100e6)
        assert not r["converged"].any() and (r["iterations"] == 2).all() and (r["final_mismatch"] > 1e-8).all()
        r0 = capi.Plan(nets, schedule=sched).solve(1e-8, 0, 100e6)
        assert not r0["converged"].any() and (r0["iterations"] == 0).all()
    # fronts above 64 slots cannot use the grid schedule
    big = gridgen.large_mesh(3000, seed=2, n_extra=900, n_trafo=300, n_shunt=60)
    p = capi.Plan([big])
    if p.stats()["max_front"] > 64:
        assert p.schedule == "levels"
        with pytest.raises(capi.GapHipError) as e:
            p.set_schedule("grid")
        assert e.value.code == capi.E_ARG


@pytest.mark.parametrize("n", [10, 60])
def test_lu_golden(device, n):
    d = np.load(GOLDEN / f"lu_dd{n}.npz")
    lu = capi.LU()
    lu.factorize(n, n, d["rp"], d["ci"], d["val"])
    assert np.abs(lu.solve(d["b"]) - d["x"]).max() < 1e-10


def test_lu_random_and_contract(device):
    rng = np.random.default_rng(3)
    lu = capi.LU()
    for n in (1, 7, 90, 400):
        a = (rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n))) * (rng.uniform(0, 1, (n, n)) < 0.05)
        a = a + a.T.conj() * 0.0 + np.diag(np.abs(a).sum(1) + 1.0)
        rows, cols = np.nonzero(a)
        rp = np.searchsorted(rows, np.arange(n + 1))
        lu.factorize(n, n, rp, cols, a[rows, cols])
        for _ in range(2):
            b = rng.normal(size=n) + 1j * rng.normal(size=n)
            assert np.allclose(lu.solve(b), np.linalg.solve(a, b), rtol=1e-10, atol=1e-12)
    with pytest.raises(capi.GapHipError) as e:
        lu.solve(np.ones(3))
    assert e.value.code == capi.E_ARG
    with pytest.raises(capi.GapHipError) as e:
        capi.LU().solve(np.ones(3))
    assert e.value.code == capi.E_STATE
    with pytest.raises(capi.GapHipError) as e:
        capi.LU().factorize(2, 3, [0, 1, 2], [0, 1], [1.0, 1.0])
    assert e.value.code == capi.E_ARG


def test_pybind_simple_api(device):
    import gap_amd.gap_solver as gs
    net, g = load_golden("pgm_radial_3feeder_8nodepf")
    bus, br = net.to_rows()
    out = gs.solve_simple_power_flow(bus, br, tolerance=1e-8, max_iterations=20,
                                     base_power=float(g["base_power"]), backend="gpu")
    meta = out[-1]
    assert meta[0] == 1.0 and int(meta[1]) == int(g["iterations"])
    v = np.array([r[0] + 1j * r[1] for r in out[:-1]])
    assert np.abs(v - g["v"]).max() <= V_TOL
    assert np.allclose([r[2] for r in out[:-1]], np.abs(v))
    with pytest.raises(RuntimeError, match="not available"):
        gs.solve_simple_power_flow(bus, br, backend="cpu")
    with pytest.raises(RuntimeError, match="at least 7"):
        gs.solve_simple_power_flow([[0, 1e4, 2]], br, backend="hip")


def _csr(a):
    rows, cols = np.nonzero(a)
    rp = np.searchsorted(rows, np.arange(a.shape[0] + 1)).astype(np.int32)
    return rp, cols.astype(np.int32), a[rows, cols]


def _sparse_matrix(a):
    import gap_amd.gap_solver as gs
    rp, ci, val = _csr(a)
    m = gs.SparseMatrix()
    m.row_ptr, m.col_idx, m.values = rp.tolist(), ci.tolist(), [complex(x) for x in val]
    m.num_rows = m.num_cols = a.shape[0]
    m.nnz = len(ci)
    return m


def _non_dominant(rng, n, complex_values=True):
    """Random sparse matrix with no diagonal dominance: entries U[-1, 1] (complex
    or real) at density ~10%, plus one entry of magnitude 0.1-1 per row at a
    random column (so it is structurally non-singular, with most diagonal
    entries zero)."""
    a = rng.uniform(-1, 1, (n, n)) * (rng.uniform(0, 1, (n, n)) < 0.1)
    if complex_values:
        a = a + 1j * rng.uniform(-1, 1, (n, n)) * (a != 0)
    p = rng.permutation(n)
    a = a.astype(complex)
    a[np.arange(n), p] += rng.uniform(0.1, 1.0, n) * (np.exp(1j * rng.uniform(0, 6.3, n)) if complex_values else 1)
    return a


def test_lu_row_exchanges(device):
    """Matrices the reference LU solves only with row exchanges
    (cpu_lu_solver.cpp:469-498): the oracle's zero-diagonal 3x3 and random
    non-dominant sparse matrices (real, as the Newton Jacobian, and complex),
    through the C ABI and through the ILUSolver plugin class, against LAPACK."""
    import gap_amd.gap_solver as gs
    rng = np.random.default_rng(17)
    cases = [np.array([[0, 2, 0], [1, 0, 0], [0, 1, 3]], complex)]
    for n, cplx in ((6, False), (40, True), (150, False), (400, True), (900, False)):
        while True:
            a = _non_dominant(rng, n, cplx)
            if np.linalg.cond(a) < 1e7:
                break
        cases.append(a)
    lu = capi.LU()
    plug = gs.create_lu_solver(gs.BackendType.GPU_HIP)
    assert isinstance(plug, gs.HipLUSolver) and plug.get_backend_type() == gs.BackendType.GPU_HIP
    for a in cases:
        n = a.shape[0]
        assert np.count_nonzero(np.diag(a)) < n  # not solvable with diagonal pivots alone
        lu.factorize(n, n, *_csr(a))
        assert plug.factorize(_sparse_matrix(a)) and plug.is_factorized()
        for _ in range(2):
            b = rng.normal(size=n) + 1j * rng.normal(size=n)
            x = np.linalg.solve(a, b)
            x1 = lu.solve(b)
            x2 = np.array(plug.solve(list(b)))
            scale = np.abs(x).max()
            print(f"n={n}: max|x - x_lapack| / max|x| = {np.abs(x1 - x).max() / scale:.2e}")
            assert np.abs(x1 - x).max() <= 1e-9 * scale
            assert np.array_equal(x1, x2)


def test_lu_singular_matrices(device):
    """Numerically singular (rank 1) and structurally singular (an empty
    column) matrices: factorize fails, as the reference's does
    (cpu_lu_solver.cpp:486-491, 663-668)."""
    import gap_amd.gap_solver as gs
    plug = gs.HipLUSolver()
    for a in (np.array([[1, 2], [2, 4]], complex),
              np.array([[1, 0, 2], [3, 0, 1], [0, 0, 5]], complex)):
        with pytest.raises(capi.GapHipError) as e:
            capi.LU().factorize(a.shape[0], a.shape[0], *_csr(a))
        assert e.value.code == capi.E_SINGULAR
        assert not plug.factorize(_sparse_matrix(a)) and not plug.is_factorized()


def test_level_schedule_graph_replay_equals_direct_launches(device):
    """The level schedule's Newton step replayed from its captured HIP graph
    gives the same iterates, bit for bit, as launching its kernels one by one:
    on a grid with fronts above 128 slots (blocked factorisation), on a batch
    of small grids, in both precisions, over consecutive solves (every replay
    is checked through the solution it leads to), and with eager launches and
    host synchronisation between the replays."""
    big = gridgen.large_mesh(3000, seed=2, n_extra=900, n_trafo=300, n_shunt=60)
    for nets in ([big], _batch(10, 31)):
        for precision in ("fp64", "mixed"):
            runs = {}
            for graph in (False, True):
                p = capi.Plan(nets, precision=precision, schedule="levels")
                p.set_graph(graph)
                runs[graph] = [p.solve(1e-8, 100, 100e6) for _ in range(3)]
                runs[graph].append(p.solve(1e-6, 100, 100e6))  # other options, same graph
                traj = [p.trajectory(k) for k in range(len(nets))]
                runs[graph].append(traj)
            for a, b in zip(runs[False][:4], runs[True][:4]):
                assert np.array_equal(a["iterations"], b["iterations"])
                assert all(np.array_equal(x, y) for x, y in zip(a["v"], b["v"]))
            assert all(np.array_equal(x, y) for x, y in zip(runs[False][4], runs[True][4]))
            assert runs[True][0]["converged"].all()
