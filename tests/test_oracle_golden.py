"""The FP64 C oracle against the reference's own CPU path (golden fixtures
written by tests/golden/make_golden.py from the reference sources) and against
what the mathematics fixes independently of both."""
import numpy as np
import pytest

from conftest import golden_cases, load_golden, max_dv, same_nan, GOLDEN
from gap_amd import gridgen
from gap_amd.network import Network, PQ, SLACK

CASES = golden_cases()


@pytest.mark.parametrize("name", CASES)
def test_ybus_bit_exact(oracle, name):
    net, g = load_golden(name)
    rp, ci, val = oracle.build_ybus(net)
    assert np.array_equal(rp, g["y_rp"]) and np.array_equal(ci, g["y_ci"])
    assert np.array_equal(val.view(np.uint64), g["y_val"].view(np.uint64))


@pytest.mark.parametrize("name", CASES)
def test_nr_matches_reference(oracle, name):
    net, g = load_golden(name)
    r = oracle.solve(net, tol=float(g["tolerance"]), max_it=int(g["max_iterations"]),
                     base_power=float(g["base_power"]), j_call=0)
    assert r["iterations"] == int(g["iterations"])
    assert r["converged"] == bool(g["converged"])
    assert max_dv(r["v"], g["v"]) <= 1e-10
    assert same_nan(r["final_mismatch"], g["final_mismatch"])
    # Jacobian of the first iteration: pattern bit-exact, values as computed
    jrp, jci, jv = r["J"]
    assert np.array_equal(jrp, g["j_rp"]) and np.array_equal(jci, g["j_ci"])
    assert np.allclose(jv, g["j_val"], rtol=1e-13, atol=1e-13)
    # the damped iteration contracts |F| by ~0.1; early entries agree tightly
    k = min(4, len(g["traj"]))
    assert np.allclose(r["traj"][:k], g["traj"][:k], rtol=1e-9, equal_nan=True)
    assert len(r["traj"]) == len(g["traj"]) and same_nan(r["traj"], g["traj"])


@pytest.mark.parametrize("name", golden_cases("pgm_"))
def test_pgm_fixture_voltages(oracle, name):
    """|V| within 5 micro-pu of Power Grid Model's own output.json, the
    reference's acceptance bar (tests/pgm_validation/test_pgm_validation.py:38-39)."""
    net, g = load_golden(name)
    r = oracle.solve(net, tol=1e-8, max_it=20, base_power=float(g["base_power"]), j_call=-1)
    assert r["converged"]
    assert np.abs(np.abs(r["v"]) - g["pgm_u_pu"]).max() < 5e-6


@pytest.mark.parametrize("n", [10, 60])
def test_lu_against_reference_lu(oracle, n):
    d = np.load(GOLDEN / f"lu_dd{n}.npz")
    rc, x = oracle.lu_solve(n, d["rp"], d["ci"], d["val"], d["b"])
    assert rc == 0
    assert np.abs(x - d["x"]).max() < 1e-10


def _dense(n, rp, ci, val):
    a = np.zeros((n, n), complex)
    for i in range(n):
        for q in range(rp[i], rp[i + 1]):
            a[i, ci[q]] += val[q]
    return a


def test_lu_random_vs_lapack(oracle):
    rng = np.random.default_rng(7)
    for n in (1, 5, 40, 200):
        a = (rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n))) * (rng.uniform(0, 1, (n, n)) < 0.1)
        a += np.diag(rng.uniform(0.5, 2.0, n))
        rows, cols = np.nonzero(a)
        rp = np.searchsorted(rows, np.arange(n + 1)).astype(np.int32)
        b = rng.normal(size=n) + 1j * rng.normal(size=n)
        rc, x = oracle.lu_solve(n, rp, cols.astype(np.int32), a[rows, cols], b)
        assert rc == 0
        assert np.allclose(x, np.linalg.solve(a, b), rtol=1e-10, atol=1e-12)


def test_lu_needs_pivoting(oracle):
    # zero diagonal: only row exchanges make this solvable
    a = np.array([[0, 2, 0], [1, 0, 0], [0, 1, 3]], complex)
    rows, cols = np.nonzero(a)
    rp = np.searchsorted(rows, np.arange(4)).astype(np.int32)
    b = np.array([2, 1, 4], complex)
    rc, x = oracle.lu_solve(3, rp, cols.astype(np.int32), a[rows, cols], b)
    assert rc == 0 and np.allclose(x, np.linalg.solve(a, b))


def test_lu_singular_reports(oracle):
    a = np.array([[1, 2], [2, 4]], complex)
    rows, cols = np.nonzero(a)
    rp = np.searchsorted(rows, np.arange(3)).astype(np.int32)
    rc, _ = oracle.lu_solve(2, rp, cols.astype(np.int32), a[rows, cols], np.ones(2))
    assert rc == 1


def two_bus(r=0.3, x=0.4, b=2e-4, p=-1e5, q=-3e4, sh=None):
    net = Network(bus_id=[0, 1], u_rated=[10e3, 10e3], bus_type=[SLACK, PQ], u_pu=[1.0, 1.0],
                  p_w=[0.0, p], q_var=[0.0, q], br_id=[2], br_from=[0], br_to=[1], br_status=[1],
                  br_r=[r], br_x=[x], br_g=[0.0], br_b=[b])
    return net


def test_ybus_closed_form(oracle):
    net = two_bus()
    rp, ci, val = oracle.build_ybus(net)
    y = 1.0 / (0.3 + 0.4j)
    expect = np.array([y + 1j * 1e-4, -y, -y, y + 1j * 1e-4])
    assert np.array_equal(rp, [0, 2, 4]) and np.array_equal(ci, [0, 1, 0, 1])
    assert np.allclose(val, expect, rtol=1e-15)


def test_mismatch_closed_forms(oracle):
    """Empty Y or zero voltage: F = -S_spec in watts
    (tests/unit/test_newton_raphson_functions.cpp:266-287, 378-395)."""
    net = two_bus(p=-2.5e5, q=-1e5)
    y = oracle.build_ybus(net)
    f = oracle.mismatches(net, np.zeros(2), y)
    assert np.allclose(f, [2.5e5, 1e5])
    f0 = oracle.mismatches(net, np.ones(2), (np.zeros(3, np.int32), np.zeros(0, np.int32), np.zeros(0)))
    assert np.allclose(f0, [2.5e5, 1e5])


def _jacobian_and_fd(oracle, net):
    """The oracle's first-iteration Jacobian (flat start, dense) and central
    differences of its mismatch in the reference variables (angles of non-slack
    buses, then |V| of PQ buses), with the tolerance the two must meet."""
    y = oracle.build_ybus(net)
    ypu = (y[0], y[1], y[2] * (10e3 ** 2 / 100e6))
    r = oracle.solve(net, max_it=1, j_call=0)
    jrp, jci, jv = r["J"]
    n_var = len(jrp) - 1
    J = np.zeros((n_var, n_var))
    for i in range(n_var):
        J[i, jci[jrp[i]:jrp[i + 1]]] = jv[jrp[i]:jrp[i + 1]]
    v0 = np.where(net.bus_type == PQ, 0.98, net.u_pu).astype(complex)
    ang = [i for i in range(net.n_bus) if net.bus_type[i] != SLACK]
    mag = [i for i in range(net.n_bus) if net.bus_type[i] == PQ]

    def F(x):
        a = np.angle(v0).copy()
        m = np.abs(v0).copy()
        a[ang] += x[:len(ang)]
        m[mag] += x[len(ang):]
        return oracle.mismatches(net, m * np.exp(1j * a), ypu) / 100e6

    h = 1e-6
    Jfd = np.zeros_like(J)
    for k in range(n_var):
        e = np.zeros(n_var)
        e[k] = h
        Jfd[:, k] = (F(e) - F(-e)) / (2 * h)
    return J, Jfd, 1e-6 * np.abs(J).max()


def test_jacobian_equals_finite_difference(oracle):
    """The analytic Jacobian against central differences of the mismatch."""
    J, Jfd, atol = _jacobian_and_fd(oracle, gridgen.small_mesh(9, seed=4))
    assert np.allclose(J, Jfd, atol=atol)


@pytest.mark.parametrize("entry", ["q_angle_diag", "p_mag_offdiag", "dropped_term"])
def test_planted_error_is_caught(oracle, entry):
    """Plausible slips planted in the oracle's Jacobian (a flipped sign on the
    Q-angle diagonal, a flipped off-diagonal dP/d|V|, a dropped diagonal term
    -B|V|^2 of dP/dtheta) must fail the same finite-difference check that the
    correct Jacobian passes."""
    net = gridgen.small_mesh(9, seed=4)
    J, Jfd, atol = _jacobian_and_fd(oracle, net)
    assert np.allclose(J, Jfd, atol=atol)
    bad = J.copy()
    n_ang = int((net.bus_type != SLACK).sum())
    if entry == "q_angle_diag":
        bad[1, 0] = -bad[1, 0]            # row 1: Q of the first PQ bus; column 0: its angle
    elif entry == "p_mag_offdiag":
        r, c = next((r, c) for r in range(0, n_ang, 1) for c in range(n_ang, J.shape[1])
                    if J[r, c] != 0.0 and c - n_ang != r)
        bad[r, c] = -bad[r, c]
    else:
        y = oracle.build_ybus(net)
        b00 = (y[2] * (10e3 ** 2 / 100e6))[y[0][1]:y[0][2]][list(y[1][y[0][1]:y[0][2]]).index(1)].imag
        bad[0, 0] += b00 * 0.98 ** 2     # J[0,0] = -Q - B V^2 with the B V^2 term dropped
    assert not np.allclose(bad, Jfd, atol=atol)
