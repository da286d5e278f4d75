"""Adjoint sensitivities on the GPU against (a) a dense FP64 reference
gradient (numpy solve of J^T lambda = dL/dx with the oracle's Jacobian, the
partial derivatives taken by central differences of explicit functions of V
and of the parameters) and (b) central finite differences of full re-solves
(h = 1e-6 relative, every perturbed network batched into one plan, solved to
the round-off floor).

L = w_loss * sum_i Re(S_i) [pu] + w_dev * sum_PQ (|V_i| - 1)^2
    + sum g_vm |V_i| + sum g_va arg V_i."""
import numpy as np
import pytest

from conftest import load_golden
from gap_amd import capi, gridgen
from gap_amd.network import PQ, SLACK

pytestmark = pytest.mark.gpu


def ypu_dense(oracle, net, base):
    rp, ci, val = oracle.build_ybus(net)
    zb = net.u_rated[0] ** 2 / base
    n = net.n_bus
    Y = np.zeros((n, n), complex)
    for i in range(n):
        for q in range(rp[i], rp[i + 1]):
            Y[i, ci[q]] += val[q] * zb
    return Y, (rp, ci, val * zb)


def loss(Y, v, net, w_loss, w_dev, gvm=None, gva=None):
    s = v * np.conj(Y @ v)
    pq = net.bus_type == PQ
    out = w_loss * s.real.sum() + w_dev * ((np.abs(v[pq]) - 1) ** 2).sum()
    if gvm is not None:
        out += (gvm * np.abs(v)).sum()
    if gva is not None:
        out += (gva * np.angle(v)).sum()
    return out


def var_maps(net):
    slack = int(np.nonzero(net.bus_type == SLACK)[0][0])
    ang = [i for i in range(net.n_bus) if i != slack]
    mag = [i for i in range(net.n_bus) if net.bus_type[i] == PQ]
    eq = []
    for i in range(net.n_bus):
        if net.bus_type[i] == SLACK:
            continue
        eq.append(("P", i))
        if net.bus_type[i] == PQ:
            eq.append(("Q", i))
    return ang, mag, eq


def mismatch(Y, v, net, base):
    s = v * np.conj(Y @ v)
    sp = (net.p_w + 1j * net.q_var) / base
    out = []
    for i in range(net.n_bus):
        if net.bus_type[i] == SLACK:
            continue
        out.append(s[i].real - sp[i].real)
        if net.bus_type[i] == PQ:
            out.append(s[i].imag - sp[i].imag)
    return np.array(out)


def d4(f, h):
    """Fourth-order central difference of a scalar or vector function at 0."""
    return (-f(2 * h) + 8 * f(h) - 8 * f(-h) + f(-2 * h)) / (12 * h)


def dense_reference(oracle, net, v, base, w_loss, w_dev, gvm, gva):
    """dL/d(p, q, r, x) with numpy: lambda = J^{-T} dL/dx (dense solve), the
    partials of the explicit functions L(V, r, x) and F(V, r, x) by fourth-order
    central differences (truncation ~1e-12, rounding ~1e-13 relative)."""
    Y, ysp = ypu_dense(oracle, net, base)
    ang, mag, eq = var_maps(net)
    jrp, jci, jv = oracle.jacobian(net, v, ysp)
    nv = len(jrp) - 1
    J = np.zeros((nv, nv))
    for r in range(nv):
        J[r, jci[jrp[r]:jrp[r + 1]]] = jv[jrp[r]:jrp[r + 1]]
    vm, va = np.abs(v), np.angle(v)

    def L_of_x(x):
        a, m = va.copy(), vm.copy()
        a[ang] += x[:len(ang)]
        m[mag] += x[len(ang):]
        return loss(Y, m * np.exp(1j * a), net, w_loss, w_dev, gvm, gva)

    dLdx = np.zeros(nv)
    for k in range(nv):
        e = np.zeros(nv)
        e[k] = 1.0
        dLdx[k] = d4(lambda t: L_of_x(t * e), 1e-4)
    lam = np.linalg.solve(J.T, dLdx)
    gp = np.zeros(net.n_bus)
    gq = np.zeros(net.n_bus)
    for k, (kind, i) in enumerate(eq):
        (gp if kind == "P" else gq)[i] = lam[k] / base
    gr = np.zeros(net.n_branch)
    gx = np.zeros(net.n_branch)
    for b in range(net.n_branch):
        for arr, out in (("br_r", gr), ("br_x", gx)):
            def at(t):
                n2 = net.copy()
                getattr(n2, arr)[b] += t
                Y2, _ = ypu_dense(oracle, n2, base)
                return np.concatenate([[loss(Y2, v, n2, w_loss, w_dev, gvm, gva)], mismatch(Y2, v, n2, base)])
            d = d4(at, 1e-3 * abs(getattr(net, arr)[b]))
            out[b] = d[0] - lam @ d[1:]
    return gp, gq, gr, gx, loss(Y, v, net, w_loss, w_dev, gvm, gva)


def rel_err(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


@pytest.mark.parametrize("name", ["mesh200_seed0", "special_pv_shunt_parallel"])
def test_adjoint_vs_dense_reference(device, oracle, name):
    net, g = load_golden(name)
    base = 100e6
    plan = capi.Plan([net])
    r = plan.solve(1e-12, 60, base)
    v = r["v"][0]
    rng = np.random.default_rng(1)
    gvm = rng.normal(size=net.n_bus) * 0.1
    gva = rng.normal(size=net.n_bus) * 0.1
    out = plan.adjoint(1.0, 2.0, gvm, gva)
    gp, gq, gr, gx, L = dense_reference(oracle, net, v, base, 1.0, 2.0, gvm, gva)
    print(name, "rel err p q r x:", rel_err(out["grad_p"], gp), rel_err(out["grad_q"], gq),
          rel_err(out["grad_r"], gr), rel_err(out["grad_x"], gx))
    assert abs(out["loss"][0] - L) <= 1e-12 * max(1.0, abs(L))
    assert rel_err(out["grad_p"], gp) < 1e-6
    assert rel_err(out["grad_q"], gq) < 1e-6
    assert rel_err(out["grad_r"], gr) < 1e-6
    assert rel_err(out["grad_x"], gx) < 1e-6


@pytest.mark.parametrize("order,h_rel,bar", [(2, 1e-6, 1e-3), (4, 1e-3, 1e-6)])
def test_adjoint_vs_finite_differences(device, oracle, order, h_rel, bar):
    """Finite differences of full re-solves, every perturbed network of every
    parameter batched into one plan and solved to the round-off floor.

    order 2, h = 1e-6 (relative): the specified central difference.  Its own
    FP64 noise reaches ~3e-4 for the parameters with the smallest gradients
    (dL ~ 1e-12 against L ~ 1e-2 known to ~1e-16), so it is held to 1e-3.
    order 4, h = 1e-3: truncation ~h^4 ~ 1e-12 and noise ~1e-10, so the 1e-6
    relative bar applies to every parameter."""
    base = 100e6
    net = gridgen.small_mesh(200, seed=0)
    rng = np.random.default_rng(5)
    pq_buses = np.nonzero(net.bus_type == PQ)[0]
    params = [("p_w", int(i)) for i in rng.choice(pq_buses, 5, replace=False)]
    params += [("q_var", int(i)) for i in rng.choice(pq_buses, 5, replace=False)]
    params += [("br_r", int(k)) for k in rng.choice(net.n_branch, 5, replace=False)]
    params += [("br_x", int(k)) for k in rng.choice(net.n_branch, 5, replace=False)]
    steps = [1, -1] if order == 2 else [2, 1, -1, -2]
    weights = [1, -1] if order == 2 else [-1, 8, -8, 1]
    denom = 2 if order == 2 else 12
    nets, hs = [], []
    for arr, k in params:
        h = h_rel * abs(getattr(net, arr)[k])
        hs.append(h)
        for st in steps:
            n2 = net.copy()
            getattr(n2, arr)[k] += st * h
            nets.append(n2)
    res = capi.Plan(nets).solve(1e-14, 60, base)
    assert (res["final_mismatch"] < 1e-12).all()
    fd = []
    ns = len(steps)
    for j in range(len(params)):
        Ls = [loss(ypu_dense(oracle, nets[ns * j + u], base)[0], res["v"][ns * j + u], nets[ns * j + u], 1.0, 1.0)
              for u in range(ns)]
        fd.append(sum(w * l for w, l in zip(weights, Ls)) / (denom * hs[j]))
    plan = capi.Plan([net])
    plan.solve(1e-12, 60, base)
    out = plan.adjoint(1.0, 1.0)
    key = {"p_w": "grad_p", "q_var": "grad_q", "br_r": "grad_r", "br_x": "grad_x"}
    errs = np.array([abs(out[key[arr]][k] - fd[j]) / max(abs(fd[j]), 1e-300) for j, (arr, k) in enumerate(params)])
    print(f"order {order} h={h_rel:g} relative errors:", errs)
    assert errs.max() < bar


def test_adjoint_10k_grid_scenarios(device, oracle):
    """10k-bus meshed grid (1,500 transformers, 300 shunts, 15% PV) x 4 load
    scenarios (+-30%) in one plan: the FP64 adjoint against the sparse FP64
    reference gradient (scenario 0: every load, 300 random branches) and
    against eighth-order central differences of full re-solves on 16
    parameters of scenario 1."""
    from adjoint_ref import reference_gradient
    base = 100e6
    net = gridgen.large_mesh(10_000, seed=1)
    scen = gridgen.load_scenarios(net, 4, seed=2000)
    plan = capi.Plan(scen)
    r = plan.solve(1e-11, 60, base)
    assert r["converged"].all()
    out = plan.adjoint(1.0, 1.0)
    nb, nr = net.n_bus, net.n_branch
    rng = np.random.default_rng(11)
    brs = rng.choice(nr, 300, replace=False)
    gp, gq, gr, gx, L = reference_gradient(oracle, scen[0], r["v"][0], base, 1.0, 1.0, branches=brs)
    errs = (rel_err(out["grad_p"][:nb], gp), rel_err(out["grad_q"][:nb], gq),
            rel_err(out["grad_r"][:nr][brs], gr[brs]), rel_err(out["grad_x"][:nr][brs], gx[brs]))
    print("10k adjoint vs sparse reference, rel err p q r x:", errs)
    assert abs(out["loss"][0] - L) <= 1e-11 * abs(L)
    assert max(errs) < 1e-6
    # eighth-order central differences of re-solves on scenario 1, at the
    # parameters with the largest gradients (where a difference quotient can
    # resolve 1e-6) and at random ones (held to the measured noise floor)
    from fd_check import GRAD_KEY, assess, fd_gradients, noise_floor
    g1 = {k: out[k][nb:2 * nb] if k in ("grad_p", "grad_q") else out[k][nr:2 * nr] for k in
          ("grad_p", "grad_q", "grad_r", "grad_x")}
    pq_buses = np.nonzero(scen[1].bus_type == PQ)[0]
    params = [(1, "br_r", int(k)) for k in np.argsort(-np.abs(g1["grad_r"]))[:3]]
    params += [(1, "br_x", int(k)) for k in np.argsort(-np.abs(g1["grad_x"]))[:3]]
    params += [(1, "p_w", int(i)) for i in pq_buses[np.argsort(-np.abs(g1["grad_p"][pq_buses]))[:2]]]
    params += [(1, "p_w", int(i)) for i in rng.choice(pq_buses, 2, replace=False)]
    params += [(1, "q_var", int(i)) for i in rng.choice(pq_buses, 2, replace=False)]
    params += [(1, arr, int(k)) for arr in ("br_r", "br_x") for k in rng.choice(nr, 2, replace=False)]
    g_fd, hs = fd_gradients(oracle, scen, params, 8, 3e-2, base)
    sigma = noise_floor(oracle, scen[1], base)
    g_adj = np.array([g1[GRAD_KEY[a]][k] for _, a, k in params])
    rep = assess(g_adj, g_fd, hs, sigma, 8)
    print("10k adjoint vs 8th-order FD (h = 3e-2 rel): noise floor of L", sigma,
          {k: v for k, v in rep.items() if k not in ("ok", "resolvable_mask")})
    assert rep["ok"].all()
    assert rep["resolvable_mask"][:6].sum() >= 2
    assert rep["resolvable_max_rel"] < 1e-6


def test_adjoint_batch_equals_single(device):
    base = 100e6
    net = gridgen.small_mesh(500, seed=12, n_extra=40)
    scen = gridgen.load_scenarios(net, 6, seed=3)
    plan = capi.Plan(scen)
    plan.solve(1e-10, 60, base)
    b = plan.adjoint(1.0, 0.5)
    nb, nr = net.n_bus, net.n_branch
    for k, n in enumerate(scen):
        p1 = capi.Plan([n])
        p1.solve(1e-10, 60, base)
        s = p1.adjoint(1.0, 0.5)
        assert np.array_equal(b["grad_p"][k * nb:(k + 1) * nb], s["grad_p"])
        assert np.array_equal(b["grad_r"][k * nr:(k + 1) * nr], s["grad_r"])
        assert b["loss"][k] == s["loss"][0]


def test_torch_autograd(device):
    import torch
    from gap_amd.autograd import PowerFlow, PowerFlowLoss, SolveOptions
    base = 100e6
    nets = [gridgen.small_mesh(120, seed=s, n_extra=24) for s in (1, 2)]
    plan = capi.Plan(nets)
    opts = SolveOptions(1e-11, 60, base)
    cat = lambda k: torch.tensor(np.concatenate([getattr(n, k) for n in nets]), dtype=torch.float64,
                                 device="cuda", requires_grad=True)
    p, q, r, x = cat("p_w"), cat("q_var"), cat("br_r"), cat("br_x")
    Lv = PowerFlowLoss.apply(p, q, r, x, plan, opts, 1.0, 1.0)
    (Lv * torch.tensor([1.0, 2.0], device="cuda", dtype=torch.float64)).sum().backward()
    plan.solve(1e-11, 60, base)
    ref = plan.adjoint(1.0, 1.0)
    nb = nets[0].n_bus
    assert torch.allclose(p.grad[:nb].cpu(), torch.tensor(ref["grad_p"][:nb]), rtol=1e-12, atol=0)
    assert torch.allclose(p.grad[nb:].cpu(), torch.tensor(2 * ref["grad_p"][nb:]), rtol=1e-12, atol=0)
    # voltage outputs with an arbitrary upstream gradient equal the g_vm / g_va adjoint
    p2, q2, r2, x2 = cat("p_w"), cat("q_var"), cat("br_r"), cat("br_x")
    vm, va = PowerFlow.apply(p2, q2, r2, x2, plan, opts)
    gvm = torch.randn_like(vm)
    gva = torch.randn_like(va)
    (vm * gvm + va * gva).sum().backward()
    ref2 = plan.adjoint(0.0, 0.0, gvm.cpu().numpy(), gva.cpu().numpy())
    assert np.allclose(q2.grad.cpu().numpy(), ref2["grad_q"], rtol=1e-12, atol=0)
    assert np.allclose(x2.grad.cpu().numpy(), ref2["grad_x"], rtol=1e-12, atol=0)
