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


def test_adjoint_10k_x256(device, oracle):
    """The adjoint workload at full size: the 10k-bus meshed grid (1,500
    transformers, 300 shunts, 15% PV) x 256 load scenarios (+-30%) in one plan,
    L = total active loss + sum over PQ buses of (|V| - 1)^2, FP64 adjoint of all
    256 scenarios in one call.  Checked against
      * the sparse FP64 reference gradient (tests/adjoint_ref.py) on scenarios 0
        and 137: every load p, q and 400 random branches r, x, to 1e-6 of the
        largest gradient of each kind;
      * central differences of full re-solves on scenario 137, eighth order at
        h = 3e-2 relative, on the largest gradients of each kind (where a
        difference quotient resolves 1e-6) and on random parameters; and the
        second-order stencil at h = 1e-6 on the largest branch gradients.  Each
        parameter is judged against 1e-6 |g| plus three times its noise bound
        (the loss of a solve carries round-off of ~5e-14, measured on copies
        whose loads differ in the 13th digit, tests/fd_check.py); the resolvable
        ones must be within 1e-6."""
    import sys
    import time
    from adjoint_ref import reference_gradient
    from fd_check import GRAD_KEY, assess, fd_gradients, noise_floor

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    base = 100e6
    net = gridgen.large_mesh(10_000, seed=1)
    scen = gridgen.load_scenarios(net, 256, seed=2000)
    nb, nr = net.n_bus, net.n_branch
    t0 = time.time()
    plan = capi.Plan(scen)
    log(f"plan of 256 x 10k-bus scenarios: {time.time() - t0:.1f} s")
    t0 = time.time()
    r = plan.solve(1e-12, 60, base)
    assert r["converged"].all()
    out = plan.adjoint(1.0, 1.0)
    log(f"solve + adjoint of 256 scenarios: {time.time() - t0:.2f} s, iterations <= {int(r['iterations'].max())}")
    rng = np.random.default_rng(11)
    brs = rng.choice(nr, 400, replace=False)
    for s in (0, 137):
        t0 = time.time()
        gp, gq, gr, gx, L = reference_gradient(oracle, scen[s], r["v"][s], base, 1.0, 1.0, branches=brs)
        sl, sr = slice(s * nb, (s + 1) * nb), slice(s * nr, (s + 1) * nr)
        errs = (rel_err(out["grad_p"][sl], gp), rel_err(out["grad_q"][sl], gq),
                rel_err(out["grad_r"][sr][brs], gr[brs]), rel_err(out["grad_x"][sr][brs], gx[brs]))
        log(f"scenario {s}: vs sparse reference rel err p q r x {errs} ({time.time() - t0:.1f} s)")
        assert abs(out["loss"][s] - L) <= 1e-11 * abs(L)
        assert max(errs) < 1e-6
    s = 137
    g = {k: out[k][s * nb:(s + 1) * nb] if k in ("grad_p", "grad_q") else out[k][s * nr:(s + 1) * nr]
         for k in ("grad_p", "grad_q", "grad_r", "grad_x")}
    pq_buses = np.nonzero(scen[s].bus_type == PQ)[0]
    top = [(s, "br_r", int(k)) for k in np.argsort(-np.abs(g["grad_r"]))[:3]]
    top += [(s, "br_x", int(k)) for k in np.argsort(-np.abs(g["grad_x"]))[:3]]
    params = top + [(s, "p_w", int(i)) for i in pq_buses[np.argsort(-np.abs(g["grad_p"][pq_buses]))[:2]]]
    params += [(s, "q_var", int(i)) for i in pq_buses[np.argsort(-np.abs(g["grad_q"][pq_buses]))[:1]]]
    params += [(s, arr, int(i)) for arr in ("p_w", "q_var") for i in rng.choice(pq_buses, 2, replace=False)]
    params += [(s, arr, int(k)) for arr in ("br_r", "br_x") for k in rng.choice(nr, 2, replace=False)]
    sigma = noise_floor(oracle, scen[s], base, seed=s)
    for order, h_rel, plist in ((8, 3e-2, params), (2, 1e-6, top)):
        t0 = time.time()
        g_fd, hs = fd_gradients(oracle, scen, plist, order, h_rel, base)
        g_adj = np.array([g[GRAD_KEY[a]][k] for _, a, k in plist])
        rep = assess(g_adj, g_fd, hs, sigma, order)
        log(f"FD order {order}, h = {h_rel:g} rel, {len(plist)} parameters ({time.time() - t0:.1f} s): noise floor "
            f"of L {sigma:.2e}, {({k: v for k, v in rep.items() if k not in ('ok', 'resolvable_mask')})}")
        assert rep["ok"].all()
        if order == 8:
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


def test_torch_autograd_interleaved_forwards(device):
    """Two forwards on one plan with different loads, backward of the first
    after the second: each output gets the gradients of its own solve (the
    wrapper re-solves the first inputs), equal to a fresh plan's."""
    import torch
    from gap_amd.autograd import PowerFlow, SolveOptions
    base = 100e6
    net = gridgen.small_mesh(150, seed=4, n_extra=20)
    plan = capi.Plan([net])
    opts = SolveOptions(1e-11, 60, base)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda", requires_grad=True)
    ins1 = [t(net.p_w), t(net.q_var), t(net.br_r), t(net.br_x)]
    ins2 = [t(net.p_w * 1.25), t(net.q_var * 0.8), t(net.br_r), t(net.br_x * 1.1)]
    vm1, va1 = PowerFlow.apply(*ins1, plan, opts)
    vm2, va2 = PowerFlow.apply(*ins2, plan, opts)
    assert not torch.equal(vm1, vm2)
    g = torch.linspace(-1.0, 1.0, net.n_bus, dtype=torch.float64, device="cuda")
    (vm1 * g).sum().backward()
    fresh = capi.Plan([net])
    fresh.solve(1e-11, 60, base)
    ref = fresh.adjoint(0.0, 0.0, g.cpu().numpy(), None)
    assert np.allclose(ins1[0].grad.cpu().numpy(), ref["grad_p"], rtol=1e-10, atol=0)
    assert np.allclose(ins1[3].grad.cpu().numpy(), ref["grad_x"], rtol=1e-10, atol=0)
    (vm2 * g).sum().backward()
    n2 = net.copy()
    n2.p_w, n2.q_var, n2.br_x = net.p_w * 1.25, net.q_var * 0.8, net.br_x * 1.1
    fresh2 = capi.Plan([n2])
    fresh2.solve(1e-11, 60, base)
    ref2 = fresh2.adjoint(0.0, 0.0, g.cpu().numpy(), None)
    assert np.allclose(ins2[0].grad.cpu().numpy(), ref2["grad_p"], rtol=1e-10, atol=0)


def test_adjoint_and_flows_refuse_stale_state(device):
    """The adjoint needs a converged solve of the current inputs, the branch
    flows a solve of the current branch data."""
    net = gridgen.small_mesh(60, seed=8)
    plan = capi.Plan([net])
    plan.solve(1e-10, 60, 100e6)
    plan.adjoint(1.0, 1.0)
    plan.branch_flows()
    plan.set_injections(0, net.p_w * 1.1, net.q_var)
    with pytest.raises(capi.GapHipError) as e:
        plan.adjoint(1.0, 1.0)
    assert e.value.code == capi.E_STATE
    plan.branch_flows()  # V and Y unchanged: still the last solution's flows
    plan.set_branch_rx(0, net.br_r * 1.1, net.br_x)
    with pytest.raises(capi.GapHipError) as e:
        plan.branch_flows()
    assert e.value.code == capi.E_STATE
    plan.solve(1e-10, 60, 100e6)
    plan.adjoint(1.0, 1.0)
    plan.branch_flows()
    plan.solve(1e-10, 2, 100e6)  # stops at the iteration cap
    with pytest.raises(capi.GapHipError) as e:
        plan.adjoint(1.0, 1.0)
    assert e.value.code == capi.E_STATE
