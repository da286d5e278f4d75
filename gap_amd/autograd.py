"""Differentiable power flow: ``torch.autograd.Function`` wrappers over the
GPU Newton-Raphson plan and its adjoint.

``PowerFlow.apply(p_w, q_var, r, x, plan, opts)`` solves every network of the
plan with the given bus injections (W / VAr, generator sign, concatenated
over the plan's networks) and branch r1 / x1 (ohm, concatenated) and returns
``(vm, va)``: bus voltage magnitude (pu) and angle (rad).  Backward solves
J^T lambda = dL/dx with the FP64 factors of the Jacobian at the solution
(one transposed solve for the whole batch) and returns dL/dp, dL/dq, dL/dr,
dL/dx.

``PowerFlowLoss.apply(p_w, q_var, r, x, plan, opts, w_loss, w_dev)`` returns
per network L = w_loss * total active loss (pu) + w_dev * sum_PQ (|V| - 1)^2,
with the same backward (the loss also depends on r, x explicitly).

Tensors may live on any device; the plan's device buffers are the compute
path, tensors only carry values in and gradients out.

The adjoint belongs to the solve that produced the outputs.  ``PowerFlow``
keeps its inputs and the plan's epoch (solves and input changes); if the plan was solved again or
its inputs changed before ``backward`` runs, backward re-loads those inputs and
solves again before taking the adjoint, so interleaved forwards on one plan
give each output its own gradients.  ``PowerFlowLoss`` takes its gradients in
forward.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .capi import Plan


@dataclass
class SolveOptions:
    tolerance: float = 1e-8
    max_iterations: int = 100
    base_power: float = 100e6


def _np(t: torch.Tensor) -> np.ndarray:
    return t.detach().to("cpu", torch.float64).numpy()


def _load(plan: Plan, p_w, q_var, r, x):
    p, q, rr, xx = (a if isinstance(a, np.ndarray) else _np(a) for a in (p_w, q_var, r, x))
    bo = 0
    ro = 0
    for k, net in enumerate(plan.nets):
        nb, nr = net.n_bus, net.n_branch
        plan.set_injections(k, p[bo:bo + nb], q[bo:bo + nb])
        if nr:
            plan.set_branch_rx(k, rr[ro:ro + nr], xx[ro:ro + nr])
        bo += nb
        ro += nr
    if bo != len(p) or ro != len(rr):
        raise ValueError("parameter vectors do not match the plan's networks")


def _solve(plan: Plan, opts: SolveOptions):
    res = plan.solve(opts.tolerance, opts.max_iterations, opts.base_power)
    if not res["converged"].all():
        bad = np.nonzero(~res["converged"])[0]
        raise RuntimeError(f"power flow did not converge for networks {bad.tolist()}")
    v = np.concatenate(res["v"])
    return v


class PowerFlow(torch.autograd.Function):
    @staticmethod
    def forward(ctx, p_w, q_var, r, x, plan: Plan, opts: SolveOptions = SolveOptions()):
        _load(plan, p_w, q_var, r, x)
        v = _solve(plan, opts)
        ctx.plan = plan
        ctx.opts = opts
        ctx.epoch = plan.epoch
        ctx.inputs = (_np(p_w).copy(), _np(q_var).copy(), _np(r).copy(), _np(x).copy())
        ctx.dev = p_w.device
        vm = torch.as_tensor(np.abs(v), device=p_w.device)
        va = torch.as_tensor(np.angle(v), device=p_w.device)
        return vm, va

    @staticmethod
    def backward(ctx, g_vm, g_va):
        gvm = None if g_vm is None else _np(g_vm)
        gva = None if g_va is None else _np(g_va)
        if ctx.plan.epoch != ctx.epoch:
            # the plan solved other inputs since this forward: restore ours
            _load(ctx.plan, *ctx.inputs)
            _solve(ctx.plan, ctx.opts)
        out = ctx.plan.adjoint(0.0, 0.0, gvm, gva)
        d = ctx.dev
        return (torch.as_tensor(out["grad_p"], device=d), torch.as_tensor(out["grad_q"], device=d),
                torch.as_tensor(out["grad_r"], device=d), torch.as_tensor(out["grad_x"], device=d), None, None)


class PowerFlowLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, p_w, q_var, r, x, plan: Plan, opts: SolveOptions = SolveOptions(), w_loss: float = 1.0,
                w_dev: float = 1.0):
        _load(plan, p_w, q_var, r, x)
        _solve(plan, opts)
        out = plan.adjoint(w_loss, w_dev)
        ctx.grads = out
        ctx.bus_n = [n.n_bus for n in plan.nets]
        ctx.br_n = [n.n_branch for n in plan.nets]
        ctx.dev = p_w.device
        return torch.as_tensor(out["loss"], device=p_w.device)

    @staticmethod
    def backward(ctx, g):
        # each network's parameters only reach its own loss: scale its slice
        w = _np(g).reshape(-1)
        wb = np.repeat(w, ctx.bus_n)
        wr = np.repeat(w, ctx.br_n)
        out, d = ctx.grads, ctx.dev
        return (torch.as_tensor(out["grad_p"] * wb, device=d), torch.as_tensor(out["grad_q"] * wb, device=d),
                torch.as_tensor(out["grad_r"] * wr, device=d), torch.as_tensor(out["grad_x"] * wr, device=d),
                None, None, None, None)


__all__ = ["PowerFlow", "PowerFlowLoss", "SolveOptions"]
