"""Log-likelihood of a compiled circuit as a differentiable function of its parameters.

``log_likelihood(pc, x)`` returns ``lls [B, num_roots]`` from the HIP forward; its backward is
the hand-written adjoint: the flow backward (``TensorCircuit.backward``) with the incoming
gradient w as root flows.  Its parameter flows are ``pf = theta * d(sum_b w_b LL_b) / d theta``:
every parameter occurrence (each tied copy counted separately) enters the circuit's value
linearly, so theta * d(value)/d(theta) of one occurrence is the flow through that edge, and the
flows of tied copies add up to the total derivative.  So

* sum and Categorical parameters: ``dL/dtheta = pf / theta`` (tied copies summed first);
* Bernoulli ``p``: ``dL/dp = f(x=1) / p - f(x=0) / (1 - p)``.

The gradient is taken with respect to the parameters as stored (no renormalisation), the
quantity the EM update and the reference's ``param_flows`` are built from
(``pyjuice/model/tensorcircuit.py:24-37, 344-538``).  The backward overwrites the circuit's
parameter-flow accumulators and needs the node marginals of this forward (run it before the
next forward).  A parameter equal to 0 gets gradient 0.  The flow kernels take logarithms of
flows, so a gradient with entries of both signs runs the flow backward twice (positive and
negative parts) and subtracts.
"""
from __future__ import annotations

from typing import Optional, Sequence

import torch

from pyjuice_amd import _native as N

__all__ = ["log_likelihood"]


def _sum_param_to_flow(pc) -> torch.Tensor:
    """Flow slot of every flat sum parameter (-1 for the dummy prefix), cached on the circuit."""
    m = getattr(pc, "_param_to_flow", None)
    if m is not None and m.device == pc.device:
        return m
    idx = torch.full([pc.num_sum_params], -1, dtype=torch.int64)
    for ns in pc.root_ns:
        if ns.is_sum() and not ns.is_tied() and ns.provided("_param_range"):
            a, b = ns._param_range
            fa, fb = ns._param_flow_range
            assert b - a == fb - fa
            idx[a:b] = torch.arange(fa, fb, dtype=torch.int64)
    pc._param_to_flow = idx.to(pc.device)
    return pc._param_to_flow


def _fused(pf: torch.Tensor, groups) -> torch.Tensor:
    """Copy of ``pf`` with the flows of tied copies added into their source's slots."""
    out = pf.clone()
    pf64 = 1 if out.dtype == torch.float64 else 0
    for t, n, ch in groups:
        N.call("pjamd_pflow_fuse", N.ptr(out), pf64, int(t), int(n), N.ptr(ch), ch.numel(), N.stream())
    return out


def _sum_grad(pc, params: torch.Tensor) -> torch.Tensor:
    pf = _fused(pc.param_flows, pc._em_dev["fuse"])
    idx = _sum_param_to_flow(pc)
    g = torch.zeros_like(params, dtype=torch.float64)
    ok = idx >= 0
    th = params.detach().to(torch.float64)[ok]
    f = pf[idx[ok]].to(torch.float64)
    g[ok] = torch.where(th > 0, f / th.clamp_min(1e-300), torch.zeros_like(th))
    return g.to(params.dtype)


def _input_grad(layer, params: torch.Tensor) -> torch.Tensor:
    t = layer._dev
    pf = _fused(layer.param_flows, [(a, n, c) for a, n, c in t["tied"]]).to(torch.float64)
    src = t["src"].long()
    s_pids, s_pfids, ncats = t["s_pids"][src], t["s_pfids"][src], t["ncats"][src].long()
    th = params.detach().to(torch.float64)
    g = torch.zeros_like(th)
    if layer._kind() == 0:
        K = int(ncats.max().item()) if ncats.numel() > 0 else 0
        k = torch.arange(K, device=th.device)
        valid = k[None, :] < ncats[:, None]
        pi = (s_pids[:, None] + k[None, :])[valid]
        fi = (s_pfids[:, None] + k[None, :])[valid]
        x = th[pi]
        g[pi] = torch.where(x > 0, pf[fi] / x.clamp_min(1e-300), torch.zeros_like(x))
    else:
        p = th[s_pids]
        g[s_pids] = pf[s_pfids] / p.clamp_min(1e-300) - pf[s_pfids + 1] / (1.0 - p).clamp_min(1e-300)
    return g.to(params.dtype)


class _CircuitLL(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pc, inputs, precision, params, *input_params):
        with torch.no_grad():
            if params.data_ptr() != pc.params.data_ptr():
                pc.params.data.copy_(params)
            for layer, p in zip(pc.input_layer_group, input_params):
                if p.data_ptr() != layer.params.data_ptr():
                    layer.params.data.copy_(p)
            lls = pc.forward(inputs, precision=precision)
        ctx.pc, ctx.inputs, ctx.precision = pc, inputs, precision
        ctx.save_for_backward(params, *input_params)
        return lls

    @staticmethod
    def backward(ctx, grad):
        pc = ctx.pc
        params, *input_params = ctx.saved_tensors
        want_sum = ctx.needs_input_grad[3]
        want_in = [ctx.needs_input_grad[4 + i] for i in range(len(input_params))]
        g, g_in = None, [None] * len(input_params)
        with torch.no_grad():
            # Flows are linear in the root weights but the flow kernels work on log(flow), so a
            # signed gradient runs the adjoint once per sign: weights max(g, 0) and max(-g, 0),
            # and the second result is subtracted.
            w = grad.t().contiguous()
            parts = [(1.0, w.clamp_min(0.0)), (-1.0, (-w).clamp_min(0.0))]
            for sign, wp in parts:
                if not bool((wp > 0).any()):
                    continue
                pc.backward(ctx.inputs, ll_weights=wp, flows_memory=0.0, precision=ctx.precision)
                if want_sum:
                    gs = _sum_grad(pc, params)
                    g = gs * sign if g is None else g + sign * gs
                for i, (layer, p) in enumerate(zip(pc.input_layer_group, input_params)):
                    if want_in[i]:
                        gi = _input_grad(layer, p)
                        g_in[i] = gi * sign if g_in[i] is None else g_in[i] + sign * gi
            if want_sum and g is None:
                g = torch.zeros_like(params)
            for i, p in enumerate(input_params):
                if want_in[i] and g_in[i] is None:
                    g_in[i] = torch.zeros_like(p)
        return (None, None, None, g, *g_in)


def log_likelihood(pc, inputs: torch.Tensor, params: Optional[torch.Tensor] = None,
                   input_params: Optional[Sequence[torch.Tensor]] = None,
                   precision: Optional[str] = None) -> torch.Tensor:
    """``lls [B, num_roots]`` of ``inputs [B, num_vars]``, differentiable with respect to
    ``params`` (default ``pc.params``) and ``input_params`` (default each input layer's
    ``params``); pass tensors with ``requires_grad=True`` to get their gradients."""
    params = pc.params if params is None else params
    if input_params is None:
        input_params = [layer.params for layer in pc.input_layer_group]
    return _CircuitLL.apply(pc, inputs, precision, params, *input_params)
