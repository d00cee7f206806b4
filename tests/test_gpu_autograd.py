"""``juice.log_likelihood`` (the adjoint behind ``torch.autograd.Function``) under losses whose
gradient is negative or of mixed sign, against the FP64 C restatement's parameter flows for the
same signed root weights (which ``tests/test_oracle_cpu.py`` pins to central differences).

Tolerance: theta * gradient within 1e-4 (FP32) / 2e-2 (BF16) of the largest oracle value.
"""
import numpy as np
import pytest
import torch

import parity_utils as U
import pyjuice_amd as juice
from oracle.pc_cpu import CompiledCPU
from pyjuice_amd.model import TensorCircuit
from pyjuice_amd.model.autograd import _sum_param_to_flow

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-4, "bf16": 2e-2}


@pytest.fixture(scope="module")
def device():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def _circuit(name):
    if name == "hclt_cat_bs64":
        ns, _, cards = U.hclt_categorical(L=128, block_size=64, K=16)
    elif name == "hclt_bern16":
        ns, _, cards = U.hclt_bernoulli()
    elif name == "pd_small":
        ns, cards = U.pd_small()
    else:
        ns, cards = U.hmm_small()
    return ns, cards


def _weights(kind, B):
    if kind == "nll_mean":      # (-ll).mean(): every root weight is -1 / B
        return np.full([B], -1.0 / B)
    g = np.random.default_rng(12)
    w = g.normal(size=B)
    w[::7] = 0.0                 # some samples without weight
    return w


@pytest.mark.parametrize("name", ["hclt_cat_bs64", "hclt_bern16", "pd_small", "hmm_small"])
@pytest.mark.parametrize("loss", ["nll_mean", "mixed_sign"])
@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_signed_gradients_match_oracle(name, loss, precision, device):
    U.say(f"[autograd] {name} {loss} {precision}")
    ns, cards = _circuit(name)
    pc = TensorCircuit(ns)
    pc.to(device)
    B = 64
    X = U.random_data(cards, B, seed=21)
    xt = torch.from_numpy(X).to(device)
    w = _weights(loss, B)
    params = pc.params.detach().clone().requires_grad_(True)
    in_params = [l.params.detach().clone().requires_grad_(True) for l in pc.input_layer_group]
    lls = juice.log_likelihood(pc, xt, params=params, input_params=in_params, precision=precision)
    if loss == "nll_mean":
        (-lls[:, 0]).mean().backward()
    else:
        (lls[:, 0] * torch.from_numpy(w).to(device=device, dtype=torch.float32)).sum().backward()

    cpu = CompiledCPU(pc, np.float64)
    cpu.forward(X)
    cpu.backward(root_flows=w[None, :])
    fused = cpu.param_flows.copy()
    for t, k, ch in cpu.fuse:
        fused[t:t + k] = sum(cpu.param_flows[c:c + k] for c in ch)
    theta = pc.params.detach().cpu().numpy().astype(np.float64)
    idx = _sum_param_to_flow(pc).cpu().numpy()
    ok = idx >= 0
    g = params.grad.detach().cpu().numpy().astype(np.float64)
    assert np.isfinite(g).all()
    got, ref = g[ok] * theta[ok], fused[idx[ok]]
    assert U.rel_err(got, ref) <= TOL[precision], U.rel_err(got, ref)
    if loss == "nll_mean":
        assert (got <= 1e-12 * np.abs(ref).max()).all()       # every flow is non-positive

    for layer, p, d in zip(pc.input_layer_group, in_params, cpu.inputs):
        pf = d["pflows"].copy()
        for t, k, ch in d["tied"]:
            pf[t:t + k] = sum(d["pflows"][c:c + k] for c in ch)
        gi = p.grad.detach().cpu().numpy().astype(np.float64)
        th = layer.params.detach().cpu().numpy().astype(np.float64)
        assert np.isfinite(gi).all()
        src = layer.source_nids.cpu().numpy()
        s_pids = layer.s_pids.cpu().numpy()[src]
        s_pfids = layer.s_pfids.cpu().numpy()[src]
        if d["kind"] == 0:
            K = layer._ncats.cpu().numpy()[src]
            pi = np.concatenate([a + np.arange(k) for a, k in zip(s_pids, K)])
            fi = np.concatenate([a + np.arange(k) for a, k in zip(s_pfids, K)])
            assert U.rel_err(gi[pi] * th[pi], pf[fi]) <= TOL[precision]
        else:
            p1 = th[s_pids]
            ref_g = pf[s_pfids] / p1 - pf[s_pfids + 1] / (1.0 - p1)
            assert U.rel_err(gi[s_pids], ref_g) <= TOL[precision]
