"""Pins the C restatement on the compiled index tables (``oracle/pc_cpu.c``, FP64 build) to the
FP64 DAG oracle, which never sees the index tables: same log-likelihoods, node marginals,
node flows, parameter flows (sum and input, including tied ones) and EM parameters, to FP64
rounding.  The FP32 build (the CPU baseline of bench.py) is checked against the FP64 build at
FP32 tolerance."""
import numpy as np
import pytest
import torch

import kat_circuits
import parity_utils as U
from oracle import dag_eval
from oracle.pc_cpu import CompiledCPU
from pyjuice_amd.model import TensorCircuit
from pyjuice_amd.nodes.nodes import node_iterator


def _circuit(name):
    if name == "simple_model":
        return kat_circuits.simple_model_dag(), [4, 4, 6, 6], dict(layer_sparsity_tol=0.1)
    if name == "tying_block1":
        return kat_circuits.tying_block1_dag(), [5, 5, 5, 5], dict(max_tied_ns_per_parflow_block=2)
    if name == "hclt_bern16":
        ns, _, cards = U.hclt_bernoulli()
        return ns, cards, dict()
    if name == "hclt_cat_bs32":
        ns, _, cards = U.hclt_categorical(L=64, block_size=32)
        return ns, cards, dict()
    if name == "pd_small":
        ns, cards = U.pd_small()
        return ns, cards, dict()
    if name == "hmm_small":
        ns, cards = U.hmm_small()
        return ns, cards, dict()
    raise KeyError(name)


CIRCUITS = ["simple_model", "tying_block1", "hclt_bern16", "hclt_cat_bs32", "pd_small", "hmm_small"]


@pytest.mark.parametrize("name", CIRCUITS)
def test_compiled_cpu_matches_dag_oracle(name):
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, **kw)
    pc.update_parameters()
    X = U.random_data(cards, 24, seed=5)
    cpu = CompiledCPU(pc)

    lls = cpu.forward(X)[0]
    vals = dag_eval.forward(ns, X)
    assert np.abs(lls - vals[ns][0]).max() <= 1e-9 * len(cards)
    for n in node_iterator(ns):
        if n.is_sum() or n.is_input():
            a, b = n._output_ind_range
            d = np.abs(cpu.node_mars[a:b] - vals[n])
            d[np.isinf(vals[n]) & np.isinf(cpu.node_mars[a:b])] = 0
            assert d.max() <= 1e-9 * len(n.scope), (n, d.max())

    cpu.backward()
    flows, pflows = dag_eval.backward(ns, X, vals)
    for n in node_iterator(ns):
        if n.is_sum() or n.is_input():
            a, b = n._output_ind_range
            assert U.rel_err(cpu.node_flows[a:b], flows[n]) <= 1e-9, n
    pc.param_flows = torch.from_numpy(cpu.param_flows.copy())
    pc.update_param_flows(origin_ns_only=False)
    for n in node_iterator(ns):
        if n.is_sum() and n.provided("_param_flows") and not n.is_tied():
            got = U.dense_pf_from_blocks(n, n._param_flows.numpy())
            ref = pflows[n]
            for m in node_iterator(ns):
                if m is not n and m.is_tied() and m.get_source_ns() is n and m._param_flow_range == n._param_flow_range:
                    ref = ref + pflows[m]
            # the per-node copies made by update_param_flows are FP32
            assert U.rel_err(got, ref) <= 1e-6, n
    for layer, d in zip(pc.input_layer_group, cpu.inputs):
        for n in layer.nodes:
            a, b = n._param_flow_range
            got = d["pflows"][a:b].reshape(n.num_nodes, n.dist.num_param_flows())
            ref = sum(pflows[m] for m in layer.nodes if m._param_flow_range == n._param_flow_range)
            assert U.rel_err(got, ref) <= 1e-9, n

    cpu.mini_batch_em(step_size=0.5, pseudocount=0.1)
    new = dag_eval.em_update(ns, pflows, step_size=0.5, pseudocount=0.1)
    pc.params.data = torch.from_numpy(cpu.params.astype(np.float32))
    pc.update_parameters()
    for layer, d in zip(pc.input_layer_group, cpu.inputs):
        for n in layer.nodes:
            if not n.is_tied():
                a, b = n._param_range
                n._params = torch.from_numpy(d["params"][a:b].copy())
    for s, p in new.items():
        got = U.dense_pf_from_blocks(s, s._params.numpy()) if s.is_sum() else \
            s._params.numpy().astype(np.float64).reshape(p.shape)
        assert np.abs(got - p).max() <= 1e-6, (s, np.abs(got - p).max())


def test_fp32_build_tracks_fp64_build():
    ns, cards, kw = _circuit("hclt_cat_bs32")
    pc = TensorCircuit(ns, **kw)
    X = U.random_data(cards, 32, seed=9)
    a, b = CompiledCPU(pc, np.float64), CompiledCPU(pc, np.float32)
    la, lb = a.forward(X)[0], b.forward(X)[0]
    assert np.abs(la - lb).max() <= 1e-4 * len(cards)
    a.backward()
    b.backward()
    assert U.rel_err(b.param_flows.astype(np.float64), a.param_flows) <= 1e-4


@pytest.mark.parametrize("name", ["simple_model", "hclt_bern16", "hmm_small"])
def test_compiled_cpu_signed_root_flows_are_weighted_gradients(name):
    """With signed root flows w, the C restatement's parameter flows are
    ``theta * d(sum_b w_b LL_b) / d theta`` (FP64 central differences of the weighted sum of its
    forward, which the test above pins to the DAG oracle): the oracle side of the signed-gradient
    adjoint."""
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, **kw)
    pc.update_parameters()
    X = U.random_data(cards, 12, seed=8)
    w = np.random.default_rng(4).normal(size=[1, 12])
    cpu = CompiledCPU(pc)
    cpu.forward(X)
    cpu.backward(root_flows=w)
    fused = cpu.param_flows.copy()
    for t, k, ch in cpu.fuse:
        fused[t:t + k] = sum(cpu.param_flows[c:c + k] for c in ch)
    theta = pc.params.detach().cpu().numpy().astype(np.float64)
    from pyjuice_amd.model.autograd import _sum_param_to_flow
    idx = _sum_param_to_flow(pc).numpy()
    g = np.random.default_rng(6)
    cand = np.nonzero((idx >= 0) & (theta > 0))[0]
    h = 1e-6

    def wll(th):
        pc.params.data = torch.from_numpy(th.astype(np.float32)).to(torch.float32)
        c2 = CompiledCPU(pc)
        c2.params = th.copy()
        return float((c2.forward(X)[0] * w[0]).sum())

    for k in g.choice(cand, size=min(8, cand.size), replace=False):
        tp, tm = theta.copy(), theta.copy()
        tp[k] *= 1 + h
        tm[k] *= 1 - h
        fd = (wll(tp) - wll(tm)) / (2 * h)
        assert abs(fd - fused[idx[k]]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, fused[idx[k]])
    pc.params.data = torch.from_numpy(theta.astype(np.float32))
