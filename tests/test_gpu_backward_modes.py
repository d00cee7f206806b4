"""``TensorCircuit.backward`` modes of the reference (``model/tensorcircuit.py:344-538``) against
the FP64 oracle:

* ``negate_pflows``: sum parameter flows are subtracted (``sum_layer.py:2166, 2279``), input
  flows unchanged;
* ``logspace_flows`` (with ``allow_modify_flows=False``, as the reference requires): the node-flow
  buffer holds log(flow) (``tests/structures/logspace_flows_test.py`` checks the same against
  its own recomputation), parameter flows unchanged;
* a per-layer ``propagation_alg`` list indexed like ``inner_layer_groups``
  (``tensorcircuit.py:270-280, 460-470``), forward and backward;
* sum-layer pre / post backward callbacks, called per sum layer top-down.

Tolerances as in ``test_gpu_parity.py``.
"""
import numpy as np
import pytest
import torch

import parity_utils as U
from oracle import dag_eval
from pyjuice_amd.model import TensorCircuit
from pyjuice_amd.nodes.nodes import node_iterator
from test_gpu_parity import TOL, _circuit

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def device():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def _sum_pf_errors(pc, ns, pflows, sign=1.0):
    pc.update_param_flows(origin_ns_only=False)
    worst = 0.0
    for n in node_iterator(ns):
        if n.is_sum() and n.provided("_param_flows") and not n.is_tied():
            got = U.dense_pf_from_blocks(n, n._param_flows.numpy())
            ref = pflows[n]
            for m in node_iterator(ns):
                if m is not n and m.is_tied() and m.get_source_ns() is n and m._param_flow_range == n._param_flow_range:
                    ref = ref + pflows[m]
            worst = max(worst, U.rel_err(got, sign * ref))
    return worst


def _input_pf_errors(pc, pflows):
    worst = 0.0
    for layer in pc.input_layer_group:
        pfi = layer.param_flows.detach().cpu().numpy().astype(np.float64)
        for n in layer.nodes:
            a, b = n._param_flow_range
            got = pfi[a:b].reshape(n.num_nodes, n.dist.num_param_flows())
            ref = sum(pflows[m] for m in layer.nodes if m._param_flow_range == n._param_flow_range)
            worst = max(worst, U.rel_err(got, ref))
    return worst


@pytest.mark.parametrize("name", ["hclt_cat_bs32", "hclt_cat_bs64", "pd_small", "hmm_small"])
@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp32_valu"])
@pytest.mark.parametrize("B", [64, 300])
def test_negate_pflows(name, precision, B, device):
    U.say(f"[negate] {name} {precision} B={B}")
    tol = TOL[precision][3]
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, precision=precision, **kw)
    pc.to(device)
    X = U.random_data(cards, B, seed=31)
    xt = torch.from_numpy(X).to(device)
    pc(xt)
    pc.backward(xt, flows_memory=0.0, negate_pflows=True)
    pc.update_parameters()
    vals = dag_eval.forward(ns, X)
    _, pflows = dag_eval.backward(ns, X, vals)
    assert _sum_pf_errors(pc, ns, pflows, sign=-1.0) <= tol
    assert _input_pf_errors(pc, pflows) <= tol
    # a plain backward on top of the negated one cancels the sum flows
    pc.backward(xt, flows_memory=1.0)
    ref_scale = max(float(np.abs(p).max()) for n, p in pflows.items() if n.is_sum())
    assert float(pc.param_flows.abs().max()) <= 2 * tol * ref_scale


@pytest.mark.parametrize("name", ["hclt_cat_bs32", "hclt_cat_bs64", "pd_small", "hmm_small", "hclt_bern16"])
@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_logspace_flows(name, precision, device):
    U.say(f"[logspace] {name} {precision}")
    tol = TOL[precision][3]
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, precision=precision, **kw)
    pc.to(device)
    X = U.random_data(cards, 64, seed=32)
    xt = torch.from_numpy(X).to(device)
    pc(xt)
    with pytest.raises(AssertionError):
        pc.backward(xt, logspace_flows=True)            # allow_modify_flows defaults to True
    pc.backward(xt, flows_memory=0.0, allow_modify_flows=False, logspace_flows=True)
    pc.update_parameters()
    vals = dag_eval.forward(ns, X)
    flows, pflows = dag_eval.backward(ns, X, vals)
    nf = pc.node_flows.detach().cpu().numpy().astype(np.float64)
    for n in node_iterator(ns):
        if n.is_sum() or n.is_input():
            a, b = n._output_ind_range
            lf = nf[a:b]
            ref = flows[n]
            assert np.array_equal(np.isneginf(lf), ref == 0.0), n
            assert U.rel_err(np.exp(lf), ref) <= tol, n
    assert _sum_pf_errors(pc, ns, pflows) <= tol
    assert _input_pf_errors(pc, pflows) <= tol


@pytest.mark.parametrize("name", ["simple_model", "hclt_cat_bs32", "hclt_cat_bs64", "hmm_small", "pd_small"])
@pytest.mark.parametrize("pattern", ["alternate", "mpe_top"])
def test_per_layer_propagation_algorithms(name, pattern, device):
    U.say(f"[per-layer alg] {name} {pattern}")
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, **kw)
    pc.to(device)
    groups = pc.inner_layer_groups
    sum_gids = [g for g in range(len(groups)) if groups[g].is_sum()]
    algs = ["LL"] * len(groups)
    for i, g in enumerate(sum_gids):
        if pattern == "alternate":
            algs[g] = ("LL", "GeneralLL", "MPE")[i % 3]
        else:
            algs[g] = "MPE" if g == sum_gids[-1] else "LL"
    alpha = 0.5
    vec_alg = dict()
    for g in sum_gids:
        for layer in groups[g]:
            for n in layer.nodes:
                vec_alg[n] = algs[g]
    X = U.random_data(cards, 48, seed=33)
    xt = torch.from_numpy(X).to(device)
    lls = pc(xt, propagation_alg=algs, alpha=alpha).detach().cpu().numpy()[:, 0]
    pc.update_parameters()
    vals = dag_eval.forward(ns, X, alg=vec_alg, alpha=alpha)
    assert np.abs(lls - vals[ns][0]).max() <= 1e-4 * len(cards)
    pc.backward(xt, flows_memory=0.0, allow_modify_flows=False, propagation_alg=algs, alpha=alpha)
    flows, pflows = dag_eval.backward(ns, X, vals, alg=vec_alg, alpha=alpha)
    nf = pc.node_flows.detach().cpu().numpy()
    for n in node_iterator(ns):
        if n.is_sum() or n.is_input():
            a, b = n._output_ind_range
            assert U.rel_err(nf[a:b], flows[n]) <= 1e-4, n
    assert _sum_pf_errors(pc, ns, pflows) <= 1e-4
    assert _input_pf_errors(pc, pflows) <= 1e-4


@pytest.mark.parametrize("name", ["hclt_cat_bs32", "hmm_small"])
def test_sum_layer_callbacks(name, device):
    U.say(f"[callbacks] {name}")
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, **kw)
    pc.to(device)
    X = U.random_data(cards, 64, seed=34)
    xt = torch.from_numpy(X).to(device)
    pc(xt)
    pc.backward(xt, flows_memory=0.0)
    want = pc.param_flows.detach().clone()
    calls = []

    def pre(layer, **kw):
        assert set(kw) == {"node_flows", "element_flows", "node_mars", "element_mars", "params", "param_flows"}
        calls.append(("pre", id(layer)))

    def post(layer, **kw):
        calls.append(("post", id(layer)))

    pc(xt)
    pc.backward(xt, flows_memory=0.0, sum_layer_pre_backward_callback=pre, sum_layer_post_backward_callback=post)
    order = [id(l) for g in reversed(pc.inner_layer_groups) if g.is_sum() for l in g]
    assert [c for k, c in calls if k == "pre"] == order
    assert [c for k, c in calls if k == "post"] == order
    assert U.rel_err(pc.param_flows.detach().cpu().numpy(), want.cpu().numpy()) <= 1e-5


@pytest.mark.parametrize("name", ["hclt_cat_bs64", "hclt_bern16"])
def test_backward_reuses_forward_transpose_only_for_the_same_unmodified_inputs(name, device):
    """The backward takes the forward's [V, B] data transpose when it is handed the very input
    tensor of the forward, unmodified; a different tensor, or the same one written in place, is
    transposed again.  Each case is compared with a backward that transposes its own copy."""
    U.say(f"[transpose reuse] {name}")
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, precision="bf16", **kw)
    pc.to(device)
    x1 = torch.from_numpy(U.random_data(cards, 64, seed=5)).to(device)
    x2 = torch.from_numpy(U.random_data(cards, 64, seed=6)).to(device)

    def input_pflows(x_fwd, x_bwd):
        pc(x_fwd)
        pc.backward(x_bwd, flows_memory=0.0)
        return [l.param_flows.detach().clone() for l in pc.input_layer_group]

    def close(a, b):
        return all(torch.allclose(u, v, rtol=1e-6, atol=1e-6) for u, v in zip(a, b))

    same = input_pflows(x1, x1)
    assert close(same, input_pflows(x1, x1.clone()))
    # x1 overwritten with x2's values after the forward: the backward must see x2
    xm = x1.clone()
    pc(xm)
    xm.copy_(x2)
    pc.backward(xm, flows_memory=0.0)
    written = [l.param_flows.detach().clone() for l in pc.input_layer_group]
    assert close(written, input_pflows(x1, x2.clone()))
    assert not close(written, same)
