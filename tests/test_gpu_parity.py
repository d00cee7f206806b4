"""HIP path against the FP64 oracle: forward, flow backward, parameter flows, EM, sampling.

Tolerances against the FP64 oracle (``TOL`` below), FP32 / reduced-precision MFMA forms:
* log-likelihood: |dLL| <= 1e-4 / 1e-3 nats per variable per sample;
* node marginals of every vector: |d| <= (1e-4 / 1e-3) * (scope size) + (1e-5 / 4e-3);
* node flows and parameter flows: max error <= 1e-4 / 2e-2 relative to the largest value;
* EM parameters: |d| <= 2e-6 / 5e-4 absolute;
* samples: bit-identical given the same uniforms.

Batches: 64 and 13 on every circuit, 300 and 512 on the circuits that run the MFMA kernels.
Edge cases: zero parameters (``zero_param_mask``), EM with ``keep_zero_params``, impossible
observations (out-of-range categories, so whole LL columns are -inf), and two Categorical vectors
with different category counts over one variable.
"""
import numpy as np
import pytest
import torch

import kat_circuits
import parity_utils as U
from oracle import dag_eval
from oracle.pc_cpu import CompiledCPU
from pyjuice_amd.model import TensorCircuit
from pyjuice_amd.nodes.nodes import node_iterator

pytestmark = pytest.mark.gpu


def _circuit(name):
    if name == "simple_model":
        return kat_circuits.simple_model_dag(), [4, 4, 6, 6], dict(layer_sparsity_tol=0.1)
    if name == "tying_block1":
        return kat_circuits.tying_block1_dag(), [5, 5, 5, 5], dict(max_tied_ns_per_parflow_block=2)
    if name == "hclt_bern16":
        ns, x, cards = U.hclt_bernoulli()
        return ns, cards, dict()
    if name == "hclt_cat_bs32":
        ns, x, cards = U.hclt_categorical(L=64, block_size=32)
        return ns, cards, dict()
    if name == "hclt_cat_bs64":
        ns, x, cards = U.hclt_categorical(L=128, block_size=64, K=16)
        return ns, cards, dict()
    if name == "hclt_cat_bs128":
        ns, x, cards = U.hclt_categorical(num_vars=6, L=128, K=10)
        return ns, cards, dict()
    if name == "hclt_cat_L256":    # one 256 x 256 group per tree node (the chain kernels' largest)
        ns, x, cards = U.hclt_categorical(num_vars=12, L=256, K=8)
        return ns, cards, dict()
    if name == "pd_small":
        ns, cards = U.pd_small()
        return ns, cards, dict()
    if name == "hmm_small":
        ns, cards = U.hmm_small()
        return ns, cards, dict()
    if name == "pd_sparse":
        ns, cards = U.pd_sparse()
        return ns, cards, dict()
    if name.startswith("wide_sums"):   # sum EM beyond the register-resident child count
        ns, cards = U.wide_sums(block_size=int(name[len("wide_sums"):]))
        return ns, cards, dict()
    if name == "hclt_pruned":      # several backward partitions in two sum layers
        ns, x, cards = U.hclt_pruned(num_vars=10, L=128)
        return ns, cards, dict()
    if name == "hclt_pruned_fw":   # several forward partitions in four sum layers
        ns, x, cards = U.hclt_pruned(num_vars=12, L=64)
        return ns, cards, dict()
    raise KeyError(name)


CIRCUITS = ["simple_model", "tying_block1", "hclt_bern16", "hclt_cat_bs32", "hclt_cat_bs64", "hclt_cat_bs128",
            "pd_small", "pd_sparse", "hmm_small", "hclt_pruned", "hclt_pruned_fw",
            "wide_sums64", "wide_sums32"]

# (LL nats/var, node-mar tol per scope variable, node-mar offset, relative flow / parameter-flow
# tol, EM abs tol).  A single BF16 sum node can be off by ~2^-8 (operand rounding), hence the
# per-node offset; the log-likelihood bound is the contract (1e-3 nats/var vs 1e-4 in FP32).
TOL = {"fp32": (1e-4, 1e-4, 1e-5, 1e-4, 2e-6), "fp32_valu": (1e-4, 1e-4, 1e-5, 1e-4, 2e-6),
       "bf16": (1e-3, 1e-3, 4e-3, 2e-2, 5e-4), "fp16": (1e-3, 1e-3, 4e-3, 2e-2, 5e-4)}


@pytest.fixture(scope="module")
def device():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


# The benchmark batches: 300 (a partial 64-column tile, several batch tiles and parameter-flow
# reduction stages) and 512 (the headline batch) on the circuits that take the MFMA kernels.
BIG = ["hclt_cat_bs64", "hclt_cat_bs128", "pd_small", "hmm_small", "hclt_pruned"]
CASES = [(n, B) for n in CIRCUITS for B in (64, 13)] + [(n, B) for n in BIG for B in (300, 512)]


@pytest.mark.parametrize("name,B", CASES)
@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp16", "fp32_valu"])
def test_forward_backward_em(name, B, precision, device):
    U.say(f"[parity] {name} B={B} {precision}")
    t_ll, t_mar, t_off, t_flow, t_em = TOL[precision]
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, precision=precision, **kw)
    pc.to(device)
    X = U.random_data(cards, B, seed=11)
    xt = torch.from_numpy(X).to(device)

    lls = pc(xt).detach().cpu().numpy()[:, 0]
    pc.update_parameters()
    vals = dag_eval.forward(ns, X)
    V = len(cards)
    assert np.abs(lls - vals[ns][0]).max() <= t_ll * V, np.abs(lls - vals[ns][0]).max()
    nm = pc.node_mars.detach().cpu().numpy()
    for n in node_iterator(ns):
        if n.is_sum() or n.is_input():
            a, b = n._output_ind_range
            tol = t_mar * len(n.scope) + t_off
            d = np.abs(nm[a:b] - vals[n])
            d[np.isinf(vals[n]) & np.isinf(nm[a:b])] = 0
            assert d.max() <= tol, (n, d.max())

    # flows without the in-place log transform, so node flows can be compared
    pc.backward(xt, allow_modify_flows=False, flows_memory=0.0)
    flows, pflows = dag_eval.backward(ns, X, vals)
    nf = pc.node_flows.detach().cpu().numpy()
    for n in node_iterator(ns):
        if n.is_sum() or n.is_input():
            a, b = n._output_ind_range
            assert U.rel_err(nf[a:b], flows[n]) <= t_flow, (n, U.rel_err(nf[a:b], flows[n]))
    pc.update_param_flows(origin_ns_only=False)
    for n in node_iterator(ns):
        if n.is_sum() and n.provided("_param_flows") and not n.is_tied():
            got = U.dense_pf_from_blocks(n, n._param_flows.numpy())
            ref = pflows[n]
            # flows of the tied copies sharing this range are accumulated in it too
            for m in node_iterator(ns):
                if m is not n and m.is_tied() and m.get_source_ns() is n and m._param_flow_range == n._param_flow_range:
                    ref = ref + pflows[m]
            assert U.rel_err(got, ref) <= t_flow, (n, U.rel_err(got, ref))
    for layer in pc.input_layer_group:
        pfi = layer.param_flows.detach().cpu().numpy().astype(np.float64)
        for n in layer.nodes:
            a, b = n._param_flow_range
            K = n.dist.num_param_flows()
            got = pfi[a:b].reshape(n.num_nodes, K)
            ref = sum(pflows[m] for m in layer.nodes if m._param_flow_range == n._param_flow_range)
            assert U.rel_err(got, ref) <= t_flow, (n, U.rel_err(got, ref))

    # default backward (in-place log(flow) - mar on sum nodes) must give the same parameter flows
    pf_plain = pc.param_flows.detach().clone()
    pc.backward(xt, flows_memory=0.0)
    assert U.rel_err(pc.param_flows.detach().cpu().numpy(), pf_plain.cpu().numpy()) <= t_flow * 0.1

    step, pcount = 0.5, 0.1
    pc.mini_batch_em(step_size=step, pseudocount=pcount)
    new = dag_eval.em_update(ns, pflows, step_size=step, pseudocount=pcount)
    pc.update_parameters()
    for s, p in new.items():
        if s.is_sum():
            got = U.dense_pf_from_blocks(s, s._params.numpy())
        else:
            got = s._params.numpy().astype(np.float64).reshape(p.shape)
        assert np.abs(got - p).max() <= t_em, (s, np.abs(got - p).max())


@pytest.mark.parametrize("name", ["simple_model", "hclt_bern16", "hclt_cat_bs32", "pd_small", "hmm_small",
                                  "wide_sums32", "hmm_wide"])
def test_sampling_bit_exact(name, device):
    """Caller uniforms: samples bit-identical to the oracle's (wide_sums32: 640 children, and
    hmm_wide: 1024 children per sum node, scanned in two 512-child chunks)."""
    U.say(f"[sample] {name}")
    if name == "hmm_wide":
        ns, cards = U.hmm_small(T=3, L=1024, V=8)
        kw = dict()
    else:
        ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, **kw)
    pc.to(device)
    S = 200
    u = torch.rand([pc.num_nodes, S], generator=torch.Generator().manual_seed(3))
    smp = pyjuice_sample(pc, u.to(device)).cpu().numpy()
    pc.update_parameters()
    ref = dag_eval.sample(ns, u.numpy(), pc.num_vars)
    assert np.array_equal(smp.astype(np.int64), ref), np.argwhere(smp != ref)[:5]


@pytest.mark.parametrize("name", ["simple_model", "hclt_bern16", "hclt_cat_bs64", "pd_sparse", "hmm_small", "wide_sums32"])
def test_sampling_seeded_and_graph_replays(name, device):
    """In-kernel Philox uniforms: bit-identical to the oracle drawing the same counter-based
    uniforms; a recorded graph replayed under new seeds, with eager passes in between, gives each
    seed's samples."""
    import pyjuice_amd as juice
    from oracle.dag_eval import philox_uniforms
    U.say(f"[sample seeded] {name}")
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, **kw)
    pc.to(device)
    pc.update_parameters()
    S = 300
    ref = lambda seed: dag_eval.sample(ns, (lambda r, c: philox_uniforms(seed, r, c), S), pc.num_vars)
    for seed in (5, (1 << 40) + 17):
        got = juice.queries.sample(pc, num_samples=S, seed=seed).cpu().numpy().astype(np.int64)
        assert np.array_equal(got, ref(seed))
    first = juice.queries.sample(pc, num_samples=S, seed=21, record_cudagraph=True).cpu().numpy().astype(np.int64)
    assert np.array_equal(first, ref(21))
    for seed in (22, 23):
        got = juice.queries.sample(pc, num_samples=S, seed=seed).cpu().numpy().astype(np.int64)   # replays
        assert np.array_equal(got, ref(seed)), seed
        eager = juice.queries.sample(pc, num_samples=S, seed=seed + 100, apply_cudagraph=False)
        assert np.array_equal(eager.cpu().numpy().astype(np.int64), ref(seed + 100))
    assert int(pc._sample_buffers["node_act"].sum()) == 0 and int(pc._sample_buffers["elem_act"].sum()) == 0


def pyjuice_sample(pc, u):
    import pyjuice_amd as juice
    return juice.queries.sample(pc, num_samples=u.size(1), uniforms=u)


@pytest.mark.parametrize("precision,ki", [(0, 2), (1, 16), (2, 16)])
def test_mfma_lane_layout(precision, ki, device):
    """The operand / accumulator lane layouts assumed by the MFMA kernels reproduce A @ B exactly
    (small integers are exact in BF16 / FP16 / FP32)."""
    import ctypes
    from pyjuice_amd import _native as N
    U.say(f"[mfma layout] precision={precision}")
    g = torch.Generator().manual_seed(precision)
    A = torch.randint(-4, 5, [32, ki], generator=g).float().to(device)
    Bm = torch.randint(-4, 5, [ki, 32], generator=g).float().to(device)
    C = torch.zeros([32, 32], device=device)
    N.call("pjamd_debug_mfma_probe", precision, N.ptr(A), N.ptr(Bm), N.ptr(C), N.stream())
    torch.cuda.synchronize()
    assert torch.equal(C.cpu(), (A.cpu() @ Bm.cpu()))


@pytest.mark.parametrize("precision", [1, 2])
def test_mfma16_lane_layout(precision, device):
    """The 16x16x32 operand / accumulator lane layouts assumed by the chain kernels reproduce
    A @ B exactly (small integers are exact in BF16 / FP16)."""
    from pyjuice_amd import _native as N
    U.say(f"[mfma16 layout] precision={precision}")
    g = torch.Generator().manual_seed(10 + precision)
    A = torch.randint(-4, 5, [16, 32], generator=g).float().to(device)
    Bm = torch.randint(-4, 5, [32, 16], generator=g).float().to(device)
    C = torch.zeros([16, 16], device=device)
    N.call("pjamd_debug_mfma16_probe", precision, N.ptr(A), N.ptr(Bm), N.ptr(C), N.stream())
    torch.cuda.synchronize()
    assert torch.equal(C.cpu(), (A.cpu() @ Bm.cpu()))


@pytest.mark.parametrize("name", ["hclt_cat_bs64", "hmm_small", "simple_model"])
def test_param_flow_accumulation(name, device):
    """init_param_flows(0) marks the accumulators zero: the next backward stores into the slots it
    owns; later backwards add.  Two identical backwards after clearing give twice one backward's
    parameter flows (sum and input layers), and reading the flows right after clearing gives 0."""
    U.say(f"[pflow accumulate] {name}")
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, precision="bf16", **kw)
    pc.to(device)
    X = torch.from_numpy(U.random_data(cards, 32, seed=4)).to(device)
    pc(X)
    pc.backward(X, flows_memory=0.0)
    one = pc.param_flows.detach().clone()
    one_in = [l.param_flows.detach().clone() for l in pc.input_layer_group]
    pc.init_param_flows(0.0)
    assert float(pc.param_flows.abs().max()) == 0.0
    assert all(float(l.param_flows.abs().max()) == 0.0 for l in pc.input_layer_group)
    pc.init_param_flows(0.0)
    pc.backward(X, flows_memory=1.0)
    pc.backward(X, flows_memory=1.0)
    assert U.rel_err(pc.param_flows.cpu().numpy(), 2 * one.cpu().numpy()) <= 1e-5
    for l, o in zip(pc.input_layer_group, one_in):
        assert U.rel_err(l.param_flows.cpu().numpy(), 2 * o.cpu().numpy()) <= 1e-5


@pytest.mark.parametrize("name", ["pd_sparse", "hclt_cat_bs32", "hmm_small"])
@pytest.mark.parametrize("precision", ["bf16", "fp32_valu"])
def test_backward_is_repeatable(name, precision, device):
    """Two identical flow backwards give the same parameter flows (up to the order of FP64 adds):
    no pass may read buffer rows that this backward did not write."""
    U.say(f"[repeatable] {name} {precision}")
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, precision=precision, **kw)
    pc.to(device)
    X = torch.from_numpy(U.random_data(cards, 48, seed=6)).to(device)
    outs = []
    for _ in range(2):
        pc(X)
        pc.backward(X, flows_memory=0.0)
        outs.append((pc.param_flows.detach().cpu().numpy().copy(),
                     [l.param_flows.detach().cpu().numpy().copy() for l in pc.input_layer_group]))
    assert U.rel_err(outs[0][0], outs[1][0]) <= 1e-6
    for a, b in zip(outs[0][1], outs[1][1]):
        assert U.rel_err(a, b) <= 1e-6


@pytest.mark.parametrize("name", ["hclt_cat_bs32", "hclt_cat_bs64", "hclt_cat_bs128", "hclt_cat_L256", "hmm_small"])
@pytest.mark.parametrize("precision", ["bf16", "fp16"])
@pytest.mark.parametrize("B", [64, 300])
def test_chain_runs_match_layerwise(name, precision, B, device):
    """Runs of narrow levels in one launch per pass (pjamd_chain_pack / _forward / _backward)
    against one launch per layer on the same circuit and data, and against the FP64 oracle:
    log-likelihoods, node flows of the modify-free backward, sum and input parameter flows of the
    default (modifying) backward.  Both paths round the same operands to 16 bits; the chain's
    column maxima are taken per group, the layer-wise kernels' per stage, so only the FP32
    accumulation order and the scaling differ.  B = 300 leaves a partial 32-column slice."""
    U.say(f"[chain vs layerwise] {name} {precision} B={B}")
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, precision=precision, **kw)
    pc.to(device)
    x = U.random_data(cards, B, seed=8)
    X = torch.from_numpy(x).to(device)
    pc.use_chains = True
    plan = pc._chain_plan(precision, B)
    assert plan is not None and plan.runs, "no run planned"
    out = dict()
    for chains in (False, True):
        pc.use_chains = chains
        lls = pc(X).detach().double().cpu().numpy()
        pc.backward(X, flows_memory=0.0, allow_modify_flows=False)
        nf = pc.node_flows.detach().double().cpu().numpy()
        pc.backward(X, flows_memory=0.0)
        out[chains] = (lls, nf, pc.param_flows.detach().double().cpu().numpy(),
                       [l.param_flows.detach().double().cpu().numpy() for l in pc.input_layer_group])
    tol = 2e-4
    (l0, f0, p0, i0), (l1, f1, p1, i1) = out[False], out[True]
    assert np.abs(l0 - l1).max() <= tol * len(cards)
    assert U.rel_err(f1, f0) <= 10 * tol
    assert U.rel_err(p1, p0) <= 10 * tol
    for a, b in zip(i1, i0):
        assert U.rel_err(a, b) <= 10 * tol
    ref = CompiledCPU(pc, np.float64).forward(x)[0]
    assert np.abs(l1[:, 0] - ref.reshape(-1)).max() <= 1e-3 * len(cards)


@pytest.mark.parametrize("name", ["hclt_cat_bs64", "hmm_small", "pd_small", "hclt_cat_L256+chains",
                                  "hclt_cat_bs64+chains"])
def test_train_step_graph_replays(name, device):
    """A whole training step (clear flows, forward, flow backward, EM) recorded as one HIP graph:
    each replay, on a new mini-batch and with eager work and synchronisations between replays,
    gives the same log-likelihoods and updated parameters as the eager step on a twin circuit
    (with chain runs in the graph: the parameter pack and the chain launches, against the
    layer-wise eager step)."""
    import pyjuice_amd as juice
    U.say(f"[step graph] {name}")
    chains = name.endswith("+chains")
    ns, cards, kw = _circuit(name.replace("+chains", ""))
    pc_g = TensorCircuit(ns, precision="bf16", **kw)
    pc_e = TensorCircuit(ns, precision="bf16", **kw)
    pc_g.to(device)
    pc_e.to(device)
    # with chains, the graph (pack, chain launches) is checked against the layer-wise eager step
    pc_g.use_chains = chains
    if chains:
        assert pc_g._chain_plan("bf16", 64) is not None, "no run planned"
    pc_e.params.data.copy_(pc_g.params)
    for a, b in zip(pc_e.input_layer_group, pc_g.input_layer_group):
        a.params.data.copy_(b.params)
    batches = [torch.from_numpy(U.random_data(cards, 64, seed=20 + i)).to(device) for i in range(4)]
    opt = juice.optim.CircuitOptimizer(pc_e, lr=0.1, pseudocount=0.1)

    def eager(x):
        opt.zero_grad()
        with torch.no_grad():
            lls = pc_e(x)
            pc_e.backward(x, flows_memory=0.0)
        opt.step()
        return lls

    # same kernels: rounding-level agreement; chains against layer-wise: the 16-bit operand
    # tolerance of test_chain_runs_match_layerwise, scaled by the EM step size
    t_ll, t_p = (2e-4, 1e-3) if chains else (1e-4, 1e-5)
    lls0 = eager(batches[0])                      # the graph's own eager warm-up step
    step = juice.optim.TrainStepGraph(pc_g, batches[0], lr=0.1, pseudocount=0.1)
    assert np.abs(step.warmup_lls.cpu().numpy() - lls0.cpu().numpy()).max() <= t_ll * len(cards)
    for x in batches[1:]:
        want = eager(x).detach().cpu().numpy()
        got = step(x).detach().cpu().numpy()
        torch.cuda.synchronize()
        _ = (pc_g.params.sum() + 0).item()        # eager work between replays
        assert np.abs(got - want).max() <= t_ll * len(cards)
        assert U.rel_err(pc_g.params.detach().cpu().numpy(), pc_e.params.detach().cpu().numpy()) <= t_p
        for a, b in zip(pc_g.input_layer_group, pc_e.input_layer_group):
            assert U.rel_err(a.params.detach().cpu().numpy(), b.params.detach().cpu().numpy()) <= t_p


@pytest.mark.parametrize("name", ["hclt_cat_bs64", "hclt_cat_L256", "hclt_bern16"])
@pytest.mark.parametrize("B,keep_zero,pcount", [(64, False, 0.1), (300, True, 0.5), (2048, False, 0.0)])
def test_fused_input_em_matches_two_passes(name, B, keep_zero, pcount, device):
    """The input layers' EM run inside the flow backward (pjamd_input_backward_em, what the
    training-step graph uses) against backward then mini_batch_em: the same stored parameter
    flows and updated parameters, up to summation order (the batch sort of the categorical
    backward places a category's samples in atomic order, so FP32 flow sums can differ in the
    last bit between any two runs)."""
    U.say(f"[fused input EM] {name} B={B} keep_zero={keep_zero}")
    ns, cards, kw = _circuit(name)
    pcs = [TensorCircuit(ns, precision="bf16", **kw) for _ in range(2)]
    for pc in pcs:
        pc.to(device)
    pcs[1].params.data.copy_(pcs[0].params)
    for a, b in zip(pcs[1].input_layer_group, pcs[0].input_layer_group):
        a.params.data.copy_(b.params)
    if keep_zero:       # some zero parameters to keep
        for pc in pcs:
            for layer in pc.input_layer_group:
                layer.params.data[::7] = 0.0
    X = torch.from_numpy(U.random_data(cards, B, seed=31)).to(device)
    fused = [l.fused_em_ok(B, l._output_ind_range[0]) for l in pcs[0].input_layer_group]
    if name != "hclt_bern16":
        assert any(fused), "no input layer takes the fused form"
    else:
        assert not any(fused)
    with torch.no_grad():
        for i, pc in enumerate(pcs):
            pc.init_param_flows(flows_memory=0.0)
            pc(X)
            pc.backward(X, flows_memory=0.0, fused_em=(0.3, pcount, keep_zero) if i == 0 else None)
            if i == 0:
                assert [bool(l.__dict__.get("_em_done")) for l in pc.input_layer_group] == fused
            pc.mini_batch_em(step_size=0.3, pseudocount=pcount, keep_zero_params=keep_zero)
    for a, b in zip(pcs[0].input_layer_group, pcs[1].input_layer_group):
        fa, fb = a.param_flows.detach().cpu().numpy(), b.param_flows.detach().cpu().numpy()
        assert np.abs(fa - fb).max() <= 1e-6 * max(1.0, np.abs(fb).max()), np.abs(fa - fb).max()
        pa, pb = a.params.detach().cpu().numpy(), b.params.detach().cpu().numpy()
        assert np.abs(pa - pb).max() <= 1e-6 * max(1.0, np.abs(pb).max()), np.abs(pa - pb).max()
        if keep_zero and a._kind() == 0:      # Categorical EM keeps zero parameters at zero
            assert np.all(pa[::7] == 0.0)
    assert torch.equal(pcs[0].params, pcs[1].params)


@pytest.mark.parametrize("name", ["hclt_cat_bs32", "hmm_small", "pd_sparse", "tying_block1"])
@pytest.mark.parametrize("precision", ["bf16", "fp32_valu"])
def test_recompute_elements_matches_resident(name, precision, device):
    """Layer-wise recomputation (one shared element buffer, product groups recomputed in the flow
    backward, per-layer parameter flows) gives the same results as resident element slabs, with
    an element buffer of the widest group only."""
    U.say(f"[recompute] {name} {precision}")
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, precision=precision, **kw)
    pc.to(device)
    X = torch.from_numpy(U.random_data(cards, 40, seed=9)).to(device)
    out = dict()
    for rec in (False, True):
        pc.recompute_elements = rec
        lls = pc(X).detach().double().cpu().numpy()
        pc.backward(X, flows_memory=0.0)
        out[rec] = (lls, pc.param_flows.detach().double().cpu().numpy(),
                    [l.param_flows.detach().double().cpu().numpy() for l in pc.input_layer_group],
                    pc.element_mars.size(0))
    assert out[True][3] == pc.num_elements and out[False][3] == pc.num_element_rows
    assert out[True][3] <= out[False][3]
    assert np.abs(out[True][0] - out[False][0]).max() <= 1e-6 * len(cards)
    assert U.rel_err(out[True][1], out[False][1]) <= 1e-5
    for a, b in zip(out[True][2], out[False][2]):
        assert U.rel_err(a, b) <= 1e-5


@pytest.mark.parametrize("shape", [(6, 128, 10), (8, 256, 6)])
@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_product_fusion_matches_layerwise(shape, precision, device):
    """The product layers computed inside the next sum forward (pjamd_sum_forward_prod) give the
    same node marginals, element rows and parameter flows, bit for bit, as separate product and
    sum launches (the same additions in the same order), and the log-likelihoods match the FP64
    oracle."""
    num_vars, L, K = shape
    U.say(f"[prod fusion] HCLT V={num_vars} L={L} {precision}")
    ns, x, cards = U.hclt_categorical(num_vars=num_vars, L=L, K=K)
    pc = TensorCircuit(ns, precision=precision)
    pc.to(device)
    assert pc._prod_fusion["groups"] and pc._fuse_now(pc._prod_fusion, min(pc._prod_fusion["groups"]), 64)
    X = U.random_data(cards, 64, seed=21)
    xt = torch.from_numpy(X).to(device)
    out = dict()
    for fuse in (True, False):    # fused first: element rows it failed to store would differ
        pc.fuse_products = fuse
        lls = pc(xt).detach().clone()
        nm, em = pc.node_mars.detach().clone(), pc.element_mars.detach().clone()
        pc.backward(xt, flows_memory=0.0)
        out[fuse] = (lls, nm, em, pc.param_flows.detach().clone())
    for a, b in zip(out[False], out[True]):
        assert torch.equal(a, b)
    pc.update_parameters()
    ref = dag_eval.forward(ns, X)[ns][0]
    assert np.abs(out[True][0].cpu().numpy()[:, 0] - ref).max() <= 1e-3 * num_vars


def _completions(X, miss, cards):
    """Every completion of each sample's missing entries: (rows [R, V], owner sample [R])."""
    import itertools
    rows, owner = [], []
    for b in range(X.shape[0]):
        mv = np.flatnonzero(miss[b])
        for vals in itertools.product(*[range(cards[v]) for v in mv]):
            x = X[b].copy()
            x[mv] = vals
            rows.append(x)
            owner.append(b)
    return np.asarray(rows), np.asarray(owner)


@pytest.mark.parametrize("name", ["hclt_bern16", "hclt_cat_bs32", "pd_small"])
@pytest.mark.parametrize("mask_kind", ["per_var", "per_sample"])
@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_marginal_missing_mask(name, mask_kind, precision, device):
    """Marginal queries (juice.queries.marginal, missing_mask) against brute force on the FP64
    oracle: log P(x_o) = logsumexp over every completion of the missing entries; the parameter
    flows of the marginal backward = the completions' flows weighted by their posterior
    P(x_m | x_o) (input categories of a missing entry take theta-weighted flows)."""
    import pyjuice_amd as juice
    U.say(f"[marginal] {name} {mask_kind} {precision}")
    t_ll, _, _, t_flow, _ = TOL[precision]
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, precision=precision, **kw)
    pc.to(device)
    V, B = len(cards), 8
    X = U.random_data(cards, B, seed=17)
    g = np.random.default_rng(5)
    if mask_kind == "per_var":
        mv = np.zeros(V, dtype=bool)
        mv[g.choice(V, size=2, replace=False)] = True
        miss = np.broadcast_to(mv, (B, V))
        mask = torch.from_numpy(mv)
    else:
        miss = np.zeros([B, V], dtype=bool)
        for b in range(B):
            miss[b, g.choice(V, size=1 + b % 2, replace=False)] = True
        mask = torch.from_numpy(miss)
    xt = torch.from_numpy(X).to(device)
    lls = juice.queries.marginal(pc, xt, missing_mask=mask.to(device)).detach().cpu().numpy()[:, 0]
    pc.backward(xt, missing_mask=mask.to(device), allow_modify_flows=False, flows_memory=0.0)
    pc.update_parameters()
    pc.update_param_flows(origin_ns_only=False)

    R, owner = _completions(X, miss, cards)
    vals = dag_eval.forward(ns, R)
    ll_r = vals[ns][0]
    ref = np.full(B, -np.inf)
    for b in range(B):
        ref[b] = np.logaddexp.reduce(ll_r[owner == b])
    assert np.abs(lls - ref).max() <= t_ll * V, np.abs(lls - ref).max()
    w = np.exp(ll_r - ref[owner])[None, :]
    flows, pflows = dag_eval.backward(ns, R, vals, root_flows=w)
    for n in node_iterator(ns):
        if n.is_sum() and n.provided("_param_flows") and not n.is_tied():
            got = U.dense_pf_from_blocks(n, n._param_flows.numpy())
            assert U.rel_err(got, pflows[n]) <= t_flow, (n, U.rel_err(got, pflows[n]))
    for layer in pc.input_layer_group:
        pfi = layer.param_flows.detach().cpu().numpy().astype(np.float64)
        for n in layer.nodes:
            a, b = n._param_flow_range
            K = n.dist.num_param_flows()
            got = pfi[a:b].reshape(n.num_nodes, K)
            if layer._kind() == 1:
                # Bernoulli: missing entries carry no flow, as in the reference (no masked flow
                # function); observed entries keep theirs (slot 0: x = 1, slot 1: x = 0)
                v = n.scope.to_list()[0]
                f = flows[n] * (~miss[owner, v])[None, :]
                want = np.stack([(f * (R[:, v] == 1)).sum(1), (f * (R[:, v] == 0)).sum(1)], axis=1)
            else:
                want = pflows[n]
            assert U.rel_err(got, want) <= t_flow, (n, U.rel_err(got, want))


def _all_assignments(cards):
    import itertools
    return np.asarray(list(itertools.product(*[range(k) for k in cards])), dtype=np.int64)


@pytest.mark.parametrize("name", ["simple_model", "hclt_cat_tiny", "hmm_tiny", "hclt_bern16"])
def test_top_down_probs_and_rescaled_em(name, device):
    """Top-down probabilities, their parameter-flow update, the partition function and
    step-size-rescaled EM (``top_down_prob.py:290-402``, ``tensorcircuit.py:566-578``) against
    enumeration on the FP64 oracle: the top-down probability of a node is its flow averaged over
    every assignment x weighted by p(x), and theta * probability is the expected parameter flow."""
    import pyjuice_amd as juice
    U.say(f"[top-down] {name}")
    if name == "hclt_cat_tiny":
        ns, _, cards = U.hclt_categorical(num_vars=4, K=3, L=32)
        kw = dict()
    elif name == "hmm_tiny":
        ns, cards = U.hmm_small(T=3, L=32, V=4)
        kw = dict()
    else:
        ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, **kw)
    pc.to(device)
    bern = any(layer._kind() == 1 for layer in pc.input_layer_group)
    Xall = _all_assignments(cards)
    vals = dag_eval.forward(ns, Xall)
    px = np.exp(vals[ns][0])
    assert abs(px.sum() - 1.0) < 1e-6      # parameters normalized in FP32
    flows, pf_exp = dag_eval.backward(ns, Xall, vals, root_flows=px[None, :])

    # partition function of the normalized circuit
    assert abs(float(juice.model.eval_partition_fn(pc)[0])) <= 1e-5

    # top-down probabilities and their parameter-flow update, from zero flows
    pc.init_param_flows(0.0)
    s = 0.375
    juice.model.eval_top_down_probs(pc, update_pflow=True, scale=s)
    tdp = pc._single_node_flows.cpu().numpy().astype(np.float64)
    pc.update_param_flows(origin_ns_only=False)
    for n in node_iterator(ns):
        if n.is_sum() or n.is_input():
            a, b = n._output_ind_range
            assert U.rel_err(tdp[a:b], flows[n].sum(axis=1)) <= 1e-5, n
        if n.is_sum() and n.provided("_param_flows") and not n.is_tied():
            got = U.dense_pf_from_blocks(n, n._param_flows.numpy())
            want = pf_exp[n] + sum((pf_exp[m] for m in node_iterator(ns) if m is not n and m.is_tied()
                                    and m.get_source_ns() is n and m._param_flow_range == n._param_flow_range), 0)
            assert U.rel_err(got, s * want) <= 1e-5, n
    if not bern:
        for layer in pc.input_layer_group:
            pfi = layer.param_flows.detach().cpu().numpy().astype(np.float64)
            for n in layer.nodes:
                a, b = n._param_flow_range
                got = pfi[a:b].reshape(n.num_nodes, n.dist.num_param_flows())
                want = sum(pf_exp[m] for m in layer.nodes if m._param_flow_range == n._param_flow_range)
                assert U.rel_err(got, s * want) <= 1e-5, n

    # step-size-rescaled EM on a mini-batch
    X = U.random_data(cards, 16, seed=31)
    xt = torch.from_numpy(X).to(device)
    pc.init_param_flows(0.0)
    pc._cum_flow = 0.0
    pc(xt)
    pc.backward(xt, flows_memory=0.0)
    vb = dag_eval.forward(ns, X)
    _, pf_b = dag_eval.backward(ns, X, vb)
    step, pcount = 0.25, 0.1
    pc.mini_batch_em(step_size=step, pseudocount=pcount, step_size_rescaling=True)
    assert pc._cum_flow == 0.0
    comb = {n: step / 16 * pf_b[n] + (1.0 - step) * pf_exp[n] for n in pf_b}
    if bern:      # Bernoulli inputs take no top-down flows (no masked flow function, as the reference)
        for layer in pc.input_layer_group:
            for n in layer.nodes:
                comb[n] = step / 16 * pf_b[n]
    new = dag_eval.em_update(ns, comb, step_size=1.0, pseudocount=pcount)
    pc.update_parameters()
    for n, p in new.items():
        got = U.dense_pf_from_blocks(n, n._params.numpy()) if n.is_sum() else \
            n._params.numpy().astype(np.float64).reshape(p.shape)
        assert np.abs(got - p).max() <= 2e-6, (n, np.abs(got - p).max())


@pytest.mark.parametrize("L", [512, 1024])
@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_split_reduction_forward(L, precision, device):
    """pjamd_sum_forward_split (a node block's reduction split over workgroups, the last one
    combining the partial log-sum-exps) against one workgroup per tile, and the log-likelihood
    against the FP64 oracle; two calls in a row give the same result."""
    from pyjuice_amd.layer.sum_layer import SumLayer
    U.say(f"[split forward] HMM L={L} {precision}")
    ns, cards = U.hmm_small(T=4, L=L, V=16)
    pc = TensorCircuit(ns, precision=precision)
    pc.to(device)
    X = U.random_data(cards, 64, seed=41)
    xt = torch.from_numpy(X).to(device)
    out = []
    saved = SumLayer.split_max_tiles
    try:
        for max_tiles in (saved, saved, 0):
            SumLayer.split_max_tiles = max_tiles
            for g in pc.inner_layer_groups:
                for layer in g:
                    layer.__dict__.pop("_split_cache", None)
            out.append(pc(xt).detach().double().cpu().numpy()[:, 0])
            split_used = any(v is not None for g in pc.inner_layer_groups for layer in g
                             for v in layer.__dict__.get("_split_cache", {}).values())
            assert split_used == (max_tiles > 0)
    finally:
        SumLayer.split_max_tiles = saved
    assert np.array_equal(out[0], out[1])
    # 16-bit operands are rounded relative to each split's own column max, so the two forms
    # differ at operand precision; both are held to the 16-bit contract against FP64 below
    assert np.abs(out[0] - out[2]).max() <= 1e-3 * len(cards)
    pc.update_parameters()
    ref = dag_eval.forward(ns, X)[ns][0]
    assert np.abs(out[0] - ref).max() <= 1e-3 * len(cards)
    assert np.abs(out[2] - ref).max() <= 1e-3 * len(cards)


@pytest.mark.parametrize("L", [512, 1024])
@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_split_reduction_backward(L, precision, device):
    """pjamd_sum_backward_ele_split (the fused element-flow pass with its parent reduction split
    over workgroups, partials combined by a second kernel that writes the children's flows):
    parameter flows against one workgroup per tile and against the FP64 oracle."""
    from pyjuice_amd.layer.sum_layer import SumLayer
    U.say(f"[split backward] HMM L={L} {precision}")
    ns, cards = U.hmm_small(T=4, L=L, V=16)
    pc = TensorCircuit(ns, precision=precision)
    pc.to(device)
    X = U.random_data(cards, 64, seed=43)
    xt = torch.from_numpy(X).to(device)
    res = dict()
    saved = SumLayer.split_max_tiles
    try:
        for max_tiles in (saved, 0):
            SumLayer.split_max_tiles = max_tiles
            for g in pc.inner_layer_groups:
                for layer in g:
                    layer.__dict__.pop("_split_cache", None)
            pc(xt)
            pc.backward(xt, flows_memory=0.0)
            used = any(v is not None for g in pc.inner_layer_groups for layer in g
                       for k, v in layer.__dict__.get("_split_cache", {}).items() if k[0] == "bk")
            assert used == (max_tiles > 0)
            res[max_tiles] = (pc.param_flows.detach().double().cpu().numpy(),
                              [l.param_flows.detach().double().cpu().numpy() for l in pc.input_layer_group])
    finally:
        SumLayer.split_max_tiles = saved
    assert U.rel_err(res[saved][0], res[0][0]) <= 2e-2
    for a, b in zip(res[saved][1], res[0][1]):
        assert U.rel_err(a, b) <= 2e-2
    # input parameter flows of the split run (they take the children flows the split pass wrote)
    vals = dag_eval.forward(ns, X)
    flows, pflows = dag_eval.backward(ns, X, vals)
    got_in = res[saved][1]
    for layer, pfi in zip(pc.input_layer_group, got_in):
        for n in layer.nodes:
            a, b = n._param_flow_range
            got = pfi[a:b].reshape(n.num_nodes, n.dist.num_param_flows())
            want = sum(pflows[m] for m in layer.nodes if m._param_flow_range == n._param_flow_range)
            assert U.rel_err(got, want) <= 2e-2, n


@pytest.mark.parametrize("shape", [(6, 32, 10), (5, 128, 12)])
@pytest.mark.parametrize("precision", ["bf16", "fp32"])
def test_tied_copies_share_parameter_flow_tiles(shape, precision, device):
    """Batched parameter flows of a homogeneous HMM with its tied time steps summed inside each
    tile (one write per slot, PJAMD_PAR_DESC_FIELDS copies field) against one tile per time step
    adding atomically, and both against the FP64 oracle's tied parameter flows."""
    from pyjuice_amd.model.tensorcircuit import PAR_DESC_FIELDS
    T, L, V = shape
    U.say(f"[tied par tiles] HMM T={T} L={L} {precision}")
    ns, cards = U.hmm_small(T=T, L=L, V=V)
    pc = TensorCircuit(ns, precision=precision)
    pc.to(device)
    X = U.random_data(cards, 48, seed=5)
    xt = torch.from_numpy(X).to(device)
    out = dict()
    for tiles in (1 << 30, 1):
        pc.par_tied_tiles = tiles
        pc._par_batches = pc._build_par_batches(xt.device)
        copies = [int(t["desc"].view(-1, PAR_DESC_FIELDS)[:, 8].max()) for t in pc._par_batches["tables"].values()]
        assert (max(copies) >= T - 2) if tiles == 1 else (max(copies) == 1), copies
        pc(xt)
        pc.backward(xt, flows_memory=0.0)
        out[tiles] = pc.param_flows.detach().double().cpu().numpy()
    assert U.rel_err(out[1], out[1 << 30]) <= 1e-5
    # the last backward ran with the copies summed inside each tile
    vals = dag_eval.forward(ns, X)
    flows, pflows = dag_eval.backward(ns, X, vals)
    tol = TOL[precision][3]
    pc.update_param_flows(origin_ns_only=False)
    checked = 0
    for n in node_iterator(ns):
        if n.is_sum() and n.provided("_param_flows") and not n.is_tied():
            got = U.dense_pf_from_blocks(n, n._param_flows.numpy())
            ref = pflows[n]
            for m in node_iterator(ns):
                if m is not n and m.is_tied() and m.get_source_ns() is n and m._param_flow_range == n._param_flow_range:
                    ref = ref + pflows[m]
            assert U.rel_err(got, ref) <= tol, (n, U.rel_err(got, ref))
            checked += 1
    assert checked > 0


@pytest.mark.parametrize("name", ["hclt_cat_bs128", "hclt_cat_L256"])
@pytest.mark.parametrize("precision,B", [("bf16", 300), ("bf16", 13), ("fp16", 64), ("fp32", 300)])
def test_parameter_flow_tile_forms_agree(name, precision, B, device):
    """The batched parameter flows from 128 x 128 tiles (16-bit operands with B % 4 == 0: one
    workgroup per tile, 64-column stages, LDS-staged epilogue; otherwise four 64 x 64 workgroups
    per tile) against the 64 x 64 tile table, and both against the FP64 oracle."""
    U.say(f"[par tile forms] {name} {precision} B={B}")
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, precision=precision, **kw)
    pc.to(device)
    X = U.random_data(cards, B, seed=21)
    xt = torch.from_numpy(X).to(device)
    out = dict()
    for rows in (128, 64):
        pc.par_tile_rows = rows
        pc._par_batches = pc._build_par_batches(xt.device)
        if rows == 128:
            assert 128 in pc._par_batches["tables"], sorted(pc._par_batches["tables"])
        pc(xt)
        pc.backward(xt, flows_memory=0.0)
        out[rows] = pc.param_flows.detach().double().cpu().numpy()
    tol = TOL[precision][3]
    assert U.rel_err(out[128], out[64]) <= tol
    vals = dag_eval.forward(ns, X)
    flows, pflows = dag_eval.backward(ns, X, vals)
    pc.par_tile_rows = 128
    pc._par_batches = pc._build_par_batches(xt.device)
    pc(xt)
    pc.backward(xt, flows_memory=0.0)
    pc.update_param_flows(origin_ns_only=False)
    checked = 0
    for n in node_iterator(ns):
        if n.is_sum() and n.provided("_param_flows") and not n.is_tied():
            got = U.dense_pf_from_blocks(n, n._param_flows.numpy())
            assert U.rel_err(got, pflows[n]) <= tol, (n, U.rel_err(got, pflows[n]))
            checked += 1
    assert checked > 0


@pytest.mark.parametrize("name", ["simple_model", "tying_block1", "hclt_cat_bs32", "hclt_cat_bs64", "hmm_small",
                                  "pd_small", "pd_sparse"])
@pytest.mark.parametrize("alg,alpha", [("MPE", None), ("GeneralLL", 0.5), ("GeneralLL", 2.0)])
def test_propagation_algorithms(name, alg, alpha, device):
    """MPE and GeneralLL forward / flow backward / parameter flows (pjamd_sum_*_alg) against the
    FP64 oracle: log-values of every vector, node flows, sum and input parameter flows."""
    U.say(f"[propagation] {name} {alg} {alpha}")
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, **kw)
    pc.to(device)
    X = U.random_data(cards, 32, seed=23)
    xt = torch.from_numpy(X).to(device)
    extra = {} if alpha is None else {"alpha": alpha}
    lls = pc(xt, propagation_alg=alg, **extra).detach().cpu().numpy()[:, 0]
    pc.backward(xt, propagation_alg=alg, flows_memory=0.0, **extra)
    pc.update_parameters()
    a = 1.0 if alpha is None else alpha
    vals = dag_eval.forward(ns, X, alg=alg, alpha=a)
    flows, pflows = dag_eval.backward(ns, X, vals, alg=alg, alpha=a)
    V = len(cards)
    assert np.abs(lls - vals[ns][0]).max() <= 1e-4 * V
    nm = pc.node_mars.detach().cpu().numpy()
    nf = pc.node_flows.detach().cpu().numpy()
    for n in node_iterator(ns):
        if n.is_sum() or n.is_input():
            a0, a1 = n._output_ind_range
            d = np.abs(nm[a0:a1] - vals[n])
            d[np.isinf(vals[n]) & np.isinf(nm[a0:a1])] = 0
            assert d.max() <= 1e-4 * len(n.scope) + 1e-5, (n, d.max())
            assert U.rel_err(nf[a0:a1], flows[n]) <= 1e-4, (n, U.rel_err(nf[a0:a1], flows[n]))
    pc.update_param_flows(origin_ns_only=False)
    for n in node_iterator(ns):
        if n.is_sum() and n.provided("_param_flows") and not n.is_tied():
            got = U.dense_pf_from_blocks(n, n._param_flows.numpy())
            ref = pflows[n]
            for m in node_iterator(ns):
                if m is not n and m.is_tied() and m.get_source_ns() is n and m._param_flow_range == n._param_flow_range:
                    ref = ref + pflows[m]
            assert U.rel_err(got, ref) <= 1e-4, (n, U.rel_err(got, ref))
    for layer in pc.input_layer_group:
        pfi = layer.param_flows.detach().cpu().numpy().astype(np.float64)
        for n in layer.nodes:
            a0, a1 = n._param_flow_range
            got = pfi[a0:a1].reshape(n.num_nodes, n.dist.num_param_flows())
            ref = sum(pflows[m] for m in layer.nodes if m._param_flow_range == n._param_flow_range)
            assert U.rel_err(got, ref) <= 1e-4, (n, U.rel_err(got, ref))


@pytest.mark.parametrize("name", ["hclt_cat_bs32", "hmm_small", "pd_small"])
def test_general_ll_alpha_one_matches_ll_path(name, device):
    """The vector-ALU propagation kernels with GeneralLL alpha = 1 against the main LL path
    (MFMA / tiled kernels, fused backward): same log-likelihoods and parameter flows to FP32
    rounding."""
    U.say(f"[propagation vs LL] {name}")
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, precision="fp32", **kw)
    pc.to(device)
    xt = torch.from_numpy(U.random_data(cards, 32, seed=29)).to(device)
    out = dict()
    for alg in ("LL", "GeneralLL"):
        lls = pc(xt, propagation_alg=alg, alpha=1.0).detach().double().cpu().numpy()
        pc.backward(xt, propagation_alg=alg, alpha=1.0, flows_memory=0.0)
        out[alg] = (lls, pc.param_flows.detach().double().cpu().numpy(),
                    [l.param_flows.detach().double().cpu().numpy() for l in pc.input_layer_group])
    assert np.abs(out["LL"][0] - out["GeneralLL"][0]).max() <= 1e-5 * len(cards)
    assert U.rel_err(out["GeneralLL"][1], out["LL"][1]) <= 1e-4
    for a, b in zip(out["GeneralLL"][2], out["LL"][2]):
        assert U.rel_err(a, b) <= 1e-4


@pytest.mark.parametrize("name", ["hclt_bern16", "hclt_cat_bs32"])
def test_input_dtypes(name, device):
    """Data given as bool / uint8 / int8 / int32 / int64 (pjamd_data_to_vb) gives bit-identical
    log-likelihoods, equal to the FP64 oracle; a negative int8 entry stays negative (marginalized)
    instead of wrapping to 255."""
    U.say(f"[dtypes] {name}")
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, precision="fp32", **kw)
    pc.to(device)
    X = U.random_data(cards, 24, seed=31)
    dts = [torch.uint8, torch.int8, torch.int32, torch.int64] + ([torch.bool] if max(cards) == 2 else [])
    out = [pc(torch.from_numpy(X).to(dt).to(device)).detach().cpu().numpy()[:, 0] for dt in dts]
    for o in out[1:]:
        assert np.array_equal(o, out[0])
    pc.update_parameters()
    assert np.abs(out[0] - dag_eval.forward(ns, X)[ns][0]).max() <= 1e-4 * len(cards)
    Xm = X.copy()
    Xm[:, 0] = -1
    a = pc(torch.from_numpy(Xm).to(torch.int8).to(device)).detach().cpu().numpy()[:, 0]
    mask = torch.zeros(len(cards), dtype=torch.bool)
    mask[0] = True
    b = pc(torch.from_numpy(X).to(device), missing_mask=mask.to(device)).detach().cpu().numpy()[:, 0]
    assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["hclt_cat_bs32", "hclt_cat_bs64", "hmm_small", "pd_sparse"])
@pytest.mark.parametrize("B", [1, 4])
@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_tiny_batches(name, B, precision, device):
    """Batches of 1 and 4 samples (partial tiles and one 16-byte column group) through forward,
    flow backward and parameter flows, against the FP64 oracle."""
    U.say(f"[tiny batch] {name} B={B} {precision}")
    t_ll, t_mar, t_off, t_flow, t_em = TOL[precision]
    ns, cards, kw = _circuit(name)
    pc = TensorCircuit(ns, precision=precision, **kw)
    pc.to(device)
    X = U.random_data(cards, B, seed=37)
    xt = torch.from_numpy(X).to(device)
    lls = pc(xt).detach().cpu().numpy()[:, 0]
    pc.backward(xt, flows_memory=0.0)
    pc.update_parameters()
    vals = dag_eval.forward(ns, X)
    flows, pflows = dag_eval.backward(ns, X, vals)
    assert np.abs(lls - vals[ns][0]).max() <= t_ll * len(cards)
    pc.update_param_flows(origin_ns_only=False)
    for n in node_iterator(ns):
        if n.is_sum() and n.provided("_param_flows") and not n.is_tied():
            got = U.dense_pf_from_blocks(n, n._param_flows.numpy())
            ref = pflows[n]
            for m in node_iterator(ns):
                if m is not n and m.is_tied() and m.get_source_ns() is n and m._param_flow_range == n._param_flow_range:
                    ref = ref + pflows[m]
            assert U.rel_err(got, ref) <= t_flow, (n, U.rel_err(got, ref))


def _edge_circuit(name):
    """A circuit with a seeded quarter of its sum parameters masked to zero (``zero_param_mask``,
    at least one child kept per node) and a seeded fifth of its Categorical parameters set to 0."""
    ns, cards, kw = _circuit(name)
    TensorCircuit(ns, **kw).update_parameters()     # every vector holds its parameters
    g = torch.Generator().manual_seed(5)
    for n in node_iterator(ns):
        if n.is_tied():
            continue
        if n.is_sum() and n.num_ch_nodes > 1:
            m = torch.rand(n._params.shape, generator=g) < 0.25
            m[..., 0] &= False   # keep child 0 of each child block: no node loses every child
            n.set_zero_param_mask(m)
            n._params[m] = 0.0
        elif n.is_input() and n.dist.get_signature() == "Categorical":
            K = n.dist.num_cats
            P = n._params.reshape(-1, K).clone()
            z = torch.rand(P.shape, generator=g) < 0.2
            z[:, 0] = False
            P[z] = 0.0
            n._params = (P / P.sum(dim=1, keepdim=True)).reshape(-1)
    return ns, cards, kw


@pytest.mark.parametrize("name", ["hclt_cat_bs32", "hclt_cat_bs64", "pd_small", "hmm_small"])
@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp32_valu"])
@pytest.mark.parametrize("B", [64, 300])
def test_zero_params_impossible_data_keep_zero_em(name, precision, B, device):
    """Zero parameters and out-of-range categories (-inf LL columns: the reference's
    ``emars_max == -inf`` guards, ``sum_layer.py:519-549``) through forward, flows and EM with
    ``keep_zero_params`` (``par_update.py:196-205, 236-253``, ``categorical.py:171-201``)."""
    U.say(f"[edge] {name} {precision} B={B}")
    t_ll, t_mar, t_off, t_flow, t_em = TOL[precision]
    ns, cards, kw = _edge_circuit(name)
    pc = TensorCircuit(ns, precision=precision, **kw)
    pc.to(device)
    X = U.random_data(cards, B, seed=13)
    X[3, 0] = cards[0]          # out of range: -inf
    X[B - 2, -1] = cards[-1] + 7
    xt = torch.from_numpy(X).to(device)
    lls = pc(xt).detach().cpu().numpy()[:, 0]
    pc.update_parameters()
    vals = dag_eval.forward(ns, X)
    ref = vals[ns][0]
    inf = np.isinf(ref)
    assert inf[3] and inf[B - 2]
    assert np.array_equal(np.isinf(lls), inf) and np.all(lls[inf] < 0)
    assert np.abs(lls[~inf] - ref[~inf]).max() <= t_ll * len(cards)
    pc.backward(xt, allow_modify_flows=False, flows_memory=0.0)
    flows, pflows = dag_eval.backward(ns, X, vals)
    nf = pc.node_flows.detach().cpu().numpy()
    assert np.isfinite(nf).all()
    for n in node_iterator(ns):
        if n.is_sum() or n.is_input():
            a, b = n._output_ind_range
            assert U.rel_err(nf[a:b], flows[n]) <= t_flow, n
    pc.update_param_flows(origin_ns_only=False)
    for n in node_iterator(ns):
        if n.is_sum() and n.provided("_param_flows") and not n.is_tied():
            got = U.dense_pf_from_blocks(n, n._param_flows.numpy())
            ref_pf = pflows[n]
            for m in node_iterator(ns):
                if m is not n and m.is_tied() and m.get_source_ns() is n and m._param_flow_range == n._param_flow_range:
                    ref_pf = ref_pf + pflows[m]
            assert U.rel_err(got, ref_pf) <= t_flow, n
    step, pcount = 0.5, 0.1
    pc.mini_batch_em(step_size=step, pseudocount=pcount, keep_zero_params=True)
    new = dag_eval.em_update(ns, pflows, step_size=step, pseudocount=pcount, keep_zero_params=True)
    pc.update_parameters()
    for s, p in new.items():
        if s.is_sum():
            got = U.dense_pf_from_blocks(s, s._params.numpy())
        else:
            got = s._params.numpy().astype(np.float64).reshape(p.shape)
        assert np.array_equal(got == 0, p == 0) or s.is_input() and s.dist.get_signature() != "Categorical"
        assert np.abs(got - p).max() <= t_em, (s, np.abs(got - p).max())


@pytest.mark.parametrize("B", [64, 512, 1024, 2048])
def test_mixed_category_counts_on_one_variable(B, device):
    """Two Categorical vectors over the same variable with different category counts, laid out in
    one input layer so that 8-node tiles of one 64-node backward group alternate between them:
    each tile's sort must use its own count."""
    from pyjuice_amd.nodes import inputs, multiply, summate
    import pyjuice_amd as juice
    U.say(f"[mixed K] B={B}")
    torch.manual_seed(2)
    with juice.set_block_size(8):
        vecs = []
        for v in range(3):
            a = inputs(v, num_node_blocks=2, dist=juice.distributions.Categorical(num_cats=12))
            b = inputs(v, num_node_blocks=2, dist=juice.distributions.Categorical(num_cats=5))
            vecs.append(summate(a, b, num_node_blocks=2))
        top = multiply(*vecs)
    root = summate(top, num_node_blocks=1, block_size=1)
    root.init_parameters()
    pc = TensorCircuit(root)
    pc.to(device)
    layer = pc.input_layer_group[0]
    assert layer._uniform_tiles(8)
    k = layer._ncats.cpu().numpy()
    v = layer.vids[:, 0].cpu().numpy()
    tiles = [(int(v[i]), int(k[i])) for i in range(0, len(k), 8)]
    assert any(t[0] == u[0] and t[1] != u[1] for t, u in zip(tiles, tiles[1:])), tiles
    X = U.random_data([12] * 3, B, seed=4)     # categories >= 5 are impossible for the 5-category vectors
    xt = torch.from_numpy(X).to(device)
    pc(xt)
    pc.backward(xt, allow_modify_flows=False, flows_memory=0.0)
    pc.update_parameters()
    vals = dag_eval.forward(root, X)
    _, pflows = dag_eval.backward(root, X, vals)
    pfi = layer.param_flows.detach().cpu().numpy().astype(np.float64)
    for n in layer.nodes:
        a, b = n._param_flow_range
        got = pfi[a:b].reshape(n.num_nodes, n.dist.num_param_flows())
        assert U.rel_err(got, pflows[n]) <= 1e-4, n
