"""The benchmark configurations at their real shapes, on the HIP path, checked against the FP64
C restatement (``oracle/pc_cpu.c``, itself pinned to the DAG oracle) on the same compiled
circuit and parameters, or against properties that hold at any size.

1. HCLT, 16 Bernoulli variables, 8 latents, 512 samples, FP32: forward / flows / EM against the
   DAG oracle; the probabilities of all 2^16 assignments sum to one.
2. HCLT-784 (256 categories), 128 latents, block size 32, FP32: |dLL| <= 1e-4 nats/var, parameter
   flows and EM against the oracle.
3. PD 28x28, 256 latents, block size 32, BF16 and FP16: |dLL| <= 1e-3 nats/var; the FP64
   parameter-flow accumulator over 60 mini-batches of 1024 equals the FP64 sum of its parts,
   and a mini-batch's accumulated parameter flows match the oracle's.
4. HMM, 1024 hidden states, 4096 tokens, length 64, tied: |dLL| <= 1e-4 nats/var (FP32); the
   autograd adjoint, with layer-wise recomputation of the element marginals, against oracle
   parameter flows and against exact finite differences.
5. HCLT-784, 512 latents, block size 32, 50 % of the sum-edge blocks pruned, batch 2048, BF16:
   |dLL| <= 1e-3 nats/var and parameter flows against the oracle.
Oracle comparisons use a few samples of the batch (the FP64 CPU pass is the slow part).
"""
import itertools

import numpy as np
import pytest
import torch

import parity_utils as U
import pyjuice_amd as juice
from oracle import dag_eval
from oracle.pc_cpu import CompiledCPU
from pyjuice_amd import synthetic
from pyjuice_amd.model import TensorCircuit
from pyjuice_amd.nodes.nodes import node_iterator

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def device():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def _images(n, seed=0):
    return synthetic.image_categorical(n, H=28, W=28, K=256, rho=0.6, seed=seed, field_seed=1)


def _hclt784(device, L, block_size=None, n_fit=2048):
    x = _images(n_fit)
    torch.manual_seed(0)
    ns = juice.structures.HCLT(torch.from_numpy(x).float().to(device), num_latents=L,
                               input_dist=juice.distributions.Categorical(256), chunk_size=64,
                               block_size=block_size)
    return ns


def _sum_pflows_close(pc, cpu, tol):
    got = pc.param_flows.detach().cpu().numpy().astype(np.float64)
    return U.rel_err(got, cpu.param_flows)


def _input_pflows_close(pc, cpu):
    worst = 0.0
    for layer, d in zip(pc.input_layer_group, cpu.inputs):
        got = layer.param_flows.detach().cpu().numpy().astype(np.float64)
        worst = max(worst, U.rel_err(got, d["pflows"][:got.size]))
    return worst


def test_config1_hclt_bernoulli16(device):
    U.say("[config 1] HCLT 16 Bernoulli vars, 8 latents")
    x = synthetic.bernoulli_chain(512, 16, seed=0).astype(np.int64)
    torch.manual_seed(0)
    ns = juice.structures.HCLT(torch.from_numpy(x).float(), num_latents=8,
                               input_dist=juice.distributions.Bernoulli(), chunk_size=16)
    pc = juice.compile(ns)
    pc.to(device)
    xt = torch.from_numpy(x).to(device)
    lls = pc(xt).detach().cpu().numpy()[:, 0]
    pc.update_parameters()
    vals = dag_eval.forward(ns, x)
    assert np.abs(lls - vals[ns][0]).max() <= 1e-4 * 16
    # all 2^16 assignments: the circuit is normalised
    allx = torch.from_numpy(np.array(list(itertools.product([0, 1], repeat=16)), dtype=np.int64)).to(device)
    with torch.no_grad():
        ll_all = pc(allx).double()[:, 0]
    assert abs(float(torch.logsumexp(ll_all, 0))) <= 1e-4
    # one full EM step (pseudocount 0.1) against the oracle
    pc(xt)
    pc.backward(xt, flows_memory=0.0)
    flows, pflows = dag_eval.backward(ns, x, vals)
    pc.mini_batch_em(step_size=1.0, pseudocount=0.1)
    new = dag_eval.em_update(ns, pflows, step_size=1.0, pseudocount=0.1)
    pc.update_parameters()
    for s, p in new.items():
        got = U.dense_pf_from_blocks(s, s._params.numpy()) if s.is_sum() else \
            s._params.numpy().astype(np.float64).reshape(p.shape)
        assert np.abs(got - p).max() <= 2e-6, s


def test_config2_hclt784_L128_bs32_fp32(device):
    U.say("[config 2] HCLT-784, 128 latents, block size 32, FP32")
    ns = _hclt784(device, 128, block_size=32)
    pc = juice.compile(ns)
    pc.to(device)
    x = _images(32, seed=5).astype(np.int64)
    xt = torch.from_numpy(x).to(device)
    cpu = CompiledCPU(pc, np.float64)
    ref = cpu.forward(x)[0]
    lls = pc(xt, precision="fp32").detach().cpu().numpy()[:, 0]
    assert np.abs(lls - ref).max() / 784 <= 1e-4
    pc.backward(xt, flows_memory=0.0, precision="fp32")
    cpu.backward()
    assert _sum_pflows_close(pc, cpu, 1e-4) <= 1e-4
    # FP32 flows pass through 53 layers: relative error grows with depth
    assert _input_pflows_close(pc, cpu) <= 1e-3
    # EM from the same flows: the oracle's EM step runs on the HIP parameter flows, so this pins
    # the EM kernels themselves (the flows were checked above)
    cpu.param_flows[:] = pc.param_flows.detach().cpu().numpy().astype(np.float64)
    for layer, d in zip(pc.input_layer_group, cpu.inputs):
        f = layer.param_flows.detach().cpu().numpy().astype(np.float64)
        d["pflows"][:f.size] = f
    pc.mini_batch_em(step_size=0.1, pseudocount=0.1)
    cpu.mini_batch_em(step_size=0.1, pseudocount=0.1)
    assert np.abs(pc.params.detach().cpu().numpy() - cpu.params).max() <= 2e-6
    for layer, d in zip(pc.input_layer_group, cpu.inputs):
        assert np.abs(layer.params.detach().cpu().numpy() - d["params"][:layer.params.numel()]).max() <= 2e-6


@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_config3_pd_28x28_L256(precision, device):
    U.say(f"[config 3] PD 28x28, 256 latents, {precision}")
    torch.manual_seed(0)
    ns = juice.structures.PD(data_shape=(28, 28), num_latents=256, split_intervals=(4, 4),
                             structure_type="sum_dominated", input_dist=juice.distributions.Categorical(256))
    pc = juice.compile(ns)
    pc.to(device)
    n_batches, bsz = 60, 1024
    x = _images(n_batches * bsz, seed=7).astype(np.int64)
    xt = torch.from_numpy(x).to(device)
    lls = pc(xt[:bsz], precision=precision).detach().cpu().numpy()[:, 0]
    cpu = CompiledCPU(pc, np.float64)
    n = 4
    ref = cpu.forward(x[:n])[0]
    assert np.abs(lls[:n] - ref).max() / 784 <= 1e-3
    # FP64 accumulator over 60 mini-batches of 1024 (61,440 samples) = the FP64 sum of the
    # single-batch flows (each batch's flows stored, then added on the device in the same order)
    want = None
    for i in range(n_batches):
        xb = xt[i * bsz:(i + 1) * bsz]
        pc(xb, precision=precision)
        pc.backward(xb, flows_memory=0.0, precision=precision)
        want = pc.param_flows.clone() if want is None else want + pc.param_flows
    pc.init_param_flows(0.0)
    for i in range(n_batches):
        xb = xt[i * bsz:(i + 1) * bsz]
        pc(xb, precision=precision)
        pc.backward(xb, flows_memory=1.0, precision=precision)
    assert pc.param_flows.dtype == torch.float64
    assert U.rel_err(pc.param_flows.cpu().numpy(), want.cpu().numpy()) <= 1e-12
    # and the accumulated flows of the first mini-batch's samples match the FP64 oracle's
    pc.init_param_flows(0.0)
    pc(xt[:n], precision=precision)
    pc.backward(xt[:n], flows_memory=1.0, precision=precision)
    cpu.backward()
    assert _sum_pflows_close(pc, cpu, 2e-2) <= 2e-2


def test_config4_hmm_adjoint(device):
    U.say("[config 4] HMM 1024 x 4096, length 64")
    torch.manual_seed(0)
    ns = juice.structures.HMM(seq_length=64, num_latents=1024, num_emits=4096, homogeneous=True)
    pc = juice.compile(ns)
    pc.to(device)
    seqs = synthetic.hmm_sequences(256, 64, 1024, 4096, alpha=0.1, seed=2).astype(np.int64)
    xt = torch.from_numpy(seqs).to(device)
    lls = pc(xt, precision="fp32").detach().cpu().numpy()[:, 0]
    cpu = CompiledCPU(pc, np.float64)
    n = 4
    ref = cpu.forward(seqs[:n])[0]
    assert np.abs(lls[:n] - ref).max() / 64 <= 1e-4

    # adjoint behind torch.autograd.Function against the oracle's parameter flows (4 sequences),
    # with layer-wise recomputation of the element marginals (one shared element buffer)
    pc.recompute_elements = True
    params = pc.params.detach().clone().requires_grad_(True)
    in_params = [l.params.detach().clone().requires_grad_(True) for l in pc.input_layer_group]
    out = juice.log_likelihood(pc, xt[:n], params=params, input_params=in_params, precision="fp32")
    out.sum().backward()
    cpu.backward()
    theta = pc.params.detach().cpu().numpy().astype(np.float64)
    fused = cpu.param_flows.copy()
    for t, k, ch in cpu.fuse:
        fused[t:t + k] = sum(cpu.param_flows[c:c + k] for c in ch)
    idx = juice.model.autograd._sum_param_to_flow(pc).cpu().numpy()
    ok = (idx >= 0) & (theta > 0)
    ref_grad = np.zeros_like(theta)
    ref_grad[ok] = fused[idx[ok]] / theta[ok]
    g = params.grad.detach().cpu().numpy().astype(np.float64)
    assert pc.element_mars.size(0) == pc.num_elements < pc.num_element_rows
    assert U.rel_err(g[ok] * theta[ok], ref_grad[ok] * theta[ok]) <= 1e-4

    # exact finite differences: the circuit is multilinear, so scaling one parameter theta_k by
    # (1 + d) changes a sample's log-likelihood by log(1 + d * theta_k * dLL/dtheta_k)
    x1 = xt[:1]
    params1 = pc.params.detach().clone().requires_grad_(True)
    base = juice.log_likelihood(pc, x1, params=params1, precision="fp32")
    base.sum().backward()
    g1 = params1.grad.detach().to(torch.float64)
    th = params1.detach().to(torch.float64)
    contrib = (g1 * th).cpu().numpy()
    rng = np.random.default_rng(3)
    cand = np.argsort(-contrib)[:256]
    pick = rng.choice(cand, size=64, replace=False)
    d = 1.0
    with torch.no_grad():
        ll0 = float(pc(x1, precision="fp32")[0, 0])
        worst = 0.0
        for k in pick:
            saved = pc.params[k].item()
            pc.params[k] = saved * (1.0 + d)
            ll1 = float(pc(x1, precision="fp32")[0, 0])
            pc.params[k] = saved
            pred = np.log1p(d * contrib[k])
            worst = max(worst, abs((ll1 - ll0) - pred) / (1e-3 + abs(pred)))
    assert contrib[pick].min() > 0 and worst <= 2e-2, worst


def test_config5_pruned_hclt512_bs32_b2048(device):
    U.say("[config 5] HCLT-784, 512 latents, block size 32, 50% blocks pruned, batch 2048, bf16")
    ns = _hclt784(device, 512, block_size=32)
    removed = juice.transformations.prune_sum_blocks(ns, 0.5, seed=4)
    assert removed > 0
    pc = juice.compile(ns)
    pc.to(device)
    x = _images(2048, seed=9).astype(np.int64)
    xt = torch.from_numpy(x).to(device)
    lls = pc(xt, precision="bf16").detach().cpu().numpy()[:, 0]
    assert np.isfinite(lls).all()
    n = 4
    cpu = CompiledCPU(pc, np.float64)
    ref = cpu.forward(x[:n])[0]
    assert np.abs(lls[:n] - ref).max() / 784 <= 1e-3
    pc(xt[:n], precision="bf16")
    pc.backward(xt[:n], flows_memory=0.0, precision="bf16")
    cpu.backward()
    assert _sum_pflows_close(pc, cpu, 2e-2) <= 2e-2
    assert _input_pflows_close(pc, cpu) <= 2e-2
