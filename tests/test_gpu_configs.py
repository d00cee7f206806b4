"""The benchmark configurations at their real shapes, on the HIP path, checked against the FP64
C restatement (``oracle/pc_cpu.c``, itself pinned to the DAG oracle) on the same compiled
circuit and parameters, or against properties that hold at any size.

1. HCLT, 16 Bernoulli variables, 8 latents, 512 samples, FP32: forward / flows / EM against the
   DAG oracle; the probabilities of all 2^16 assignments sum to one.
2. HCLT-784 (256 categories), 128 latents, block size 32, FP32: |dLL| <= 1e-4 nats/var, parameter
   flows and EM against the oracle.
3. PD 28x28, 256 latents, block size 32, batch 1024, BF16 and FP16: |dLL| <= 1e-3 nats/var;
   parameter flows and EM against the oracle; the FP64 parameter-flow accumulator over 60
   mini-batches of 1024 equals the FP64 sum of its parts.
4. HMM, 1024 hidden states, 4096 tokens, length 64, tied: |dLL| <= 1e-4 nats/var (FP32); the
   autograd adjoint, with layer-wise recomputation of the element marginals, against oracle
   parameter flows and against FP32 central differences on 80 random parameters.
5. HCLT-784, 512 latents, block size 32, 50 % of the sum-edge blocks pruned, batch 2048, BF16:
   |dLL| <= 1e-3 nats/var on every sample and parameter flows against the oracle.
Headline: HCLT-784, 256 latents, block size 256, batch 512, BF16: every LL column, full-batch
parameter flows and the EM step against the oracle.
LL is compared on every column of the batch.  Flows are compared on the whole batch (configs 2
and headline) or, where the FP64 pass over the whole batch is too slow (configs 3 and 5), with root
weight 1 on two columns of every 64-column tile and 0 elsewhere, so the device still runs the
full-size launches.
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


def _spread_columns(B, per_tile=2, tile=64, seed=0):
    """Sorted column indices with ``per_tile`` columns in every ``tile``-wide batch tile (the first
    and last column always included)."""
    g = np.random.default_rng(seed)
    cols = {0, B - 1}
    for t0 in range(0, B, tile):
        n = min(tile, B - t0)
        cols.update(int(c) for c in t0 + g.choice(n, size=min(per_tile, n), replace=False))
    return np.array(sorted(cols), dtype=np.int64)


def _ll_err_per_var(lls, ref, V):
    d = np.abs(lls - ref) / V
    return float(d.mean()), float(d.max())


def _copy_flows_to_cpu(pc, cpu):
    cpu.param_flows[:] = pc.param_flows.detach().cpu().numpy().astype(np.float64)
    for layer, d in zip(pc.input_layer_group, cpu.inputs):
        f = layer.param_flows.detach().cpu().numpy().astype(np.float64)
        d["pflows"][:f.size] = f


def _em_matches(pc, cpu, step, pcount, tol):
    """The EM kernels against the oracle's EM, both from the HIP parameter flows."""
    _copy_flows_to_cpu(pc, cpu)
    pc.mini_batch_em(step_size=step, pseudocount=pcount)
    cpu.mini_batch_em(step_size=step, pseudocount=pcount)
    worst = float(np.abs(pc.params.detach().cpu().numpy() - cpu.params).max())
    for layer, d in zip(pc.input_layer_group, cpu.inputs):
        worst = max(worst, float(np.abs(layer.params.detach().cpu().numpy() - d["params"][:layer.params.numel()]).max()))
    assert worst <= tol, worst


def _full_batch_parity(pc, cpu, x, precision, t_ll, t_flow, t_in):
    """Every LL column, then parameter flows (sum and input) of unit root flows over the whole
    batch, against the FP64 oracle on the same compiled circuit."""
    xt = torch.from_numpy(x).to(pc.device)
    B, V = x.shape
    lls = pc(xt, precision=precision).detach().cpu().numpy()[:, 0]
    ref = cpu.forward(x)[0]
    mean, mx = _ll_err_per_var(lls, ref, V)
    U.say(f"  |dLL|/var over {B} samples: mean {mean:.2e}, max {mx:.2e}")
    assert mx <= t_ll, mx
    pc.backward(xt, flows_memory=0.0, precision=precision)
    cpu.backward()
    e_sum, e_in = _sum_pflows_close(pc, cpu, t_flow), _input_pflows_close(pc, cpu)
    U.say(f"  parameter flows rel. err: sum {e_sum:.2e}, input {e_in:.2e}")
    assert e_sum <= t_flow and e_in <= t_in, (e_sum, e_in)


def _weighted_columns_parity(pc, x, cols, precision, t_flow, t_in):
    """Flows of the whole batch on the device with root weight 1 on ``cols`` and 0 elsewhere (so
    every batch tile and reduction stage of the full-size launch runs) against the oracle's flows
    of just those columns."""
    xt = torch.from_numpy(x).to(pc.device)
    B = x.shape[0]
    w = torch.zeros([1, B], device=pc.device)
    w[0, torch.from_numpy(cols).to(pc.device)] = 1.0
    pc(xt, precision=precision)
    pc.backward(xt, ll_weights=w, flows_memory=0.0, precision=precision)
    cpu = CompiledCPU(pc, np.float64)
    cpu.forward(x[cols])
    cpu.backward()
    e_sum, e_in = _sum_pflows_close(pc, cpu, t_flow), _input_pflows_close(pc, cpu)
    U.say(f"  weighted-column flows ({cols.size} of {B} columns) rel. err: sum {e_sum:.2e}, input {e_in:.2e}")
    assert e_sum <= t_flow and e_in <= t_in, (e_sum, e_in)
    return cpu


def test_config2_hclt784_L128_bs32_fp32(device):
    U.say("[config 2] HCLT-784, 128 latents, block size 32, batch 512, FP32")
    ns = _hclt784(device, 128, block_size=32)
    pc = juice.compile(ns)
    pc.to(device)
    x = _images(512, seed=5).astype(np.int64)
    cpu = CompiledCPU(pc, np.float64)
    # FP32 flows pass through 53 layers: relative error of the input flows grows with depth
    _full_batch_parity(pc, cpu, x, "fp32", 1e-4, 1e-4, 1e-3)
    _em_matches(pc, cpu, 0.1, 0.1, 2e-6)


def test_headline_hclt784_L256_bs256_bf16(device):
    U.say("[headline] HCLT-784, 256 latents, block size 256, batch 512, BF16")
    ns = _hclt784(device, 256)
    pc = juice.compile(ns)
    pc.to(device)
    assert all(l.block_size in (1, 256) for g in pc.inner_layer_groups if g.is_sum() for l in g)
    x = _images(512, seed=9).astype(np.int64)
    cpu = CompiledCPU(pc, np.float64)
    _full_batch_parity(pc, cpu, x, "bf16", 1e-3, 2e-2, 2e-2)
    _em_matches(pc, cpu, 0.1, 0.1, 2e-6)


def test_headline_training_step_as_benchmarked(device):
    """The headline step as bench.py runs it (chain runs on, the input layers' EM inside the
    backward, TrainStepGraph.run_step) against the FP64 oracle's forward, flow backward and EM on
    the same batch and starting parameters: every LL column, the stored sum and input parameter
    flows, and the updated parameters."""
    U.say("[headline step] HCLT-784, 256 latents, block size 256, batch 512, BF16, chains, fused input EM")
    ns = _hclt784(device, 256)
    pc = juice.compile(ns)
    pc.to(device)
    pc.use_chains = True
    x = _images(512, seed=13).astype(np.int64)
    assert pc._chain_plan("bf16", 512) is not None, "no chain run planned"
    assert all(l.fused_em_ok(512, l._output_ind_range[0]) for l in pc.input_layer_group)
    cpu = CompiledCPU(pc, np.float64)          # the starting parameters
    xt = torch.from_numpy(x).to(device)
    step, pcount = 0.1, 0.1
    lls = juice.optim.TrainStepGraph.run_step(pc, xt, step, pcount, "bf16").detach().cpu().numpy()[:, 0]
    torch.cuda.synchronize()
    assert all(not l.__dict__.get("_em_done", False) for l in pc.input_layer_group)   # consumed by the EM
    ref = cpu.forward(x)[0]
    mean, mx = _ll_err_per_var(lls, ref, x.shape[1])
    U.say(f"  |dLL|/var over 512 samples: mean {mean:.2e}, max {mx:.2e}")
    assert mx <= 1e-3, mx
    cpu.backward()
    e_sum, e_in = _sum_pflows_close(pc, cpu, 2e-2), _input_pflows_close(pc, cpu)
    U.say(f"  parameter flows rel. err: sum {e_sum:.2e}, input {e_in:.2e}")
    assert e_sum <= 2e-2 and e_in <= 2e-2, (e_sum, e_in)
    cpu.mini_batch_em(step_size=step, pseudocount=pcount)
    worst = float(np.abs(pc.params.detach().cpu().numpy() - cpu.params).max())
    for layer, d in zip(pc.input_layer_group, cpu.inputs):
        worst = max(worst, float(np.abs(layer.params.detach().cpu().numpy() - d["params"][:layer.params.numel()]).max()))
    U.say(f"  EM parameters max abs err: {worst:.2e}")
    assert worst <= 1e-3, worst


@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_config3_pd_28x28_L256(precision, device):
    U.say(f"[config 3] PD 28x28, 256 latents, batch 1024, {precision}")
    torch.manual_seed(0)
    ns = juice.structures.PD(data_shape=(28, 28), num_latents=256, split_intervals=(4, 4),
                             structure_type="sum_dominated", input_dist=juice.distributions.Categorical(256))
    pc = juice.compile(ns)
    pc.to(device)
    n_batches, bsz = 60, 1024
    x = _images(n_batches * bsz, seed=7).astype(np.int64)
    xt = torch.from_numpy(x).to(device)
    # every LL column of the first mini-batch
    lls = pc(xt[:bsz], precision=precision).detach().cpu().numpy()[:, 0]
    cpu = CompiledCPU(pc, np.float64)
    ref = cpu.forward(x[:bsz])[0]
    mean, mx = _ll_err_per_var(lls, ref, 784)
    U.say(f"  |dLL|/var over {bsz} samples: mean {mean:.2e}, max {mx:.2e}")
    assert mx <= 1e-3, mx
    # the full-size backward launches (batch 1024) against the oracle's flows of two columns per
    # 64-column tile
    cpu = _weighted_columns_parity(pc, x[:bsz], _spread_columns(bsz), precision, 2e-2, 2e-2)
    _em_matches(pc, cpu, 0.1, 0.1, 2e-6)
    # FP64 accumulator over 60 mini-batches of 1024 (61,440 samples): equals the host's FP64
    # sum of the per-batch flows (summed in another order), and its first batch's unit-weight
    # flows are the full-batch flows checked above
    parts = []
    for i in range(n_batches):
        xb = xt[i * bsz:(i + 1) * bsz]
        pc(xb, precision=precision)
        pc.backward(xb, flows_memory=0.0, precision=precision)
        parts.append(pc.param_flows.detach().cpu().numpy().copy())
    want = np.sum(np.stack(parts[::-1]), axis=0)
    pc.init_param_flows(0.0)
    for i in range(n_batches):
        xb = xt[i * bsz:(i + 1) * bsz]
        pc(xb, precision=precision)
        pc.backward(xb, flows_memory=1.0, precision=precision)
    assert pc.param_flows.dtype == torch.float64
    assert U.rel_err(pc.param_flows.cpu().numpy(), want) <= 1e-12


def test_config4_hmm_adjoint(device):
    U.say("[config 4] HMM 1024 x 4096, length 64, batch 256")
    torch.manual_seed(0)
    ns = juice.structures.HMM(seq_length=64, num_latents=1024, num_emits=4096, homogeneous=True)
    pc = juice.compile(ns)
    pc.to(device)
    seqs = synthetic.hmm_sequences(256, 64, 1024, 4096, alpha=0.1, seed=2).astype(np.int64)
    xt = torch.from_numpy(seqs).to(device)
    lls = pc(xt, precision="fp32").detach().cpu().numpy()[:, 0]
    cpu = CompiledCPU(pc, np.float64)
    n = 16
    ref = cpu.forward(seqs[:n])[0]
    dll = np.abs(lls[:n] - ref)
    assert dll.max() / 64 <= 1e-4

    # adjoint behind torch.autograd.Function against the oracle's parameter flows (16 sequences),
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
    g = params.grad.detach().cpu().numpy().astype(np.float64)
    assert pc.element_mars.size(0) == pc.num_elements < pc.num_element_rows
    assert U.rel_err(g[ok] * theta[ok], fused[idx[ok]]) <= 1e-4

    # FP32 central differences of sum_b LL_b over all 256 sequences, on 64 uniformly drawn sum
    # parameters (almost all of them transition parameters, each tied across the 63 steps) and
    # 16 emission parameters (uniform over the tokens that occur, tied across the 64 positions).
    # A parameter is scaled by exp(+-h); for one occurrence, LL is log(a + b theta), whose
    # log-space central quotient (LL(h) - LL(-h)) / (2 sinh h) equals theta dLL/dtheta up to a
    # relative O(p) with p = b theta / (a + b theta) the sample's flow through the edge (< 1e-2
    # here, so 2 % covers it); the FP32 rounding of the two evaluations adds
    # 2 sqrt(B) eps_LL / (2 sinh h) with eps_LL the per-sample FP32 error measured above.
    B, h = seqs.shape[0], 1.0
    params1 = pc.params.detach().clone().requires_grad_(True)
    in1 = [l.params.detach().clone().requires_grad_(True) for l in pc.input_layer_group]
    juice.log_likelihood(pc, xt, params=params1, input_params=in1, precision="fp32").sum().backward()
    pred_sum = (params1.grad.detach().double() * params1.detach().double()).cpu().numpy()
    layer = pc.input_layer_group[0]
    pred_in = (in1[0].grad.detach().double() * in1[0].detach().double()).cpu().numpy()
    eps_ll = max(float(dll.max()), 1e-5)
    noise = 2.0 * np.sqrt(B) * 2.0 * eps_ll / (2.0 * np.sinh(h))
    rng = np.random.default_rng(3)
    pick_sum = rng.choice(np.nonzero(ok)[0], size=64, replace=False)
    src = layer.source_nids.cpu().numpy()
    s_pids = layer.s_pids.cpu().numpy()[src]
    seen = np.unique(seqs)
    cand_in = (s_pids[:, None] + seen[None, :]).reshape(-1)
    pick_in = rng.choice(cand_in, size=16, replace=False)

    def total_ll():
        return float(pc(xt, precision="fp32").double().sum())

    worst = 0.0
    with torch.no_grad():
        for buf, picks, pred in ((pc.params, pick_sum, pred_sum), (layer.params, pick_in, pred_in)):
            for k in picks:
                saved = buf[k].item()
                buf[k] = saved * np.exp(h)
                lp = total_ll()
                buf[k] = saved * np.exp(-h)
                lm = total_ll()
                buf[k] = saved
                fd = (lp - lm) / (2.0 * np.sinh(h))
                tol = noise + 2e-2 * abs(pred[k])
                worst = max(worst, abs(fd - pred[k]) / tol)
    U.say(f"  finite differences: worst |fd - pred| / tol = {worst:.3f} (noise floor {noise:.2e})")
    assert worst <= 1.0, worst


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
    cpu = CompiledCPU(pc, np.float64)
    ref = cpu.forward(x)[0]
    mean, mx = _ll_err_per_var(lls, ref, 784)
    U.say(f"  |dLL|/var over 2048 samples: mean {mean:.2e}, max {mx:.2e}")
    assert mx <= 1e-3, mx
    # batch 2048 runs the 2-node x 2048-column Categorical backward; two columns per 64-column
    # tile carry root weight
    _weighted_columns_parity(pc, x, _spread_columns(2048), "bf16", 2e-2, 2e-2)
    # and at batch 1100 (the 8-node x 1024-column form's range ends at 1024)
    _weighted_columns_parity(pc, x[:1100], _spread_columns(1100, seed=1), "bf16", 2e-2, 2e-2)


def test_headline_sampling_bit_exact(device):
    """Ancestral sampling at the headline size (HCLT-784, 256 latents, 512 samples), in-kernel
    counter-based uniforms, the pass recorded as a HIP graph and replayed under a second seed:
    every sample bit-identical to the oracle drawing the same uniforms."""
    from oracle.dag_eval import philox_uniforms
    U.say("[headline sampling] HCLT-784, 256 latents, 512 samples")
    ns = _hclt784(device, 256)
    pc = juice.compile(ns)
    pc.to(device)
    pc.update_parameters()
    S = 512
    for seed, rec in ((123, True), (456, False)):
        got = juice.queries.sample(pc, num_samples=S, seed=seed, record_cudagraph=rec).cpu().numpy().astype(np.int64)
        ref = dag_eval.sample(ns, (lambda r, c, sd=seed: philox_uniforms(sd, r, c), S), pc.num_vars)
        assert np.array_equal(got, ref), (seed, int((got != ref).sum()))
