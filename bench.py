#!/usr/bin/env python
"""Training-step throughput of a compiled HCLT on one MI355X.

Workload (the headline configuration): HCLT over 784 categorical variables (28 x 28 grid,
256 categories), ``num_latents = 256`` (block size 256), batch 512 of seeded synthetic
uint8 images (Gaussian-copula, AR(1) rho = 0.6).  One step is the training loop of the
reference's examples: ``zero_grad`` (clear parameter flows), ``lls = pc(x)``,
``lls.mean().backward()`` (flow backward incl. parameter flows) and ``step`` (mini-batch EM,
step 0.1, pseudocount 0.1).  Inputs and parameters are resident on the device.  Each step is
recorded once as a HIP graph and replayed (``--no-graph``: launched kernel by kernel from the
host); the per-kernel HIP-event timings behind ``roofline`` and ``kernels_ms_per_step`` come from
the same kernels launched eagerly.

Prints one JSON line.  ``value`` is samples/s of the precision given by ``--precision``
(default BF16 on the matrix cores).  The plain FP32 HIP form (vector ALUs, ``fp32_valu``) and
FP32 operands on the matrix cores (``fp32``) are timed beside it on the same circuit and
starting parameters, with the |dLL| per variable of every form against the FP64 oracle.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp32": 157.3, "fp32_valu": 157.3, "bf16": 2500.0, "fp16": 2500.0}
# HBM bytes per launch (MB) measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on the headline workload
# at block sizes 256 (chain runs on) and 64, scaled by the counters' reading of a 1 GiB copy; commands
# and tables in
# profiles/hclt784_L256_B512_bf16_hbm_traffic.txt and profiles/hclt784_L256_bs64_B512_bf16_hbm_traffic.txt.
# Reported as roofline.traffic for the dominant kernel (counters cannot be read from inside the timed
# process).
TRAFFIC_MB_PER_LAUNCH = {
    256: {"pjamd_sum_backward_ele_to_children": 133.36, "pjamd_sum_forward": 59.32,
          "pjamd_sum_backward_par_batched": 1008.62, "pjamd_prod_forward": 53.75, "pjamd_input_backward": 835.17,
          "pjamd_input_forward": 636.21, "pjamd_sum_em": 569.85, "pjamd_input_em": 847.86,
          "pjamd_chain_forward": 190.0, "pjamd_chain_backward": 261.34, "pjamd_chain_pack": 145.43},
    64: {"pjamd_sum_backward_ele_to_children": 35.21, "pjamd_sum_forward": 16.56,
         "pjamd_sum_backward_par_batched": 1197.8, "pjamd_prod_forward": 16.12, "pjamd_input_backward": 835.14,
         "pjamd_input_forward": 636.21, "pjamd_sum_em": 569.95, "pjamd_input_em": 847.86},
}
HBM_KERNELS = ("pjamd_prod_forward", "pjamd_prod_backward", "pjamd_input_forward", "pjamd_input_backward",
               "pjamd_input_em", "pjamd_sum_em", "pjamd_sum_modify_flows", "pjamd_pflow_fuse")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--structure", default="hclt", choices=["hclt", "pd", "hmm"],
                    help="hclt: HCLT-784 (the default workload); pd: PD over the same 28x28 images "
                         "(split intervals 4 x 4, sum-dominated); hmm: homogeneous HMM over seeded token "
                         "sequences (--latents hidden states, --vocab tokens, --seq-length)")
    ap.add_argument("--latents", type=int, default=256)
    ap.add_argument("--vocab", type=int, default=4096, help="HMM vocabulary")
    ap.add_argument("--seq-length", type=int, default=64, help="HMM sequence length")
    ap.add_argument("--block-size", type=int, default=-1,
                    help="HCLT block size: -1 = measured choice among num_latents / 2^k >= 64 (pyjuice_amd.autotune), "
                         "0 = the builder's default (num_latents)")
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--prune", type=float, default=0.0,
                    help="fraction of sum-edge blocks removed at seeded random (seed 4) after the structure is "
                         "built, e.g. 0.5 with --latents 512 --block-size 32 --batch 2048")
    ap.add_argument("--chains", default="auto", choices=["auto", "on", "off"],
                    help="runs of narrow levels in one launch per pass (pjamd_chain_forward / _backward): "
                         "auto = measured choice (with the block size for HCLT), on, off")
    ap.add_argument("--fit", type=int, default=2048, help="samples used to learn the Chow-Liu tree")
    ap.add_argument("--pflow-dtype", default="fp64", choices=["fp32", "fp64"],
                    help="parameter-flow accumulators (sum and input layers): fp64 (the library's default, for "
                         "accumulating over many mini-batches) or fp32 as PyJuice keeps them")
    ap.add_argument("--precision", default="auto", choices=["auto", "fp32", "fp32_valu", "bf16", "fp16"])
    ap.add_argument("--no-graph", dest="graph", action="store_false",
                    help="launch each step's kernels from the host instead of replaying the step's HIP graph")
    ap.add_argument("--no-tune-launches", dest="tune_launches", action="store_false",
                    help="keep the library's default launch form of every sum layer instead of the measured "
                         "choice (pyjuice_amd.autotune.tune_launches)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-samples", type=int, default=128,
                    help="samples of the bench batch in the CPU leg (baseline step and FP64 accuracy check)")
    ap.add_argument("--only", default="", help="profiling aid: run only this precision, no extra legs")
    ap.add_argument("--sample-count", type=int, default=512,
                    help="samples per pass of the sampling leg (0: no sampling leg)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="profiling aid: run one device-to-device copy of this many MiB first, the known-size "
                         "reference for hardware byte counters")
    return ap.parse_args()


def build_circuit(args, dev):
    import torch

    import pyjuice_amd as juice
    from pyjuice_amd import synthetic

    t0 = time.time()
    if args.structure != "hclt":
        args.fit = 0      # only the HCLT learns its tree from data
    if args.structure == "hmm":
        data = synthetic.hmm_sequences(args.fit + args.batch, args.seq_length, args.latents, args.vocab, alpha=0.1,
                                       seed=2)
    else:
        data = synthetic.image_categorical(args.fit + args.batch, H=28, W=28, K=256, rho=0.6, seed=0, field_seed=1)
    log(f"[bench] data {data.shape} in {time.time() - t0:.1f}s")
    xb = torch.from_numpy(data[args.fit:args.fit + args.batch]).to(dev)
    x_fit = torch.from_numpy(data[:args.fit]).float().to(dev)
    if args.structure != "hclt":
        torch.manual_seed(0)
        if args.structure == "pd":
            ns = juice.structures.PD(data_shape=(28, 28), num_latents=args.latents, split_intervals=(4, 4),
                                     structure_type="sum_dominated", input_dist=juice.distributions.Categorical(256))
        else:
            ns = juice.structures.HMM(seq_length=args.seq_length, num_latents=args.latents, num_emits=args.vocab,
                                      homogeneous=True)
        args.block_size, args.block_size_ms = 0, None
        t0 = time.time()
        pc = juice.compile(ns, verbose=False, pflow_dtype=pflow_dtype(args))
        pc.to(dev)
        log(f"[bench] compiled {args.structure}: {pc.num_nodes} nodes, {pc.num_edges} edges in {time.time() - t0:.1f}s")
        choose_chains(pc, xb, args)
        return ns, pc, xb, data[args.fit:args.fit + args.batch]

    def make_nodes(bs):
        torch.manual_seed(0)
        ns = juice.structures.HCLT(x_fit, num_latents=args.latents, input_dist=juice.distributions.Categorical(256),
                                   chunk_size=64, block_size=bs)
        if args.prune > 0:
            juice.transformations.prune_sum_blocks(ns, args.prune, seed=4)
        return ns

    args.block_size_ms = None
    tuned_chains = None
    if args.block_size < 0:
        from pyjuice_amd import autotune
        t0 = time.time()
        cands = autotune.block_size_candidates(args.latents)
        prec = "bf16" if args.precision == "auto" else args.precision
        opts = {"auto": (False, True), "on": (True,), "off": (False,)}[args.chains]
        (best, tuned_chains), ms = autotune.tune_build(make_nodes, xb, cands, opts, steps=3, warmup=1,
                                                       precision=None if prec == "auto" else prec,
                                                       compile_kwargs={"pflow_dtype": pflow_dtype(args)})
        args.block_size = best
        args.block_size_ms = {f"{k[0]}{'+chains' if k[1] else ''}": round(v, 4) for k, v in ms.items()}
        log(f"[bench] block size {best}, chains {'on' if tuned_chains else 'off'}, chosen from "
            f"{args.block_size_ms} ms/step in {time.time() - t0:.1f}s")
    t0 = time.time()
    ns = make_nodes(args.block_size or None)
    log(f"[bench] HCLT structure in {time.time() - t0:.1f}s")
    t0 = time.time()
    pc = juice.compile(ns, verbose=False, pflow_dtype=pflow_dtype(args))
    pc.to(dev)
    log(f"[bench] compiled: {pc.num_nodes} nodes, {pc.num_edges} edges, {pc.num_sum_params} sum params, "
        f"{len(pc.inner_layer_groups)} layer groups in {time.time() - t0:.1f}s")
    if tuned_chains is not None:
        pc.use_chains = tuned_chains
        args.chains_ms = None
    else:
        choose_chains(pc, xb, args)
    return ns, pc, xb, data[args.fit:args.fit + args.batch]


def pflow_dtype(args):
    import torch
    return {"fp32": torch.float32, "fp64": torch.float64}[args.pflow_dtype]


def choose_chains(pc, xb, args):
    """``pc.use_chains`` from ``--chains`` (auto: measured on recorded steps, kept when faster)."""
    args.chains_ms = None
    if args.chains != "auto":
        pc.use_chains = args.chains == "on"
        return
    from pyjuice_amd import autotune
    prec = "bf16" if args.precision == "auto" else args.precision
    if prec not in ("bf16", "fp16"):
        pc.use_chains = False
        return
    t0 = time.time()
    ms = autotune.tune_chains(pc, xb, prec, steps=3, warmup=1)
    args.chains_ms = {("on" if k else "off"): round(v, 4) for k, v in ms.items()}
    log(f"[bench] chains {'on' if pc.use_chains else 'off'} chosen from {args.chains_ms} ms/step in "
        f"{time.time() - t0:.1f}s")


def make_step(pc, xb, prec, graph):
    import pyjuice_amd as juice

    if graph:
        # the whole step (clear flows, forward, flow backward, EM) as one HIP graph
        g = juice.optim.TrainStepGraph(pc, xb, lr=0.1, pseudocount=0.1, precision=prec)
        return lambda: g()

    opt = juice.optim.CircuitOptimizer(pc, lr=0.1, pseudocount=0.1)

    def step():
        opt.zero_grad()
        lls = pc(xb, precision=prec)
        lls.mean().backward()
        opt.step()
        return lls

    return step


def time_steps(step, steps, warmup):
    import torch

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    b.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1000.0 / steps
    return a.elapsed_time(b) / steps, wall


def kernel_breakdown(step, steps):
    """Per-entry-point HIP-event timing of eager steps.  Before each step the stream is held by a
    bounded spin kernel (~20 ms), so the host has enqueued the whole step before the device starts
    it: the launches then run back to back and each event pair brackets its kernel, not the host's
    submission gap."""
    import torch
    from pyjuice_amd import _native as N

    sleep = getattr(torch.cuda, "_sleep", None)
    with N.LaunchTimer() as T:
        for _ in range(steps):
            if sleep is not None:
                sleep(50_000_000)
            step()
    s = T.summary()
    for d in s.values():
        d["ms"] /= steps
        d["bytes"] /= steps
        d["flops"] /= steps
        d["launches"] //= steps
    return s


def chain_report(pc, args):
    """Whether chain runs were on, their measured alternative and, when on, the runs planned at
    the bench batch (levels, components)."""
    out = {"on": bool(pc.use_chains), "chains_ms_per_step": getattr(args, "chains_ms", None)}
    plan = pc._chain_plan("bf16", args.batch) if pc.use_chains else None
    if plan is not None:
        out["runs"] = [{"levels": (r.hi - r.lo) // 2, "components": r.C} for r in plan.runs]
    return out


def roofline_of(name, d, prec, args):
    """Roofline of one kernel from its algorithmic bytes and flops per step: the bound is the
    resource whose time at peak is larger (the sum-layer contractions at B = 512 and block size
    256 carry ~50 flop/byte, below the BF16 ridge point of ~310, so they are HBM-bound too)."""
    peak_f = PEAK_TFLOPS[prec]
    headline = (args.structure == "hclt" and args.latents == 256 and args.batch == 512
                and args.block_size in TRAFFIC_MB_PER_LAUNCH and prec == "bf16" and args.prune == 0)
    traffic = TRAFFIC_MB_PER_LAUNCH[args.block_size].get(name) if headline else None
    traffic = None if traffic is None else round(traffic * 1e6)   # HBM bytes per launch
    t_hbm = d["bytes"] / (HBM_PEAK_GBS * 1e9)
    t_mfma = d["flops"] / (peak_f * 1e12)
    if name in HBM_KERNELS or t_hbm >= t_mfma:
        ach = d["bytes"] / (d["ms"] * 1e-3) / 1e9
        return {"kernel": name, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes_per_launch": round(d["bytes"] / max(d["launches"], 1))}
    ach = d["flops"] / (d["ms"] * 1e-3) / 1e12
    return {"kernel": name, "bound": "mfma", "achieved": round(ach, 3), "peak": peak_f, "unit": "TFLOP/s",
            "frac": round(ach / peak_f, 4), "traffic": traffic,
            "algorithmic_bytes_per_launch": round(d["bytes"] / max(d["launches"], 1))}


def sampling_leg(pc, args):
    """Ancestral sampling throughput: ``--sample-count`` samples per pass with in-kernel
    counter-based uniforms, the pass recorded once as a HIP graph and replayed under a new seed
    each time (``--no-graph``: launched eagerly); plus the eager per-kernel breakdown."""
    import torch

    import pyjuice_amd as juice
    from pyjuice_amd import _native as N

    S = args.sample_count
    graph = bool(args.graph)
    juice.queries.sample(pc, num_samples=S, seed=1, record_cudagraph=graph)
    for i in range(args.warmup):
        juice.queries.sample(pc, num_samples=S, seed=2 + i, apply_cudagraph=graph)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for i in range(args.steps):
        juice.queries.sample(pc, num_samples=S, seed=100 + i, apply_cudagraph=graph)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / args.steps
    with N.LaunchTimer(["pjamd_sample_sum", "pjamd_sample_prod", "pjamd_input_sample"]) as T:
        for i in range(3):
            juice.queries.sample(pc, num_samples=S, seed=200 + i, apply_cudagraph=False)
    kern = T.summary()
    log(f"[bench] sampling: {ms:.3f} ms per {S}-sample pass ({S / (ms * 1e-3):.0f} samples/s)")
    return {"samples_per_s": round(S / (ms * 1e-3), 1), "ms_per_pass": round(ms, 4), "samples_per_pass": S,
            "hip_graph": graph, "uniforms": "in-kernel Philox4x32-10, one per (node, sample)",
            "kernels_ms_per_pass": {k: {"ms": round(v["ms"] / 3, 4), "launches": v["launches"] // 3}
                                    for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["ms"])}}


def _host_threads():
    n = os.environ.get("OMP_NUM_THREADS")
    if n and n.isdigit():
        return int(n)
    return len(os.sched_getaffinity(0))


def cpu_baseline(pc, x_np, lls_by_prec, args):
    """CPU leg, run on the box's host cores in this process:

    * baseline: the FP32 C / OpenMP restatement of the same step (forward, flow backward, EM,
      on the same compiled index tables and parameters), timed on a bounded sample of the
      batch (``--cpu-samples`` samples per step, steps repeated until ~10 s of work);
    * accuracy: the FP64 build of the same restatement on the first ``--cpu-samples``
      samples, against which every device precision's log-likelihoods, parameter flows (sum and
      input) and EM step are compared.
    """
    from oracle.pc_cpu import CompiledCPU

    n = min(args.cpu_samples, x_np.shape[0])
    xs = x_np[:n]
    ref = CompiledCPU(pc, np.float64).forward(xs)[0]
    V = x_np.shape[1]
    acc = dict()
    for prec, lls in lls_by_prec.items():
        d = np.abs(lls[:n] - ref) / V
        acc[prec] = {"mean": float(d.mean()), "max": float(d.max()), "samples": int(n)}

    # parameter flows (unit root flows over the same samples) and one EM step (0.1, 0.1) of every
    # precision against the FP64 restatement, from the same starting parameters
    import torch
    o64 = CompiledCPU(pc, np.float64)
    o64.forward(xs)
    o64.backward()
    pf_ref = o64.param_flows.copy()
    in_ref = [d["pflows"].copy() for d in o64.inputs]
    o64.mini_batch_em(step_size=0.1, pseudocount=0.1)
    p0 = pc.params.detach().clone()
    in0 = [l.params.detach().clone() for l in pc.input_layer_group]
    xd = torch.from_numpy(xs).to(pc.device)

    def rel(a, b):
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))

    for prec in lls_by_prec:
        pc.params.data.copy_(p0)
        for l, p in zip(pc.input_layer_group, in0):
            l.params.data.copy_(p)
        with torch.no_grad():
            pc(xd, precision=prec)
            pc.backward(xd, flows_memory=0.0, precision=prec)
        acc[prec]["param_flows_rel_err"] = rel(pc.param_flows.detach().cpu().numpy().astype(np.float64), pf_ref)
        acc[prec]["input_param_flows_rel_err"] = max(
            rel(l.param_flows.detach().cpu().numpy().astype(np.float64), r[:l.param_flows.numel()])
            for l, r in zip(pc.input_layer_group, in_ref))
        pc.mini_batch_em(step_size=0.1, pseudocount=0.1)
        em = float(np.abs(pc.params.detach().cpu().numpy() - o64.params).max())
        for l, d in zip(pc.input_layer_group, o64.inputs):
            em = max(em, float(np.abs(l.params.detach().cpu().numpy() - d["params"][:l.params.numel()]).max()))
        acc[prec]["em_params_max_abs_err"] = em
    pc.params.data.copy_(p0)
    for l, p in zip(pc.input_layer_group, in0):
        l.params.data.copy_(p)

    cpu = CompiledCPU(pc, np.float32)
    steps, t_tot = 0, 0.0
    while steps < 1 or (t_tot < 10.0 and steps < 200):
        t0 = time.perf_counter()
        cpu.zero_param_flows()
        cpu.forward(xs)
        cpu.backward()
        cpu.mini_batch_em(step_size=0.1, pseudocount=0.1)
        t_tot += time.perf_counter() - t0
        steps += 1
    return {"value": round(n * steps / t_tot, 2), "unit": "samples/s", "cores": _host_threads(), "kind": "port",
            "sample": f"{steps} step(s) of {n} samples of the bench batch: FP32 C/OpenMP forward + flow backward "
                      f"+ EM on the same compiled circuit ({t_tot:.1f} s)"}, acc


def workload_name(args):
    step = "forward + flow backward + mini-batch EM"
    if args.structure == "pd":
        return (f"PD over 28x28 images (256 categories, split intervals 4x4, sum-dominated), "
                f"num_latents={args.latents}, batch={args.batch}; {step}")
    if args.structure == "hmm":
        return (f"homogeneous HMM, {args.latents} hidden states, {args.vocab} tokens, length {args.seq_length}, "
                f"batch={args.batch}; {step}")
    return (f"HCLT-784 (28x28, 256 categories), num_latents={args.latents}, block_size={args.block_size or 'default'}, "
            + (f"{args.prune:.0%} of the sum-edge blocks pruned, " if args.prune > 0 else "")
            + f"batch={args.batch}; {step}")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the data-parallel design for "
              f"several GPUs (data-parallel over the batch, parameter flows reduce-scattered) is not run "
              f"here.", file=sys.stderr)
        sys.exit(2)
    import torch

    import pyjuice_amd as juice
    from pyjuice_amd.layer.sum_layer import PRECISIONS

    assert torch.cuda.is_available(), "bench.py needs the HIP device"
    dev = torch.device("cuda:0")
    if args.calibrate_copy > 0:
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {args.calibrate_copy} MiB ({n} floats) done")
        del src, dst
    ns, pc, xb, x_np = build_circuit(args, dev)

    # fp32_valu: the plain FP32 HIP form (vector ALUs); fp32: FP32 operands on the matrix cores
    precs = ["fp32_valu", "fp32", "bf16"]
    main_prec = "bf16" if args.precision == "auto" else args.precision
    if args.only:
        precs = [args.only]
        main_prec = args.only
    results = dict()
    lls_by_prec = dict()
    state = {k: v.detach().clone() for k, v in [("params", pc.params)]}
    in_state = [l.params.detach().clone() for l in pc.input_layer_group]
    for prec in precs:
        # same starting parameters for every precision
        pc.params.data.copy_(state["params"])
        for l, p in zip(pc.input_layer_group, in_state):
            l.params.data.copy_(p)
        try:
            # log-likelihoods at the starting parameters, before any step (recording the step
            # graph runs one eager step)
            with torch.no_grad():
                lls_by_prec[prec] = pc(xb, precision=prec).detach().cpu().numpy()[:, 0]
            forms = None
            if args.tune_launches:
                from pyjuice_amd import autotune
                t0 = time.time()
                forms = autotune.tune_launches(pc, xb, prec)
                pc.params.data.copy_(state["params"])
                for l, p in zip(pc.input_layer_group, in_state):
                    l.params.data.copy_(p)
                log(f"[bench] {prec}: launch forms {forms['fw']} / {forms['bk']} in {time.time() - t0:.1f}s")
            step = make_step(pc, xb, prec, args.graph)
            ms, wall = time_steps(step, args.steps, args.warmup)
        except Exception as e:  # an unsupported precision is reported, not fatal
            log(f"[bench] {prec}: {e}")
            continue
        results[prec] = {"ms_per_step": ms, "wall_ms_per_step": wall, "samples_per_s": args.batch / (ms * 1e-3),
                         "launch_forms": forms}
        log(f"[bench] {prec}: {ms:.3f} ms/step ({args.batch / (ms * 1e-3):.0f} samples/s)")
        # per-launch HIP-event timing needs eager launches (a graph replay has no host calls)
        # eager steps of the graph's own sequence (TrainStepGraph.run_step), else the timed step
        eager = (lambda: juice.optim.TrainStepGraph.run_step(pc, xb, 0.1, 0.1, prec)) if args.graph else step
        results[prec]["kernels"] = kernel_breakdown(eager,
                                                    min(args.steps, 5))
    if main_prec not in results:
        main_prec = "fp32"
    r = results[main_prec]
    kern = r["kernels"]
    dom = max(kern.items(), key=lambda kv: kv[1]["ms"])
    roof = roofline_of(dom[0], dom[1], main_prec, args)

    sampling = None
    if args.sample_count > 0:
        pc.params.data.copy_(state["params"])
        for l, p in zip(pc.input_layer_group, in_state):
            l.params.data.copy_(p)
        sampling = sampling_leg(pc, args)

    cpu, acc = (None, None)
    if not args.no_cpu_baseline and not args.only:
        pc.params.data.copy_(state["params"])
        for l, p in zip(pc.input_layer_group, in_state):
            l.params.data.copy_(p)
        cpu, acc = cpu_baseline(pc, x_np, lls_by_prec, args)

    out = {
        "metric": "fwd+bwd+EM samples/s, BF16-MFMA vs FP32 HIP at |dLL/var|<=1e-3",
        "value": round(r["samples_per_s"], 1),
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "fp32_valu": "f32", "bf16": "bf16", "fp16": "f16"}[main_prec],
        "data": "synthetic",
        "config": {"workload": workload_name(args),
                   "num_nodes": pc.num_nodes, "num_edges": pc.num_edges, "num_sum_params": pc.num_sum_params,
                   "precision": main_prec, "param_flow_accumulator": args.pflow_dtype, "hip_graph": bool(args.graph),
                   "chain_runs": chain_report(pc, args),
                   "block_size_ms_per_step": args.block_size_ms},
        "roofline": roof,
        "cpu_baseline": cpu,
        "precisions": {p: {"samples_per_s": round(v["samples_per_s"], 1), "ms_per_step": round(v["ms_per_step"], 4)}
                       for p, v in results.items()},
        "launch_forms": r.get("launch_forms"),
        "dll_per_var_vs_fp64": acc,
        "sampling": sampling,
        "kernels_ms_per_step": {k: {"ms": round(v["ms"], 4), "launches": v["launches"],
                                    "GB/s": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1),
                                    "TFLOP/s": round(v["flops"] / max(v["ms"], 1e-9) / 1e9, 2),
                                    "largest_launch_GB/s": round(v["big"][0] / max(v["big"][1], 1e-9) / 1e6, 1),
                                    "largest_launch_MB": round(v["big"][0] / 1e6, 2)}
                                for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["ms"])},
    }
    for base in ("fp32_valu", "fp32"):
        if base in results and "bf16" in results:
            out[f"bf16_over_{base}"] = round(results["bf16"]["samples_per_s"] / results[base]["samples_per_s"], 3)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
