"""Top-down ancestral sampling (``pyjuice/queries/sample.py:335-474``).

All layers run on the device without host round-trips, top-down: each sum layer picks one
child element per active (node, sample), each product layer activates its active elements'
children, and the input layers draw the values.  The sampled sub-circuit is kept as 0/1 activity
masks over ``[num_nodes, S]`` and ``[num_elements, S]``; every kernel clears the flags it
consumes, so the masks are zero again after a pass and are kept on the circuit for the next one
(no clearing launches, and the pass can be recorded as one HIP graph).

Randomness: one uniform per (node, sample).  Either the caller's ``uniforms[node_id, s]`` in
[0, 1), or, by default, the counter-based uniform of (node, s) under a 64-bit ``seed``, drawn
inside the kernels (Philox4x32-10 word 0 as 24-bit fraction; ``oracle.dag_eval.philox_uniforms``
restates it), as the reference draws ``tl.rand(seed, offset)`` in its kernels.  A sum node scans
its children in chunks of 512 (``sample.py:115-131``): inside a chunk the pick is the first
child whose FP32 running sum exceeds ``u``, and a ``u`` past the chunk's total moves on with the
total subtracted; a Categorical input node picks the first category whose running sum reaches
``u`` (``categorical.py:137-161``) and a Bernoulli input node returns ``u < theta``
(``bernoulli.py:79-88``).  Given the same uniforms the samples are bit-identical to the oracle's.
"""
from __future__ import annotations

from typing import Optional

import torch

from pyjuice_amd import _native as N

__all__ = ["sample"]


def _prefix_plan(pc):
    """Row tables of ``pjamd_sample_prefix`` for the sum parameters (one row per sum node, stride =
    block size, from the EM records) and for each Categorical input layer (one row per distinct
    parameter vector), with their output buffers; built once per circuit.  The sum rows are used
    only when every forward table row's parameters are the contiguous run its EM record covers
    (checked here), so that the running sums follow the table's child order."""
    plan = getattr(pc, "_sample_prefix", None)
    if plan is not None and plan["device"] == pc.device:
        return plan
    import numpy as np
    dev = pc.device
    recs = pc.par_update_kwargs[7]["node_blocks"].detach().cpu().numpy().astype(np.int64)
    runs = {int(r[0]): (int(r[3]), int(r[4])) for r in recs}
    ok = True
    for g in pc.inner_layer_groups:
        if not g.is_sum():
            continue
        for layer in g:
            bs = layer.block_size
            for pids, nch in zip(layer.partitioned_pids, layer._fw_n_chs):
                P = pids.detach().cpu().numpy().astype(np.int64)
                n = nch.detach().cpu().numpy().astype(np.int64)
                k = np.arange(P.shape[1])[None, :]
                real = k < n[:, None]
                if not np.all(np.where(real, P == P[:, :1] + bs * k, True)):
                    ok = False
                elif any(runs.get(int(a)) != (bs, int(c)) for a, c in zip(P[:, 0], n)):
                    ok = False
    plan = {"device": dev, "sum": None, "inputs": []}
    if ok and recs.shape[0] > 0:
        bsz, nch = recs[:, 3], recs[:, 4]
        base = np.repeat(recs[:, 0], bsz) + (np.arange(int(bsz.sum())) - np.repeat(np.cumsum(bsz) - bsz, bsz))
        plan["sum"] = dict(base=torch.from_numpy(base).to(dev),
                           stride=torch.from_numpy(np.repeat(bsz, bsz).astype(np.int32)).to(dev),
                           len=torch.from_numpy(np.repeat(nch, bsz).astype(np.int32)).to(dev),
                           out=torch.zeros([pc.num_sum_params], device=dev, dtype=torch.float32))
    for layer in pc.input_layer_group:
        if layer._kind() != 0:
            plan["inputs"].append(None)
            continue
        sp = layer.s_pids.detach().cpu().numpy().astype(np.int64)
        kk = layer._ncats.detach().cpu().numpy().astype(np.int64)
        ub, first = np.unique(sp, return_index=True)
        plan["inputs"].append(dict(base=torch.from_numpy(ub).to(dev),
                                   stride=torch.ones(ub.size, dtype=torch.int32, device=dev),
                                   len=torch.from_numpy(kk[first].astype(np.int32)).to(dev),
                                   out=torch.zeros_like(layer.params)))
    pc._sample_prefix = plan
    return plan


def _buffers(pc, S: int):
    """Activity masks (all zero between passes), the key buffer and the output of ``S``-sample
    passes, kept on the circuit."""
    key = (pc.device, S)
    b = getattr(pc, "_sample_buffers", None)
    if b is None or b["key"] != key:
        dev = pc.device
        dtype = pc.input_layer_group[0].get_data_dtype()
        b = {
            "key": key,
            "node_act": torch.zeros([pc.num_nodes, S], device=dev, dtype=torch.uint8),
            "elem_act": torch.zeros([pc.num_elements, S], device=dev, dtype=torch.uint8),
            "seed": torch.zeros([2], device=dev, dtype=torch.int32),
            "out": torch.zeros([pc.num_vars, S], device=dev,
                               dtype=torch.int64 if dtype != torch.bool else torch.uint8),
            "graph": None,
        }
        pc._sample_buffers = b
    return b


def _run_pass(pc, b, S: int, uniforms: Optional[torch.Tensor]):
    """Kernel launches of one pass (graph-capturable: no allocation, no host synchronisation)."""
    r0, r1 = pc.root_ns._output_ind_range
    node_act, elem_act = b["node_act"], b["elem_act"]
    node_act[r0:r1].fill_(1)
    u_ptr = N.ptr(uniforms) if uniforms is not None else None
    seed_ptr = N.ptr(b["seed"])
    s = N.stream()
    plan = _prefix_plan(pc)
    # running sums of the current parameters (recomputed every pass: the parameters may change)
    sp = plan["sum"]
    if sp is not None:
        N.call("pjamd_sample_prefix", N.ptr(sp["out"]), N.ptr(pc.params), N.ptr(sp["base"]), N.ptr(sp["stride"]),
               N.ptr(sp["len"]), sp["base"].numel(), 512, s, algo=(8 * pc.num_sum_params, 0))
    for layer, ip in zip(pc.input_layer_group, plan["inputs"]):
        if ip is not None:
            N.call("pjamd_sample_prefix", N.ptr(ip["out"]), N.ptr(layer.params), N.ptr(ip["base"]),
                   N.ptr(ip["stride"]), N.ptr(ip["len"]), ip["base"].numel(), 0, s,
                   algo=(8 * layer.params.numel(), 0))
    cdf = N.ptr(sp["out"]) if sp is not None else None
    for lid in range(len(pc.inner_layer_groups) - 1, -1, -1):
        g = pc.inner_layer_groups[lid]
        if g.is_sum():
            for layer in g:
                for nids, cids, pids, pfids, nch, _ in layer._dev["fw"]:
                    N.call("pjamd_sample_sum", N.ptr(elem_act), N.ptr(node_act), u_ptr, seed_ptr, N.ptr(pc.params),
                           cdf, N.ptr(nids), N.ptr(cids), N.ptr(pids), N.ptr(nch), nids.numel(), cids.size(1),
                           layer.block_size, S, s, algo=(2 * nids.numel() * layer.block_size * S, 0))
        else:
            for layer in g:
                for nids, cids, _ in layer._dev["fw"]:
                    N.call("pjamd_sample_prod", N.ptr(node_act), N.ptr(elem_act), N.ptr(nids), N.ptr(cids),
                           nids.numel(), cids.size(1), layer.block_size, S, s,
                           algo=(nids.numel() * layer.block_size * S, 0))
    for layer, ip in zip(pc.input_layer_group, plan["inputs"]):
        layer.sample(b["out"], node_act, uniforms, seed=b["seed"], cdf=ip["out"] if ip is not None else None)


@torch.no_grad()
def sample(pc, num_samples: Optional[int] = None, conditional: bool = False, uniforms: Optional[torch.Tensor] = None,
           seed: Optional[int] = None, record_cudagraph: bool = False, apply_cudagraph: bool = True,
           **kwargs) -> torch.Tensor:
    """Draw ``num_samples`` samples ``[S, num_vars]`` from a single-root circuit.

    ``uniforms`` ``[num_nodes, S]`` float32 fixes every draw; otherwise the draws are the
    counter-based uniforms of ``seed`` (a random one when None).  ``record_cudagraph`` records
    the seeded pass as a HIP graph on first use; later calls with the same ``S`` replay it
    (``apply_cudagraph``) under their own seed."""
    assert not conditional, "conditional sampling is not implemented"
    assert num_samples is not None and num_samples > 0
    r0, r1 = pc.root_ns._output_ind_range
    assert r1 - r0 == 1, "sampling needs a circuit with a single root node"
    dev = pc.device
    assert dev.type == "cuda", "sampling runs on the HIP device; call `pc.to('cuda')` first"
    S = int(num_samples)
    b = _buffers(pc, S)
    if uniforms is not None:
        assert uniforms.shape == (pc.num_nodes, S) and uniforms.dtype == torch.float32
        uniforms = uniforms.to(dev).contiguous()
    else:
        if seed is None:
            seed = int(torch.randint(0, 2 ** 62, (1,)).item())
        seed = int(seed) & ((1 << 64) - 1)
        lo, hi = seed & 0xFFFFFFFF, seed >> 32
        # two uint32 key words as int32 bit patterns
        b["seed"].copy_(torch.tensor([lo - (1 << 32) if lo >= 1 << 31 else lo,
                                      hi - (1 << 32) if hi >= 1 << 31 else hi], dtype=torch.int32))
    try:
        if uniforms is None and (record_cudagraph or (apply_cudagraph and b["graph"] is not None)):
            if b["graph"] is None:
                _run_pass(pc, b, S, None)        # this call's samples; the capture executes nothing
                torch.cuda.synchronize()
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    _run_pass(pc, b, S, None)
                torch.cuda.synchronize()
                b["graph"] = g
            else:
                b["graph"].replay()
        else:
            _run_pass(pc, b, S, uniforms)
    except Exception:
        pc._sample_buffers = None            # the masks may hold flags of the failed pass
        raise
    out = b["out"].permute(1, 0).contiguous()
    dtype = pc.input_layer_group[0].get_data_dtype()
    return out.bool() if dtype == torch.bool else out
