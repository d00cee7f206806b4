"""Top-down ancestral sampling (``pyjuice/queries/sample.py:335-474``).

All layers run on the device without host round-trips.  The sampled sub-circuit is kept as
dense 0/1 activity masks over ``[num_nodes, S]`` and ``[num_elements, S]``.  Randomness comes
from one uniform per (node, sample), ``uniforms[node_id, s]`` in [0, 1): a sum node picks child
``#{k : u >= cumsum_k(theta)}`` (``sample.py:125-131``; running sums in FP32, left to right,
clamped to the last child when ``u`` reaches the total), a Categorical input node picks the
first category whose running sum reaches ``u`` (``categorical.py:137-161``) and a Bernoulli
input node returns ``u < theta`` (``bernoulli.py:79-88``).  Given the same uniforms the samples
are bit-identical to the oracle's.
"""
from __future__ import annotations

from typing import Optional

import torch

from pyjuice_amd import _native as N

__all__ = ["sample"]


@torch.no_grad()
def sample(pc, num_samples: Optional[int] = None, conditional: bool = False, uniforms: Optional[torch.Tensor] = None,
           seed: Optional[int] = None, **kwargs) -> torch.Tensor:
    """Draw ``num_samples`` samples ``[S, num_vars]`` from a single-root circuit."""
    assert not conditional, "conditional sampling is not implemented"
    assert num_samples is not None and num_samples > 0
    r0, r1 = pc.root_ns._output_ind_range
    assert r1 - r0 == 1, "sampling needs a circuit with a single root node"
    dev = pc.device
    assert dev.type == "cuda", "sampling runs on the HIP device; call `pc.to('cuda')` first"
    S = int(num_samples)
    if uniforms is None:
        gen = torch.Generator(device=dev)
        gen.manual_seed(int(seed) if seed is not None else int(torch.randint(0, 2 ** 31 - 1, (1,)).item()))
        uniforms = torch.rand([pc.num_nodes, S], device=dev, generator=gen, dtype=torch.float32)
    assert uniforms.shape == (pc.num_nodes, S) and uniforms.dtype == torch.float32
    uniforms = uniforms.to(dev).contiguous()

    node_act = torch.zeros([pc.num_nodes, S], device=dev, dtype=torch.uint8)
    elem_act = torch.zeros([pc.num_elements, S], device=dev, dtype=torch.uint8)
    node_act[r0:r1, :] = 1
    s = N.stream()
    for lid in range(len(pc.inner_layer_groups) - 1, -1, -1):
        g = pc.inner_layer_groups[lid]
        if g.is_sum():
            elem_act.zero_()
            for layer in g:
                for nids, cids, pids, pfids, nch, _ in layer._dev["fw"]:
                    N.call("pjamd_sample_sum", N.ptr(elem_act), N.ptr(node_act), N.ptr(uniforms), N.ptr(pc.params),
                           N.ptr(nids), N.ptr(cids), N.ptr(pids), N.ptr(nch), nids.numel(), cids.size(1),
                           layer.block_size, S, s)
        else:
            for layer in g:
                for nids, cids, _ in layer._dev["fw"]:
                    N.call("pjamd_sample_prod", N.ptr(node_act), N.ptr(elem_act), N.ptr(nids), N.ptr(cids),
                           nids.numel(), cids.size(1), layer.block_size, S, s)
    dtype = pc.input_layer_group[0].get_data_dtype()
    out = torch.zeros([pc.num_vars, S], device=dev, dtype=torch.int64 if dtype != torch.bool else torch.uint8)
    for layer in pc.input_layer_group:
        layer.sample(out, node_act, uniforms)
    pc._sample_node_active = node_act
    out = out.permute(1, 0).contiguous()
    return out.bool() if dtype == torch.bool else out
