"""Top-down probabilities, the partition function and step-size-rescaled mini-batch EM
(``pyjuice/model/backend/top_down_prob.py:290-402``, ``model/backend/eval_partition.py:200-230``,
``model/tensorcircuit.py:566-578``).

The top-down probability of a node is the flow it receives from a root flow of 1 when every
input variable is summed out.  In a circuit with normalized input distributions every marginal
of that query is a partition function, so the flow backward over one all-marginalized sample is
exactly the reference's top-down pass: a child's probability is sum over parents of
theta * parent probability * Z_child / Z_parent (the ``pc_is_normalized = False`` form; Z = 1
when the sum parameters are normalized too), its parameter flows are theta * probability, and
an input node spreads its probability over its categories by theta (the masked input flows).
The pass runs on the circuit's own HIP kernels over a separate 4-column buffer set (every
column the same sample, the root flow split over the columns), so the training buffers, and any
graph recorded on them, are left alone.
"""
from __future__ import annotations

from typing import Optional

import torch

__all__ = ["eval_top_down_probs", "eval_partition_fn"]

_COLS = 4     # batch columns of the scratch pass (a multiple of 4 takes the 16-byte kernels)


class _ScratchBuffers:
    """Swaps the circuit's node / element buffers for a cached ``_COLS``-column set and back."""

    def __init__(self, pc):
        self.pc = pc

    def __enter__(self):
        pc = self.pc
        self.saved = (pc.node_mars, pc.element_mars, pc.node_flows, pc.element_flows, pc._data_vb,
                      pc._data_marginalized, pc._cum_flow)
        bufs = getattr(pc, "_tdp_buffers", None)
        rows = pc.num_elements if pc.recompute_elements else pc.num_element_rows
        if bufs is None or bufs[1].size(0) != rows:
            pc.node_mars = pc.element_mars = pc.node_flows = pc.element_flows = None
            pc._ensure_buffers(_COLS, need_flows=True)
            bufs = (pc.node_mars, pc.element_mars, pc.node_flows, pc.element_flows)
            pc._tdp_buffers = bufs
        pc.node_mars, pc.element_mars, pc.node_flows, pc.element_flows = bufs
        return self

    def __exit__(self, *exc):
        pc = self.pc
        (pc.node_mars, pc.element_mars, pc.node_flows, pc.element_flows, pc._data_vb,
         pc._data_marginalized, pc._cum_flow) = self.saved
        return False


def _all_missing_forward(pc, precision: Optional[str]):
    x = torch.zeros([_COLS, pc.num_vars], dtype=torch.int64, device=pc.device)
    mask = torch.ones([pc.num_vars], dtype=torch.bool, device=pc.device)
    return pc.forward(x, missing_mask=mask, precision=precision)


def eval_partition_fn(pc, precision: Optional[str] = "fp32") -> torch.Tensor:
    """log Z of every root node ``[num_roots]``: the forward pass with every input variable
    summed out (``eval_partition.py:200-230``; the input distributions are taken as normalized,
    which the EM updates keep them)."""
    with torch.no_grad(), _ScratchBuffers(pc):
        lls = _all_missing_forward(pc, precision)
        return lls[0].detach().clone()


def eval_top_down_probs(pc, update_pflow: bool = True, scale: float = 1.0, pc_is_normalized: bool = True,
                        use_cudagraph: bool = True, precision: Optional[str] = "fp32") -> None:
    """Top-down probability of every node, left in ``pc._single_node_flows`` ``[num_nodes]``;
    with ``update_pflow`` also ``param_flows += scale * theta * probability`` for the sum edges and
    the input categories (``top_down_prob.py:290-402``).  ``pc_is_normalized`` is accepted for
    the reference's signature: the pass divides by the partition functions either way."""
    with torch.no_grad(), _ScratchBuffers(pc):
        _all_missing_forward(pc, precision)
        w = torch.full([pc.num_root_nodes, _COLS], (scale if update_pflow else 1.0) / _COLS, dtype=torch.float32,
                       device=pc.device)
        pc.backward(inputs=None, ll_weights=w, compute_param_flows=update_pflow and scale != 0.0,
                    flows_memory=1.0, allow_modify_flows=False, precision=precision)
        norm = _COLS / (scale if update_pflow and scale != 0.0 else 1.0)
        pc._single_node_flows = (pc.node_flows[:, 0] * norm).detach().clone()
    return None
