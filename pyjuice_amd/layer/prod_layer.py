"""Product layer (``pyjuice/layer/prod_layer.py``): compile in ``__init__`` and the
gather-add forward / scatter-free backward through the HIP library."""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np
import torch
import torch.nn as nn

from pyjuice_amd import _native as N
from .compilation import (assign_partitions, flatten_c_ids, get_prod_layer_parstats, get_prod_layer_stats,
                          prod_layer_backward_compilation, prod_layer_forward_compilation)
from .layer import Layer
from .partition import next_power_of_2, partition_nodes_by_n_edges

__all__ = ["ProdLayer"]


def _unique_pow2(x: torch.Tensor) -> torch.Tensor:
    return torch.unique(torch.from_numpy(next_power_of_2(x)))


class ProdLayer(Layer, nn.Module):
    def __init__(self, nodes: Sequence, global_nid_start: Optional[int] = None,
                 layer_sparsity_tol: Optional[float] = None, max_num_partitions: Optional[int] = None,
                 disable_gpu_compilation: bool = False, force_gpu_compilation: bool = False):
        Layer.__init__(self, nodes)
        nn.Module.__init__(self)
        assert len(self.nodes) > 0, "a product layer needs at least one node vector"
        assert len(self.nodes) == len(set(self.nodes)), "a vector appears twice in one layer"

        # prod_layer.py:39-46: block-sparse only if every vector is block-sparse with one block size.
        bse = all((not ns.is_sparse()) and ns.block_size == self.nodes[0].block_size for ns in self.nodes)
        self.use_block_sparse_edges = bse
        self.block_size = self.nodes[0].block_size if bse else 1
        if global_nid_start is None:
            global_nid_start = self.block_size

        nblk, nedges, n_chgs = get_prod_layer_stats(self.nodes, self.block_size, global_nid_start, bse)
        self.num_nodes = nblk * self.block_size
        self.num_edges = nedges
        self._layer_nid_range = (global_nid_start, global_nid_start + self.num_nodes)

        fw_max = _unique_pow2(partition_nodes_by_n_edges(n_chgs, sparsity_tolerance=layer_sparsity_tol,
                                                         max_num_partitions=max_num_partitions))
        self.num_fw_partitions = len(fw_max)
        part = torch.zeros([len(self.nodes)], dtype=torch.long)
        loc = torch.zeros([len(self.nodes)], dtype=torch.long)
        cnt = torch.zeros([self.num_fw_partitions], dtype=torch.long)
        for i, ns in enumerate(self.nodes):
            p = int((ns.num_chs > fw_max).sum().item())
            part[i] = p
            loc[i] = cnt[p]
            cnt[p] += ns.num_node_blocks if bse else ns.num_nodes
        nids, cids = prod_layer_forward_compilation(self.nodes, fw_max, part, loc, cnt, self.block_size, bse)
        self.partitioned_nids = nids
        self.partitioned_cids = cids

        flat_cids, flat_cid2nid = flatten_c_ids(nids, cids)
        u_cids, par_counts = get_prod_layer_parstats(flat_cids, global_nid_start=global_nid_start)
        bk_max = _unique_pow2(partition_nodes_by_n_edges(par_counts, sparsity_tolerance=layer_sparsity_tol,
                                                         max_num_partitions=max_num_partitions))
        self.num_bk_partitions = len(bk_max)
        bpart, bloc, bcnt = assign_partitions(par_counts, bk_max, start_min=0)
        u, parids = prod_layer_backward_compilation(u_cids, flat_cids, flat_cid2nid, bk_max, bpart, bloc, bcnt)
        self.partitioned_u_cids = u
        self.partitioned_parids = parids
        self._dev = None

    def is_prod(self):
        return True

    def __repr__(self):
        return f"ProdLayer[{self._layer_nid_range[0]}:{self._layer_nid_range[1]}, {self.num_nodes} nodes, {self.num_edges} edges]"

    def to(self, device):
        nn.Module.to(self, device)
        device = torch.device(device)
        self.partitioned_nids = [t.to(device) for t in self.partitioned_nids]
        self.partitioned_cids = [t.to(device) for t in self.partitioned_cids]
        self.partitioned_u_cids = [t.to(device) for t in self.partitioned_u_cids]
        self.partitioned_parids = [t.to(device) for t in self.partitioned_parids]
        self.device = device
        self._dev = None
        if device.type == "cuda":
            self.build_device_tables()
        return self

    def build_device_tables(self):
        """int32 launch tables (``_dev``) on the layer's device, with the algorithmic FP32 words per
        sample of each partition: rows gathered + rows written (+ read for accumulation)."""
        i32 = lambda t: t.to(torch.int32).contiguous()
        bs = self.block_size
        self._dev = {
            "fw": [(i32(n), i32(c), int((c != 0).sum()) * bs + n.numel() * bs)
                   for n, c in zip(self.partitioned_nids, self.partitioned_cids) if n.numel() > 0],
            "bk": [(i32(n), i32(c), int((c != 0).sum()) * bs + 2 * n.numel() * bs)
                   for n, c in zip(self.partitioned_u_cids, self.partitioned_parids) if n.numel() > 0],
        }

    def forward(self, node_mars: torch.Tensor, element_mars: torch.Tensor, _for_backward: bool = False, **kwargs):
        """``element_mars[nid + j] = sum_k node_mars[cid_k + j]`` (``prod_layer.py:160-204``)."""
        B = node_mars.size(1)
        for nids, cids, words in self._dev["fw"]:
            N.call("pjamd_prod_forward", N.ptr(element_mars), N.ptr(node_mars), N.ptr(nids), N.ptr(cids),
                   nids.numel(), cids.size(1), self.block_size, B, 0, N.stream(), algo=(4 * words * B, 0))

    def backward(self, node_flows: torch.Tensor, element_flows: torch.Tensor, logspace_flows: bool = False, **kwargs):
        """``node_flows[u + j] += sum_p element_flows[par_p + j]`` (``prod_layer.py:206-237``)."""
        assert not logspace_flows, "logspace flows are not supported by the HIP product layers"
        B = node_flows.size(1)
        for u, parids, words in self._dev["bk"]:
            N.call("pjamd_prod_backward", N.ptr(node_flows), N.ptr(element_flows), N.ptr(u), N.ptr(parids),
                   u.numel(), parids.size(1), self.block_size, B, N.stream(), algo=(4 * words * B, 0))
