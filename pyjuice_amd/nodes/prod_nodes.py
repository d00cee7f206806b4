"""Vectors of product nodes (``pyjuice/nodes/prod_nodes.py``).

``edge_ids`` is ``[num_node_blocks, num_chs]`` (block-sparse: every entry names a child
node block) or ``[num_nodes, num_chs]`` (sparse: every entry names a child node).
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np
import torch

from pyjuice_amd.graph import PartitionNode
from .nodes import CircuitNodes

__all__ = ["ProdNodes"]


class ProdNodes(CircuitNodes):
    SPARSE = 0
    BLOCK_SPARSE = 1

    def __init__(self, num_node_blocks: int, chs: Sequence[CircuitNodes], edge_ids: Optional[torch.Tensor] = None,
                 block_size: int = 0, **kwargs):
        rg = PartitionNode([c.region_node for c in chs])
        super().__init__(num_node_blocks, rg, block_size=block_size, **kwargs)
        self.chs = list(chs)
        self._construct_edges(edge_ids)
        self._run_init_callbacks(**kwargs)

    def is_prod(self):
        return True

    @property
    def num_edges(self):
        return self.num_nodes * self.num_chs

    @property
    def edge_type(self):
        if self.edge_ids.size(0) == self.num_node_blocks:
            return self.BLOCK_SPARSE
        if self.edge_ids.size(0) == self.num_nodes:
            return self.SPARSE
        raise RuntimeError(f"product edge table must be [num_node_blocks, num_children], got {tuple(self.edge_ids.shape)}")

    def is_block_sparse(self):
        return self.edge_type == self.BLOCK_SPARSE

    def is_sparse(self):
        return self.edge_type == self.SPARSE

    def duplicate(self, *args, tie_params: bool = False, allow_type_mismatch: bool = False):
        chs = list(args) if len(args) > 0 else self.chs
        if len(args) > 0:
            assert len(chs) == self.num_chs
            for o, n in zip(self.chs, chs):
                if not allow_type_mismatch:
                    assert type(o) == type(n)
                else:
                    assert not n.is_prod()
                assert o.num_node_blocks == n.num_node_blocks and o.block_size == n.block_size
        return ProdNodes(self.num_node_blocks, chs, self.edge_ids.clone(), block_size=self.block_size,
                         source_node=self if tie_params else None)

    def init_parameters(self, perturbation: float = 2.0, recursive: bool = True, is_root: bool = True, **kwargs):
        super().init_parameters(perturbation=perturbation, recursive=recursive, is_root=is_root, **kwargs)

    def _construct_edges(self, edge_ids):
        if edge_ids is None:
            for c in self.chs:
                assert self.num_node_blocks == c.num_node_blocks and self.block_size == c.block_size, \
                    "without edge_ids every child must have as many node blocks as the product"
            edge_ids = torch.arange(self.num_node_blocks).unsqueeze(1).repeat(1, self.num_chs)
        if isinstance(edge_ids, np.ndarray):
            edge_ids = torch.from_numpy(edge_ids)
        edge_ids = edge_ids.long()
        if edge_ids.size(0) == self.num_node_blocks:
            assert edge_ids.size(1) == self.num_chs, f"edge_ids must have shape ({self.num_node_blocks}, {self.num_chs}) for a block-sparse product"
            for i, c in enumerate(self.chs):
                assert torch.all(edge_ids[:, i] >= 0), "negative child index in the edge table"
                assert torch.all(edge_ids[:, i] < c.num_node_blocks), "child index in the edge table exceeds the child vector"
        elif edge_ids.size(0) == self.num_nodes:
            assert edge_ids.size(1) == self.num_chs, f"edge_ids must have shape ({self.num_nodes}, {self.num_chs}) for a sparse product"
            for i, c in enumerate(self.chs):
                assert torch.all(edge_ids[:, i] >= 0), "negative child index in the edge table"
                assert torch.all(edge_ids[:, i] < c.num_nodes), "child index in the edge table exceeds the child vector"
        else:
            raise RuntimeError(f"product edge table must be [num_node_blocks, num_children], got {tuple(edge_ids.shape)}")
        self.edge_ids = edge_ids

    def __repr__(self):
        t = "sparse" if self.edge_type == self.SPARSE else "block-sparse"
        return (f"<ProdNodes {self.num_node_blocks}x{self.block_size}, {t}, {self.num_chs} child vector(s), "
                f"{len(self.scope)} var(s)>")
