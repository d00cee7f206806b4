"""DAG construction API: ``inputs``, ``multiply``, ``summate`` and the context managers
``set_block_size`` / ``structural_properties`` (``pyjuice/nodes/construction.py``)."""
from __future__ import annotations

from typing import Callable, Optional, Sequence, Union

import numpy as np
import torch

from pyjuice_amd.graph import RegionGraph
from pyjuice_amd.utils import DecoratorContextManager
from .distributions import Distribution
from .input_nodes import InputNodes
from .nodes import CircuitNodes
from .prod_nodes import ProdNodes
from .sum_nodes import SumNodes

__all__ = ["inputs", "multiply", "summate", "set_block_size", "structural_properties"]


def inputs(var: Union[int, Sequence[int]], num_node_blocks: int = 0, dist: Distribution = None,
           params: Optional[torch.Tensor] = None, num_nodes: int = 0, block_size: int = 0, **kwargs) -> InputNodes:
    """Vector of input nodes over the same variable(s) with one distribution (construction.py:22)."""
    assert block_size == 0 or block_size & (block_size - 1) == 0, "block_size must be a power of two"
    if num_nodes > 0:
        assert num_node_blocks == 0, "give num_nodes or num_node_blocks, not both"
        if block_size == 0:
            block_size = CircuitNodes.DEFAULT_BLOCK_SIZE
        assert num_nodes % block_size == 0
        num_node_blocks = num_nodes // block_size
    return InputNodes(num_node_blocks=num_node_blocks, scope=[var] if isinstance(var, int) else list(var),
                      dist=dist, params=params, block_size=block_size, **kwargs)


def multiply(nodes1, *args, edge_ids: Optional[torch.Tensor] = None, sparse_edges: bool = False, **kwargs) -> ProdNodes:
    """Vector of product nodes over children with disjoint scopes (construction.py:69)."""
    assert nodes1.is_sum() or nodes1.is_input(), "multiply() takes input or sum vectors only"
    chs = [nodes1]
    num_node_blocks = nodes1.num_node_blocks
    block_size = nodes1.block_size
    scope = nodes1.scope
    for ns in args:
        assert ns.is_input() or ns.is_sum(), \
            f"Children of product nodes must be input or sum nodes, but found input of type {type(ns)}."
        if edge_ids is None:
            assert ns.num_node_blocks == num_node_blocks, \
                f"Input nodes should have the same `num_node_blocks`, but got {ns.num_node_blocks} and {num_node_blocks}."
        assert ns.block_size == block_size, "Input nodes should have the same `block_size`."
        if not RegionGraph.ALLOW_NONDECOMPOSABLE:
            assert not scope.contains_any(ns.scope), "the children of multiply() must cover disjoint variables"
        chs.append(ns)
        scope = scope | ns.scope
    if edge_ids is not None:
        if isinstance(edge_ids, np.ndarray):
            edge_ids = torch.from_numpy(edge_ids)
        if sparse_edges:
            assert edge_ids.shape[0] % block_size == 0
            num_node_blocks = edge_ids.shape[0] // block_size
        else:
            num_node_blocks = edge_ids.shape[0]
    return ProdNodes(num_node_blocks, chs, edge_ids, block_size=block_size, **kwargs)


def summate(nodes1, *args, num_node_blocks: int = 0, num_nodes: int = 0, edge_ids: Optional[torch.Tensor] = None,
            block_size: int = 0, sum_edge_ids_constructor: Optional[Callable] = None, **kwargs) -> SumNodes:
    """Vector of sum nodes over children with the same scope (construction.py:122)."""
    assert block_size == 0 or block_size & (block_size - 1) == 0, "block_size must be a power of two"
    if num_nodes > 0:
        assert num_node_blocks == 0, "give num_nodes or num_node_blocks, not both"
        if block_size == 0:
            block_size = CircuitNodes.DEFAULT_BLOCK_SIZE
        assert num_nodes % block_size == 0
        num_node_blocks = num_nodes // block_size
    assert nodes1.is_prod() or nodes1.is_input(), \
        f"summate() takes input or product vectors, not {type(nodes1).__name__}"
    if sum_edge_ids_constructor is not None:
        edge_ids = sum_edge_ids_constructor(nodes1, *args, num_node_blocks=num_node_blocks,
                                            block_size=block_size if block_size > 0 else CircuitNodes.DEFAULT_BLOCK_SIZE,
                                            **kwargs)
    if edge_ids is not None and num_node_blocks == 0:
        num_node_blocks = int(torch.as_tensor(edge_ids)[0, :].max().item()) + 1
    assert num_node_blocks > 0, "a vector needs at least one node block"
    chs = [nodes1]
    for ns in args:
        assert ns.is_prod() or ns.is_input(), \
            f"Children of sum nodes must be input or product nodes, but found input of type {type(ns)}."
        if not RegionGraph.ALLOW_NONSMOOTH:
            assert ns.scope == nodes1.scope, "the children of summate() must cover the same variables"
        chs.append(ns)
    return SumNodes(num_node_blocks, chs, edge_ids, block_size=block_size, **kwargs)


class set_block_size(DecoratorContextManager):
    """Sets ``CircuitNodes.DEFAULT_BLOCK_SIZE`` inside the ``with`` block (construction.py:193)."""

    def __init__(self, block_size: int = 1):
        assert block_size & (block_size - 1) == 0, "block_size must be a power of two"
        self.block_size = block_size
        self.original_block_size = None

    def __enter__(self):
        self.original_block_size = CircuitNodes.DEFAULT_BLOCK_SIZE
        CircuitNodes.DEFAULT_BLOCK_SIZE = self.block_size

    def __exit__(self, *exc):
        CircuitNodes.DEFAULT_BLOCK_SIZE = self.original_block_size


class structural_properties(DecoratorContextManager):
    """Relaxes the smoothness / decomposability assertions inside the ``with`` block (construction.py:223)."""

    def __init__(self, allow_nonsmooth: bool = False, allow_nondecomposable: bool = False):
        self.allow_nonsmooth = allow_nonsmooth
        self.allow_nondecomposable = allow_nondecomposable

    def __enter__(self):
        RegionGraph.ALLOW_NONSMOOTH = self.allow_nonsmooth
        RegionGraph.ALLOW_NONDECOMPOSABLE = self.allow_nondecomposable

    def __exit__(self, *exc):
        RegionGraph.ALLOW_NONSMOOTH = False
        RegionGraph.ALLOW_NONDECOMPOSABLE = False
