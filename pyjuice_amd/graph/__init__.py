"""Region graph: scope bookkeeping and smoothness / decomposability checks.

Mirrors the behaviour of ``pyjuice/graph/region_graph.py`` (classes of the same names,
class-level flags ``ALLOW_NONSMOOTH`` / ``ALLOW_NONDECOMPOSABLE``).
"""
from __future__ import annotations

from typing import List, Union

from pyjuice_amd.utils import BitSet

__all__ = ["RegionGraph", "PartitionNode", "InnerRegionNode", "InputRegionNode"]


class RegionGraph:
    ALLOW_NONSMOOTH = False
    ALLOW_NONDECOMPOSABLE = False

    def __init__(self, scope: BitSet, children: List["RegionGraph"]):
        self.scope = scope
        self.children = children
        self.num_chs = len(children)
        self.parents = []
        for c in children:
            c.parents.append(self)


class PartitionNode(RegionGraph):
    """Region of a product: children scopes must be disjoint."""

    def __init__(self, children):
        assert len(children) > 0, "a PartitionNode needs at least one child region"
        scope = BitSet()
        for c in children:
            if not RegionGraph.ALLOW_NONDECOMPOSABLE:
                assert not scope.contains_any(c.scope), "the child regions of a PartitionNode must have disjoint scopes"
            scope = scope | c.scope
        super().__init__(scope, children)


class InnerRegionNode(RegionGraph):
    """Region of a sum: children scopes must be identical."""

    def __init__(self, children):
        assert len(children) > 0, "an InnerRegionNode needs at least one child partition"
        scope = BitSet(children[0].scope._mask)
        for c in children[1:]:
            if not RegionGraph.ALLOW_NONSMOOTH:
                assert scope == c.scope, "the child partitions of an InnerRegionNode must share one scope"
            else:
                scope = scope | c.scope
        super().__init__(scope, children)


class InputRegionNode(RegionGraph):
    def __init__(self, scope: Union[List[int], BitSet]):
        if not isinstance(scope, BitSet):
            scope = BitSet.from_array(scope)
        super().__init__(scope, [])
