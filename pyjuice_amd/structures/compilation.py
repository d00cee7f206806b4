"""Tree Bayesian network -> hidden region graph (``pyjuice/structures/compilation.py:12-102``)."""
from __future__ import annotations

from typing import Callable, Optional, Type

import networkx as nx

from pyjuice_amd.nodes import inputs, multiply, set_block_size, summate
from pyjuice_amd.utils import max_cdf_power_of_2

__all__ = ["BayesianTreeToHiddenRegionGraph"]


def BayesianTreeToHiddenRegionGraph(tree: nx.Graph, root, num_latents: int, InputDist: Type, dist_params: dict,
                                    num_root_ns: int = 1, block_size: Optional[int] = None,
                                    tie_input_params: bool = False,
                                    sum_edge_ids_constructor: Optional[Callable] = None):
    """Every variable ``v`` gets a latent ``z_v`` with ``num_latents`` states: leaves become input
    vectors, an inner variable becomes ``summate(multiply(children..., inputs(v)))``; the root
    sum has one node (block size 1) unless ``num_root_ns > 1``.  Children are taken in BFS-tree
    successor order and variables are built in DFS post-order from ``root``."""
    if block_size is None:
        block_size = min(1024, max_cdf_power_of_2(num_latents))
    num_node_blocks = num_latents // block_size

    clt = nx.bfs_tree(tree, root)
    for n in clt.nodes:
        assert len(list(clt.predecessors(n))) <= 1

    template = [None]

    def get_input_ns(v):
        if tie_input_params:
            if template[0] is None:
                template[0] = inputs(v, num_node_blocks=num_node_blocks, dist=InputDist(**dist_params))
                return template[0]
            return template[0].duplicate(v, tie_params=True)
        return inputs(v, num_node_blocks=num_node_blocks, dist=InputDist(**dist_params))

    var2rnode = dict()
    with set_block_size(block_size):
        for v in nx.dfs_postorder_nodes(tree, root):
            chs = list(clt.successors(v))
            if len(chs) == 0:
                var2rnode[v] = get_input_ns(v)
                continue
            regions = [var2rnode[c] for c in chs]
            regions.append(get_input_ns(v))
            rp = multiply(*regions)
            if v == root:
                if num_root_ns == 1:
                    r = summate(rp, num_node_blocks=num_root_ns, block_size=1,
                                sum_edge_ids_constructor=sum_edge_ids_constructor)
                else:
                    r = summate(rp, num_node_blocks=num_root_ns // block_size, block_size=block_size,
                                sum_edge_ids_constructor=sum_edge_ids_constructor)
            else:
                r = summate(rp, num_node_blocks=num_node_blocks, sum_edge_ids_constructor=sum_edge_ids_constructor)
            var2rnode[v] = r
    return var2rnode[root]
