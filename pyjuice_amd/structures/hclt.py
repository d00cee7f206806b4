"""Hidden Chow-Liu trees (``pyjuice/structures/hclt.py``).

The backbone tree is the maximum-mutual-information spanning tree (networkx minimum spanning
tree over negated MI, rooted at ``nx.center(T)[0]``).  MI is estimated with Gaussian soft
binning: each chunk of variables is min-max scaled to [0, 1] (over the whole chunk), softly
assigned to ``num_bins`` evenly spaced bins with width ``sigma``, and
``MI = H(x1) + H(x2) - H(x1, x2)`` with ``log(p + 1e-4)`` entropies (``hclt.py:12-51``).
"""
from __future__ import annotations

from typing import Callable, Optional, Type

import networkx as nx
import numpy as np
import torch

from pyjuice_amd.nodes.distributions import Categorical, Distribution
from .compilation import BayesianTreeToHiddenRegionGraph

__all__ = ["HCLT", "mutual_information", "mutual_information_chunked", "chow_liu_tree"]


def _soft_bins(x: torch.Tensor, bins: torch.Tensor, sigma: float) -> torch.Tensor:
    x = (x - torch.min(x)) / (torch.max(x) - torch.min(x) + 1e-8)
    return torch.exp(-0.5 * (x.unsqueeze(2) - bins.view(1, 1, -1)).pow(2) / sigma ** 2)  # [B, K, bins]


def mutual_information(x1: torch.Tensor, x2: torch.Tensor, num_bins: int, sigma: float) -> torch.Tensor:
    assert x1.device == x2.device
    B, K1 = x1.shape
    K2 = x2.size(1)
    bins = torch.linspace(0, 1, num_bins, device=x1.device)
    p1 = _soft_bins(x1, bins, sigma)
    p2 = _soft_bins(x2, bins, sigma)
    # The joint histogram is formed with the same torch contraction (subscripts and operand shapes)
    # as hclt.py:27, so that the same torch build rounds it identically and near-tied spanning-tree
    # edges resolve the same way.
    joint = torch.einsum("bia,baj->ij", p1.reshape(B, K1 * num_bins, 1),
                         p2.reshape(B, 1, K2 * num_bins)).reshape(K1, num_bins, K2, num_bins) / B
    m1 = (p1 / p1.sum(dim=2, keepdim=True)).mean(dim=0)
    m2 = (p2 / p2.sum(dim=2, keepdim=True)).mean(dim=0)
    j = joint / joint.sum(dim=(1, 3), keepdim=True)
    h1 = -(m1 * torch.log(m1 + 1e-4)).sum(dim=1)
    h2 = -(m2 * torch.log(m2 + 1e-4)).sum(dim=1)
    h12 = -(j * torch.log(j + 1e-4)).sum(dim=(1, 3))
    return h1.unsqueeze(1) + h2.unsqueeze(0) - h12


def mutual_information_chunked(x1: torch.Tensor, x2: torch.Tensor, num_bins: int, sigma: float,
                               chunk_size: int) -> torch.Tensor:
    K = x1.size(1)
    mi = torch.zeros([K, K])
    for xs in range(0, K, chunk_size):
        xe = min(xs + chunk_size, K)
        for ys in range(0, K, chunk_size):
            ye = min(ys + chunk_size, K)
            mi[xs:xe, ys:ye] = mutual_information(x1[:, xs:xe], x2[:, ys:ye], num_bins, sigma).cpu()
    return mi


def chow_liu_tree(mi: np.ndarray) -> nx.Graph:
    K = mi.shape[0]
    G = nx.Graph()
    for v in range(K):
        G.add_node(v)
        for u in range(v):
            G.add_edge(u, v, weight=-mi[u, v])
    return nx.minimum_spanning_tree(G)


def HCLT(x: torch.Tensor, num_latents: int, num_bins: int = 32, sigma: float = 0.5 / 32, chunk_size: int = 64,
         num_root_ns: int = 1, block_size: Optional[int] = None, input_dist: Optional[Distribution] = None,
         input_node_type: Type[Distribution] = Categorical, input_node_params: dict = {"num_cats": 256},
         tie_input_params: bool = False, sum_edge_ids_constructor: Optional[Callable] = None):
    """Build an HCLT over the columns of ``x`` ``[num_samples, num_vars]`` (``hclt.py:67-125``)."""
    if input_dist is not None:
        input_node_type, input_node_params = input_dist._get_constructor()
    mi = mutual_information_chunked(x, x, num_bins, sigma, chunk_size=chunk_size).detach().cpu().numpy()
    T = chow_liu_tree(mi)
    root = nx.center(T)[0]
    return BayesianTreeToHiddenRegionGraph(T, root, num_latents, input_node_type, input_node_params,
                                           num_root_ns=num_root_ns, block_size=block_size,
                                           tie_input_params=tie_input_params,
                                           sum_edge_ids_constructor=sum_edge_ids_constructor)
