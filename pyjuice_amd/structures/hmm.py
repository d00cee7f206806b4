"""Hidden Markov models as circuits (``pyjuice/structures/hmm.py:15-86``).

Built backwards from the last position: ``z_{t}`` is a sum over ``multiply(x_{t+1}, z_{t+1})``;
with ``homogeneous=True`` every transition sum and every emission vector is tied to the first
one created."""
from __future__ import annotations

from typing import Callable, Optional

import torch

from pyjuice_amd.nodes import inputs, multiply, set_block_size, summate
from pyjuice_amd.nodes.distributions import Categorical
from pyjuice_amd.utils import max_cdf_power_of_2

__all__ = ["HMM"]


def HMM(seq_length: int, num_latents: int, num_emits: int, homogeneous: bool = True,
        block_size: Optional[int] = None, alpha: Optional[torch.Tensor] = None, beta: Optional[torch.Tensor] = None,
        gamma: Optional[torch.Tensor] = None, sum_edge_ids_constructor: Optional[Callable] = None):
    if block_size is None:
        block_size = min(max_cdf_power_of_2(num_latents), 1024)
    nblk = num_latents // block_size
    with set_block_size(block_size=block_size):
        ns_input = inputs(seq_length - 1, num_node_blocks=nblk, dist=Categorical(num_cats=num_emits))
        if beta is not None:
            assert beta.size(0) == num_latents and beta.size(1) == num_emits
            ns_input.set_params(beta)
        ns_sum = None
        curr_zs = ns_input
        for var in range(seq_length - 2, -1, -1):
            curr_xs = ns_input.duplicate(var, tie_params=homogeneous)
            if ns_sum is None:
                ns = summate(curr_zs, num_node_blocks=nblk, sum_edge_ids_constructor=sum_edge_ids_constructor)
                if alpha is not None:
                    assert alpha.size(0) == num_latents and alpha.size(1) == num_latents
                    ns.set_params(alpha)
                ns_sum = ns
            else:
                ns = ns_sum.duplicate(curr_zs, tie_params=homogeneous)
            curr_zs = multiply(curr_xs, ns)
        ns = summate(curr_zs, num_node_blocks=1, block_size=1, sum_edge_ids_constructor=sum_edge_ids_constructor)
        if gamma is not None:
            assert gamma.dim() == 1 and gamma.size(0) == num_latents
            ns.set_params(gamma.unsqueeze(0))
    return ns
