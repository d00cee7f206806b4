"""Marginal queries (``pyjuice/queries/marginal.py:17-41``): log P(e) of the observed part of each
sample, the variables under ``missing_mask`` summed out."""
from __future__ import annotations

from typing import Callable, Optional, Union

import torch

__all__ = ["marginal"]


def marginal(pc, data: torch.Tensor, missing_mask: Optional[torch.Tensor] = None,
             fw_input_fn: Optional[Union[str, Callable]] = None, **kwargs) -> torch.Tensor:
    """``data [B, num_vars]`` hard evidence; ``missing_mask`` ``[num_vars]`` or ``[B, num_vars]``
    (True = marginalized).  Returns ``lls [B, num_roots]``: the input forward gives marginalized
    entries log-probability 0 and the inner layers run unchanged, which sums those variables out
    of a smooth, decomposable circuit with normalized input distributions."""
    return pc.forward(data, input_layer_fn=fw_input_fn, missing_mask=missing_mask, **kwargs)
