"""Input distributions.

Host-side mirror of the distribution plugin interface of
``pyjuice/nodes/distributions/distributions.py:8``.  The reference inlines Triton snippets
(``fw_mar_fn`` / ``bk_flow_fn`` / ``em_fn`` / ``sample_fn``) into kernel templates; here
each distribution names a native kernel family by ``native_kind`` and the HIP library
holds the fused kernels for that family (Categorical, Bernoulli).
"""
from __future__ import annotations

from typing import Any, Optional

import torch

__all__ = ["Distribution", "Categorical", "Bernoulli"]


class Distribution:
    """Base class; see ``pyjuice/nodes/distributions/distributions.py:8-207``."""

    native_kind = -1

    def __init__(self):
        self.post_fw_fns = []
        self.post_bp_fns = []
        self.sampling_fns = []

    def get_signature(self) -> str:
        raise NotImplementedError()

    def requires_external_inputs(self) -> bool:
        return "Extern" in self.get_signature()

    def get_metadata(self):
        return []

    def normalize_parameters(self, params: torch.Tensor, **kwargs):
        return params

    def num_parameters(self) -> int:
        raise NotImplementedError()

    def num_param_flows(self) -> int:
        raise NotImplementedError()

    def init_parameters(self, num_nodes: int, perturbation: float = 2.0, params: Optional[Any] = None, **kwargs):
        raise NotImplementedError()

    @property
    def need_meta_parameters(self) -> bool:
        return False

    def get_data_dtype(self):
        return torch.float32

    def _get_constructor(self):
        raise NotImplementedError()


class Categorical(Distribution):
    """Categorical over ``num_cats`` values (``pyjuice/nodes/distributions/categorical.py:12``).

    Per node: ``num_cats`` parameters and ``num_cats`` parameter flows; data dtype long.
    """

    native_kind = 0

    def __init__(self, num_cats: int):
        super().__init__()
        self.num_cats = int(num_cats)

    def get_signature(self):
        return "Categorical"

    def get_metadata(self):
        return [self.num_cats]

    def normalize_parameters(self, params: torch.Tensor, **kwargs):
        params = params.reshape(-1, self.num_cats)
        params = params / params.sum(dim=1, keepdim=True)
        return params.reshape(-1)

    def num_parameters(self):
        return self.num_cats

    def num_param_flows(self):
        return self.num_cats

    def init_parameters(self, num_nodes: int, perturbation: float = 2.0, params: Optional[Any] = None, **kwargs):
        # Same draw as categorical.py:57-67: exp(-perturbation * U), rows normalised.
        if params is not None:
            assert isinstance(params, torch.Tensor) and params.numel() == num_nodes * self.num_parameters()
            return params
        p = torch.exp(torch.rand([num_nodes, self.num_cats]) * -perturbation)
        p /= p.sum(dim=1, keepdim=True)
        return p.reshape(-1)

    def get_data_dtype(self):
        return torch.long

    def _get_constructor(self):
        return Categorical, {"num_cats": self.num_cats}

    def __repr__(self):
        return f"Categorical(num_cats={self.num_cats})"


class Bernoulli(Distribution):
    """Bernoulli (``pyjuice/nodes/distributions/bernoulli.py:12``): one parameter P(x=1),
    two parameter-flow slots per node (slot 0 for x=1, slot 1 for x=0)."""

    native_kind = 1

    def get_signature(self):
        return "Bernoulli"

    def get_metadata(self):
        return []

    def num_parameters(self):
        return 1

    def num_param_flows(self):
        return 2

    def init_parameters(self, num_nodes: int, perturbation: float = 2.0, params: Optional[Any] = None, **kwargs):
        # Same draw as bernoulli.py:43-57.
        if params is not None:
            assert isinstance(params, torch.Tensor) and params.numel() == num_nodes * self.num_parameters()
            return params
        p = torch.exp(torch.rand([num_nodes, 2]) * -perturbation)
        p /= p.sum(dim=1, keepdim=True)
        return p[:, 0].contiguous()

    def get_data_dtype(self):
        return torch.bool

    def _get_constructor(self):
        return Bernoulli, {}

    def __repr__(self):
        return "Bernoulli()"
