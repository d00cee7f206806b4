"""Vectors of input nodes (``pyjuice/nodes/input_nodes.py``)."""
from __future__ import annotations

from copy import deepcopy
from typing import Optional, Sequence, Union

import torch

from pyjuice_amd.graph import InputRegionNode
from pyjuice_amd.utils import BitSet
from .distributions import Distribution
from .nodes import CircuitNodes

__all__ = ["InputNodes"]


class InputNodes(CircuitNodes):
    def __init__(self, num_node_blocks: int, scope: Union[Sequence[int], BitSet], dist: Distribution,
                 params: Optional[torch.Tensor] = None, block_size: int = 0, **kwargs):
        rg = InputRegionNode(scope)
        super().__init__(num_node_blocks, rg, block_size=block_size, **kwargs)
        self.chs = []
        self.dist = dist
        if params is not None:
            self.set_params(params)
        self._run_init_callbacks(**kwargs)
        # Set by ``set_params`` only (input_nodes.py:55-56).
        self._param_initialized = params is not None

    def is_input(self):
        return True

    @property
    def num_edges(self):
        return 0

    def duplicate(self, scope=None, tie_params: bool = False) -> "InputNodes":
        if scope is None:
            scope = self.scope
        else:
            if isinstance(scope, int):
                scope = [scope]
            assert len(scope) == len(self.scope)
        ns = InputNodes(self.num_node_blocks, scope=scope, dist=deepcopy(self.dist), block_size=self.block_size,
                        source_node=self if tie_params else None)
        if getattr(self, "_params", None) is not None and not tie_params:
            ns._params = self._params.clone()
        return ns

    def get_params(self):
        return self._params if self.provided("_params") else None

    def get_param_flows(self):
        return self._param_flows if self.provided("_param_flows") else None

    def set_params(self, params: torch.Tensor, normalize: bool = True):
        assert params.numel() == self.num_nodes * self.dist.num_parameters()
        params = params.reshape(-1)
        if normalize:
            params = self.dist.normalize_parameters(params)
        self._param_initialized = True
        self._params = params

    def _draw_own_params(self, perturbation: float, **kwargs):
        if not self.is_tied() and not self.has_params():
            self._params = self.dist.init_parameters(num_nodes=self.num_nodes, perturbation=perturbation,
                                                     params=self.get_params(), **kwargs)

    def init_parameters(self, perturbation: float = 2.0, recursive: bool = True, is_root: bool = True,
                        ret_params: bool = False, **kwargs):
        if not self.is_tied() and not self.has_params():
            self._params = self.dist.init_parameters(num_nodes=self.num_nodes, perturbation=perturbation,
                                                     params=self.get_params(), **kwargs)
            if ret_params:
                return self._params
        elif self.is_tied() and ret_params:
            return self.get_source_ns().init_parameters(perturbation=perturbation, recursive=False,
                                                         is_root=True, ret_params=True, **kwargs)
        return None

    def get_data_dtype(self):
        return self.dist.get_data_dtype()

    def __repr__(self):
        return f"<InputNodes {self.num_node_blocks}x{self.block_size} over {len(self.scope)} var(s), {type(self.dist).__name__}>"
