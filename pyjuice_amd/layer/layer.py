"""Layer base class (``pyjuice/layer/layer.py``)."""
from __future__ import annotations

from typing import Sequence

import torch

__all__ = ["Layer"]


class Layer:
    propagation_alg_mapping = {"LL": 0, "MPE": 1, "GeneralLL": 2}

    def __init__(self, nodes: Sequence, disable_block_size_check: bool = False):
        self.nodes = list(nodes)
        self.scopes = []
        seen = set()
        for ns in self.nodes:
            if ns.scope not in seen:
                seen.add(ns.scope)
                self.scopes.append(ns.scope)
        if disable_block_size_check:
            self.block_size = None
        else:
            for i in range(1, len(self.nodes)):
                assert self.nodes[i].block_size == self.nodes[0].block_size, \
                    "all vectors of one layer must share a block size"
            self.block_size = self.nodes[0].block_size
        self.device = torch.device("cpu")

    def provided(self, name):
        return getattr(self, name, None) is not None

    def is_input(self):
        return False

    def is_prod(self):
        return False

    def is_sum(self):
        return False
