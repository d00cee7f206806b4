"""Group of same-kind layers evaluated together (``pyjuice/layer/layer_group.py``)."""
from __future__ import annotations

from typing import Callable, Optional, Sequence

import torch.nn as nn

__all__ = ["LayerGroup"]


class LayerGroup(nn.Module):
    def __init__(self, layers: Sequence):
        super().__init__()
        assert len(layers) >= 1, "a layer group needs at least one layer"
        for i in range(1, len(layers)):
            assert type(layers[i]) == type(layers[0])
        if layers[0].is_input():
            self.layer_type = "input"
        elif layers[0].is_prod():
            self.layer_type = "prod"
        else:
            assert layers[0].is_sum()
            self.layer_type = "sum"
        self.num_layers = len(layers)
        self.layers = list(layers)
        for i, layer in enumerate(self.layers):
            self.add_module(f"layer_{i}", layer)

    def to(self, device):
        for layer in self.layers:
            layer.to(device)
        return self

    def forward(self, *args, **kwargs):
        for layer in self.layers:
            layer.forward(*args, **kwargs)

    def backward(self, *args, **kwargs):
        for layer in self.layers:
            layer.backward(*args, **kwargs)

    def callback(self, callback_fn: Optional[Callable], *args, **kwargs):
        if callback_fn is not None:
            for layer in self.layers:
                callback_fn(layer, *args, **kwargs)

    def is_input(self):
        return self.layer_type == "input"

    def is_prod(self):
        return self.layer_type == "prod"

    def is_sum(self):
        return self.layer_type == "sum"

    def __len__(self):
        return self.num_layers

    def __getitem__(self, idx):
        return self.layers[idx]

    def __iter__(self):
        return iter(self.layers)
