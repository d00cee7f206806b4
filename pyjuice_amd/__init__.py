"""pyjuice_amd: MI355X-native probabilistic-circuit engine with PyJuice's API surface.

``import pyjuice_amd as juice`` provides ``juice.compile`` / ``juice.TensorCircuit``,
``juice.inputs`` / ``multiply`` / ``summate``, the HCLT / PD / HMM builders under
``juice.structures``, ``juice.optim.CircuitOptimizer`` and ``juice.queries.sample``.
"""
from . import utils, graph, nodes
from .nodes import (inputs, multiply, summate, set_block_size, structural_properties, distributions,
                    CircuitNodes, InputNodes, ProdNodes, SumNodes)
from .layer import InputLayer, ProdLayer, SumLayer, LayerGroup
from .model import TensorCircuit, compile, log_likelihood
from . import structures, optim, queries, transformations, autotune, io

__all__ = ["compile", "TensorCircuit", "log_likelihood", "inputs", "multiply", "summate", "set_block_size",
           "structural_properties", "distributions", "structures", "optim", "queries", "transformations",
           "autotune", "InputLayer", "ProdLayer", "SumLayer", "LayerGroup"]
