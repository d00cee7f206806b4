from .nodes import CircuitNodes, node_iterator
from .input_nodes import InputNodes
from .prod_nodes import ProdNodes
from .sum_nodes import SumNodes, normalize_ns_parameters
from .construction import inputs, multiply, summate, set_block_size, structural_properties
from . import distributions

__all__ = ["CircuitNodes", "InputNodes", "ProdNodes", "SumNodes", "inputs", "multiply", "summate",
           "set_block_size", "structural_properties", "distributions", "normalize_ns_parameters", "node_iterator"]
