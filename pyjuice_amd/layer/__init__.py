from .layer import Layer
from .input_layer import InputLayer
from .prod_layer import ProdLayer
from .sum_layer import SumLayer
from .layer_group import LayerGroup
from .partition import partition_nodes_by_n_edges

__all__ = ["Layer", "InputLayer", "ProdLayer", "SumLayer", "LayerGroup", "partition_nodes_by_n_edges"]
