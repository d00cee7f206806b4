"""Circuits of the reference's known-answer tests, rebuilt with this package's API.

Each builder returns a dict of named objects used by the expressions in
``tests/golden/kat_compile.json`` (and by the GPU parity tests)."""
import torch

import pyjuice_amd as juice
import pyjuice_amd.nodes.distributions as dists
from pyjuice_amd.layer import InputLayer, ProdLayer, SumLayer
from pyjuice_amd.model import TensorCircuit
from pyjuice_amd.nodes import inputs, multiply, summate


def simple_model_dag(seed=82392, block_size=16):
    """``tests/model/simple_model_test.py:19-47``."""
    torch.manual_seed(seed)
    with juice.set_block_size(block_size):
        ni0 = inputs(0, num_node_blocks=2, dist=dists.Categorical(num_cats=4))
        ni1 = inputs(1, num_node_blocks=2, dist=dists.Categorical(num_cats=4))
        ni2 = inputs(2, num_node_blocks=2, dist=dists.Categorical(num_cats=6))
        ni3 = inputs(3, num_node_blocks=2, dist=dists.Categorical(num_cats=6))
        np0 = multiply(ni0, ni1)
        np1 = multiply(ni2, ni3)
        np2 = multiply(ni1, ni2)
        np3 = multiply(ni0, ni1)
        ns0 = summate(np0, np3, num_node_blocks=2)
        ns1 = summate(np1, num_node_blocks=2)
        ns2 = summate(np2, num_node_blocks=2)
        np4 = multiply(ns0, ni2, ni3)
        np5 = multiply(ns1, ni0, ni1)
        np6 = multiply(ns2, ni0, ni3)
        ns = summate(np4, np5, np6, num_node_blocks=1, block_size=1)
    ns.init_parameters()
    return ns


def simple_model():
    ns = simple_model_dag()
    pc = TensorCircuit(ns, layer_sparsity_tol=0.1)
    ps, pf, bsz, itv, gn, nch, cum, meta = pc.par_update_kwargs

    class EM:
        pass

    em = EM()
    em.par_start_ids, em.pflow_start_ids, em.blk_sizes, em.blk_intervals = ps, pf, bsz, itv
    em.global_nids, em.nchs, em.cum_pflows = gn, nch, cum
    return {"pc": pc, "il": pc.input_layer_group[0], "pl0": pc.inner_layer_groups[0][0],
            "sl0": pc.inner_layer_groups[1][0], "pl1": pc.inner_layer_groups[2][0],
            "sl1": pc.inner_layer_groups[3][0], "em": em}


def _four_inputs(block_size, num_cats=2):
    with juice.set_block_size(block_size):
        return [inputs(v, num_node_blocks=2, dist=dists.Categorical(num_cats=num_cats)) for v in range(4)]


def sum_layer(block_size=16):
    """``tests/layer/sum_layer_test.py:19-47`` (also the corner cases of ``:148-179``)."""
    ni0, ni1, ni2, ni3 = _four_inputs(block_size)
    with juice.set_block_size(block_size):
        np0 = multiply(ni0, ni1)
        np1 = multiply(ni2, ni3)
        np2 = multiply(ni1, ni2)
        ns0 = summate(np0, num_node_blocks=2)
        ns1 = summate(np1, num_node_blocks=2)
        ns2 = summate(np2, num_node_blocks=2)
    il = InputLayer([ni0, ni1, ni2, ni3], cum_nodes=block_size)
    pl = ProdLayer([np0, np1, np2])
    layer = SumLayer([ns0, ns1, ns2], global_nid_start=block_size, global_pid_start=block_size ** 2,
                     global_pfid_start=0, node2tiednodes=dict())
    return {"layer": layer, "il": il, "pl": pl}


def prod_layer(block_size):
    """``tests/layer/prod_layer_test.py:22-36``."""
    ni0, ni1, ni2, ni3 = _four_inputs(block_size)
    with juice.set_block_size(block_size):
        np0 = multiply(ni0, ni1)
        np1 = multiply(ni2, ni3)
        np2 = multiply(ni1, ni2)
    InputLayer([ni0, ni1, ni2, ni3], cum_nodes=block_size)
    return {"layer": ProdLayer([np0, np1, np2])}


def input_layer(tied=False):
    """``tests/layer/input_layer_test.py:21-32`` and ``:144-154``."""
    with juice.set_block_size(4):
        ni0 = inputs(0, num_node_blocks=2, dist=dists.Categorical(num_cats=2))
        ni1 = inputs(1, num_node_blocks=2, dist=dists.Categorical(num_cats=2))
        ni2 = inputs(2, num_node_blocks=2, dist=dists.Categorical(num_cats=2))
        ni3 = ni1.duplicate(3, tie_params=True) if tied else inputs(3, num_node_blocks=2,
                                                                    dist=dists.Categorical(num_cats=2))
    if tied:
        layer = InputLayer([ni0, ni1, ni2, ni3], cum_nodes=1, pc_num_vars=3, max_tied_ns_per_parflow_block=1.0)
    else:
        layer = InputLayer([ni0, ni1, ni2, ni3], cum_nodes=1, pc_num_vars=4)
    layer._init_parameters(perturbation=2.0)
    return {"layer": layer}


def tying_block1_dag(seed=892910):
    """``tests/model/parameter_tying_test.py:17-46``."""
    torch.manual_seed(seed)
    with juice.set_block_size(1):
        ni0, ni1, ni2, ni3 = [inputs(v, num_node_blocks=2, dist=dists.Categorical(num_cats=5)) for v in range(4)]
        np01 = multiply(ni0, ni1)
        np12 = multiply(ni1, ni2)
        np23 = multiply(ni2, ni3)
        ns01 = summate(np01, num_node_blocks=2)
        ns12 = ns01.duplicate(np12, tie_params=True)
        ns23 = ns01.duplicate(np23, tie_params=True)
        np012_0 = multiply(ns01, ni2)
        np012_1 = multiply(ns12, ni0)
        ns012 = summate(np012_0, np012_1, num_node_blocks=2)
        np123_0 = multiply(ns12, ni3)
        np123_1 = multiply(ns23, ni1)
        ns123 = ns012.duplicate(np123_0, np123_1, tie_params=True)
        np0123_0 = multiply(ns012, ni3)
        np0123_1 = multiply(ns123, ni0)
        ns0123 = ns123.duplicate(np0123_0, np0123_1, tie_params=True)
    return ns0123


def tying_block1():
    pc = TensorCircuit(tying_block1_dag(), max_tied_ns_per_parflow_block=2)
    return {"pc": pc, "il": pc.input_layer_group[0], "g": [g[0] for g in pc.inner_layer_groups]}


BUILDERS = {
    "simple_model": simple_model,
    "sum_layer": lambda: sum_layer(16),
    "prod_layer_bs1": lambda: prod_layer(1),
    "prod_layer_bs8": lambda: prod_layer(8),
    "input_layer": lambda: input_layer(False),
    "input_layer_tied": lambda: input_layer(True),
    "tying_block1": tying_block1,
}
