"""Poon-Domingos structures (``pyjuice/structures/pd.py:17-228``).

The data hypercube is split recursively at the given split points along every axis; each
split yields a product of the two halves and a region's sum node mixes all its splits.  A
region with more than ``max_prod_block_conns`` splits connects every node block to
``max_prod_block_conns`` of them chosen by ``torch.topk(torch.rand(...))`` (torch CPU RNG,
same draws as the reference)."""
from __future__ import annotations

from copy import deepcopy
from typing import Callable, Dict, Optional, Sequence, Tuple, Type, Union

import torch

from pyjuice_amd.nodes import inputs, multiply, set_block_size, summate
from pyjuice_amd.nodes.distributions import Categorical, Distribution
from pyjuice_amd.utils import BitSet, max_cdf_power_of_2

__all__ = ["PD"]


def PD(data_shape: Tuple, num_latents: int, split_intervals: Optional[Union[int, Tuple[int]]] = None,
       split_points: Optional[Sequence[Sequence[int]]] = None, max_split_depth: Optional[int] = None,
       max_prod_block_conns: int = 4, structure_type: str = "sum_dominated",
       input_ns_fn: Optional[Callable] = None, input_dist: Optional[Distribution] = None,
       input_node_type: Type[Distribution] = Categorical, input_node_params: Dict = {"num_cats": 256},
       tie_homogeneous_params: bool = False, block_size: Optional[int] = None):
    assert structure_type in ["sum_dominated", "prod_dominated"]
    if input_dist is not None:
        input_node_type, input_node_params = input_dist._get_constructor()
    if block_size is None:
        block_size = min(16, max_cdf_power_of_2(num_latents)) if num_latents <= 32 else \
            min(32, max_cdf_power_of_2(num_latents))
    nblk = num_latents // block_size
    naxes = len(data_shape)
    source_ns = dict()

    if split_intervals is not None:
        if isinstance(split_intervals, int):
            split_intervals = (split_intervals,) * naxes
        split_points = [[j for j in range(split_intervals[a], data_shape[a], split_intervals[a])]
                        for a in range(naxes)]
    else:
        assert split_points is not None and len(split_points) == naxes
        for a in range(naxes):
            assert all(0 < x < data_shape[a] - 1 for x in split_points[a])
    gsp = split_points
    cache = dict()

    def cube_scope(cube):
        vars_ = None
        cum = 1
        for a in range(naxes - 1, -1, -1):
            s, e = cube[0][a], cube[1][a]
            if vars_ is None:
                vars_ = torch.arange(s, e)
            else:
                vars_ = (vars_[None, :].repeat(e - s, 1) + (torch.arange(s, e) * cum)[:, None]).reshape(-1)
            cum *= data_shape[a]
        return BitSet.from_array(vars_.tolist())

    def upd(cube, a, s=None, e=None):
        lo, hi = list(cube[0]), list(cube[1])
        if s is not None:
            lo[a] = s
        if e is not None:
            hi[a] = e
        return (tuple(lo), tuple(hi))

    def shape_of(cube):
        return tuple(cube[1][i] - cube[0][i] for i in range(len(cube[0])))

    def signature(cube, *chs):
        return (shape_of(cube), tuple(len(ns.scope) for ns in chs))

    def create_input_ns(cube):
        scope = cube_scope(cube)
        if input_ns_fn is not None:
            return input_ns_fn(scope, num_latents, block_size)
        ins = []
        for v in scope:
            if not tie_homogeneous_params:
                ns = inputs(v, num_node_blocks=nblk, dist=input_node_type(**input_node_params))
            elif "input" in source_ns:
                ns = source_ns["input"].duplicate(v, tie_params=True)
            else:
                ns = inputs(v, num_node_blocks=nblk, dist=input_node_type(**input_node_params))
                source_ns["input"] = ns
            ins.append(ns)
        eids = torch.arange(0, nblk)[None, :].repeat(2, 1)
        pns = multiply(*ins)
        if not tie_homogeneous_params:
            return summate(pns, num_node_blocks=nblk, edge_ids=eids)
        sig = signature(cube, pns)
        if sig in source_ns:
            return source_ns[sig].duplicate(pns, tie_params=True)
        ns = summate(pns, num_node_blocks=nblk, edge_ids=eids)
        source_ns[sig] = ns
        return ns

    def build(cube, depth=1):
        if cube in cache:
            return cache[cube]
        if max_split_depth is not None and depth > max_split_depth:
            ns = create_input_ns(cube)
            cache[cube] = ns
            return ns
        pns = []
        for a in range(naxes):
            s, e = cube[0][a], cube[1][a]
            for sp in [p for p in gsp[a] if s < p < e]:
                ns1 = build(upd(cube, a, e=sp))
                ns2 = build(upd(cube, a, s=sp))
                pns.append(multiply(ns1, ns2) if structure_type == "sum_dominated"
                           else multiply(ns1.duplicate(), ns2.duplicate()))
        if len(pns) == 0:
            ns = create_input_ns(cube)
        elif cube == root_cube:
            ns = summate(*pns, num_node_blocks=1, block_size=1)
        elif len(pns) <= max_prod_block_conns:
            if not tie_homogeneous_params:
                ns = summate(*pns, num_node_blocks=nblk)
            else:
                sig = signature(cube, *pns)
                if sig in source_ns:
                    ns = source_ns[sig].duplicate(*pns, tie_params=True)
                else:
                    ns = summate(*pns, num_node_blocks=nblk)
                    source_ns[sig] = ns
        else:
            block_ids = torch.topk(torch.rand([nblk, len(pns)]), k=max_prod_block_conns, dim=1).indices
            par = torch.arange(0, nblk)[:, None, None].repeat(1, max_prod_block_conns, nblk)
            chs = block_ids[:, :, None] * nblk + torch.arange(0, nblk)[None, None, :]
            eids = torch.stack((par.reshape(-1), chs.reshape(-1)), dim=0)
            if not tie_homogeneous_params:
                ns = summate(*pns, num_node_blocks=nblk, edge_ids=eids)
            else:
                sig = signature(cube, *pns)
                if sig in source_ns:
                    ns = source_ns[sig].duplicate(*pns, tie_params=True)
                else:
                    ns = summate(*pns, num_node_blocks=nblk, edge_ids=eids)
                    source_ns[sig] = ns
        cache[cube] = ns
        return ns

    with set_block_size(block_size=block_size):
        root_cube = ((0,) * naxes, deepcopy(tuple(data_shape)))
        root_ns = build(root_cube)
    return root_ns
