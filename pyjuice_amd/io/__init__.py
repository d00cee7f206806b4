"""Circuit files (``.jpc``): the reference's on-disk format (``pyjuice/io/io.py:13-53``,
``io/serialization.py:11-118``), so the same node DAG, with its parameters, moves between the
reference and this package.

A file is a pickled list with one record per node vector in DFS post-order (children first, the
root last): ``type`` ("Input" / "Product" / "Sum"), ``num_node_blocks``, ``block_size``,
``chs`` (record indices of the children), ``edge_ids`` and ``params`` / ``zero_param_mask``
(numpy arrays), ``scope`` and a pickled ``dist`` for input vectors, ``source_node`` for tied
vectors.  The record layout is the reference's, so either program reads the other's files.

``save`` writes each distribution under the reference's class path
(``pyjuice.nodes.distributions.categorical.Categorical``, ``...bernoulli.Bernoulli``).  ``load``
unpickles with an allow-list: numpy arrays and the two distribution classes (under either
package's path); any other global in the file raises ``pickle.UnpicklingError`` before it is
imported, so loading a file cannot run code.
"""
from __future__ import annotations

import importlib
import io as _io
import pickle
from typing import List, Union

import numpy as np
import torch

from pyjuice_amd.nodes import CircuitNodes, inputs, multiply, summate
from pyjuice_amd.nodes.distributions import Bernoulli, Categorical

__all__ = ["save", "load", "serialize_nodes", "deserialize_nodes"]

_REF_DIST_PATH = {"Categorical": "pyjuice.nodes.distributions.categorical",
                  "Bernoulli": "pyjuice.nodes.distributions.bernoulli"}
_DISTS = {("pyjuice.nodes.distributions.categorical", "Categorical"): Categorical,
          ("pyjuice.nodes.distributions.bernoulli", "Bernoulli"): Bernoulli,
          ("pyjuice_amd.nodes.distributions", "Categorical"): Categorical,
          ("pyjuice_amd.nodes.distributions", "Bernoulli"): Bernoulli}
_NUMPY = {("numpy.core.multiarray", "_reconstruct"), ("numpy._core.multiarray", "_reconstruct"),
          ("numpy.core.multiarray", "scalar"), ("numpy._core.multiarray", "scalar"),
          ("numpy.core.numeric", "_frombuffer"), ("numpy._core.numeric", "_frombuffer"),
          ("numpy", "ndarray"), ("numpy", "dtype"),
          ("_codecs", "encode")}            # bytes in protocol-2 pickles


class _AllowListUnpickler(pickle.Unpickler):
    def find_class(self, module, name):
        if (module, name) in _DISTS:
            return _DISTS[(module, name)]
        if (module, name) in _NUMPY:
            return getattr(importlib.import_module(module), name)
        if module.startswith("numpy") and name.endswith("DType"):      # numpy >= 1.25 dtype classes
            return getattr(importlib.import_module(module), name)
        raise pickle.UnpicklingError(f"global {module}.{name} is not allowed in a circuit file")


def _safe_loads(data: bytes):
    return _AllowListUnpickler(_io.BytesIO(data)).load()


def _dist_bytes(dist) -> bytes:
    """A distribution pickled as the reference's class, built from opcodes (the reference package
    need not be importable): GLOBAL, EMPTY_TUPLE, NEWOBJ, the state dict, BUILD, STOP."""
    sig = dist.get_signature()
    assert sig in _REF_DIST_PATH, f"distribution {sig} has no circuit-file form"
    state = {"post_fw_fns": [], "post_bp_fns": [], "sampling_fns": []}
    if sig == "Categorical":
        state["num_cats"] = int(dist.num_cats)
    body = pickle.dumps(state, protocol=2)
    assert body[:2] == b"\x80\x02" and body[-1:] == b"."
    return (b"\x80\x02" + f"c{_REF_DIST_PATH[sig]}\n{sig}\n".encode() + b")\x81" + body[2:-1] + b"b.")


def _dist_from(obj):
    """A fresh distribution of this package from an unpickled one (whichever class path)."""
    if isinstance(obj, Categorical):
        return Categorical(num_cats=int(obj.__dict__["num_cats"]))
    if isinstance(obj, Bernoulli):
        return Bernoulli()
    raise TypeError(f"unsupported distribution {type(obj)}")


def serialize_nodes(root_ns: CircuitNodes) -> List[dict]:
    """One record per node vector, children before parents (``serialization.py:11-52``)."""
    records, index = [], dict()
    for ns in root_ns:
        rec = {"type": "Input" if ns.is_input() else ("Product" if ns.is_prod() else "Sum"),
               "num_node_blocks": ns.num_node_blocks, "block_size": ns.block_size,
               "chs": tuple(index[c] for c in ns.chs)}
        if not ns.is_input():
            rec["edge_ids"] = ns.edge_ids.detach().cpu().numpy().copy()
        if getattr(ns, "_params", None) is not None:
            rec["params"] = ns._params.detach().cpu().numpy().copy()
        if getattr(ns, "_zero_param_mask", None) is not None:
            rec["zero_param_mask"] = ns._zero_param_mask.detach().cpu().numpy().copy()
        if ns.is_input():
            rec["scope"] = ns.scope.to_list()
            rec["dist"] = _dist_bytes(ns.dist)
        index[ns] = len(records)
        records.append(rec)
    for ns in root_ns:
        if getattr(ns, "_source_node", None) is not None:
            records[index[ns]]["source_node"] = index[ns._source_node]
    return records


def deserialize_nodes(records: List[dict]) -> CircuitNodes:
    """The node DAG of a record list (``serialization.py:55-118``); parameters are taken as
    stored (no renormalisation)."""
    built = []
    for rec in records:
        if "num_node_blocks" in rec:
            nblocks, bs = int(rec["num_node_blocks"]), int(rec["block_size"])
        elif "num_node_groups" in rec:          # the format's earlier key names
            nblocks, bs = int(rec["num_node_groups"]), int(rec["group_size"])
        else:
            nblocks, bs = int(rec["num_nodes"]), 1
        chs = [built[i] for i in rec["chs"]]
        if rec["type"] == "Input":
            dist = rec["dist"]
            dist = _dist_from(_safe_loads(dist) if isinstance(dist, (bytes, bytearray)) else dist)
            ns = inputs(rec["scope"], num_node_blocks=nblocks, dist=dist, block_size=bs)
            if "params" in rec:
                ns._params = torch.from_numpy(np.asarray(rec["params"])).clone()
        elif rec["type"] == "Product":
            edge_ids = torch.from_numpy(np.asarray(rec["edge_ids"]))
            ns = multiply(*chs, edge_ids=edge_ids, sparse_edges=edge_ids.size(0) != chs[0].num_node_blocks)
        else:
            assert rec["type"] == "Sum", f"unknown record type {rec['type']}"
            ns = summate(*chs, edge_ids=torch.from_numpy(np.asarray(rec["edge_ids"])), block_size=bs)
            if "zero_param_mask" in rec:
                ns.set_zero_param_mask(torch.from_numpy(np.asarray(rec["zero_param_mask"])))
            if "params" in rec:
                ns._params = torch.from_numpy(np.asarray(rec["params"])).clone()
        if "params" not in rec and not rec["type"] == "Product":
            ns._params = None        # as saved: tied vectors read their source's parameters
        built.append(ns)
    for ns, rec in zip(built, records):
        if "source_node" in rec:
            ns._source_node = built[rec["source_node"]]
    return built[-1]


def save(fname: str, model: Union[CircuitNodes, "TensorCircuit"]):
    """Writes an uncompiled (``CircuitNodes``) or compiled (``TensorCircuit``, parameters copied
    back first) circuit to ``fname`` (``.jpc``; ``io.py:13-33``)."""
    from pyjuice_amd.model import TensorCircuit
    if isinstance(model, TensorCircuit):
        model.update_parameters()
        model = model.root_ns
    if not fname.endswith(".jpc"):
        raise ValueError(f"circuit files end in .jpc (got {fname!r})")
    with open(fname, "wb") as f:
        pickle.dump(serialize_nodes(model), f, protocol=4)


def load(fname: str) -> CircuitNodes:
    """Reads a ``.jpc`` file (``io.py:36-53``) with the allow-list unpickler."""
    if not fname.endswith(".jpc"):
        raise ValueError(f"circuit files end in .jpc (got {fname!r})")
    with open(fname, "rb") as f:
        records = _AllowListUnpickler(f).load()
    return deserialize_nodes(records)
