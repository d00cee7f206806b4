"""Circuit files (``.jpc``, ``pyjuice/io/io.py``, ``io/serialization.py``), CPU only: a saved and
reloaded circuit compiles to the same index tables and parameters, bit for bit, and evaluates the
same under the FP64 oracle; the records keep the reference's layout (distributions under the
reference's class paths); the loader refuses any other global in a file."""
import pickle

import numpy as np
import pytest
import torch

import kat_circuits
import parity_utils as U
import pyjuice_amd as juice
from oracle import dag_eval
from pyjuice_amd.model import TensorCircuit


def _circuits():
    yield "simple_model", kat_circuits.simple_model_dag(), [4, 4, 6, 6]
    yield "tying_block1", kat_circuits.tying_block1_dag(), [5, 5, 5, 5]
    ns, x, cards = U.hclt_bernoulli()
    yield "hclt_bern16", ns, cards
    ns, x, cards = U.hclt_pruned(num_vars=8, L=64)
    yield "hclt_pruned", ns, cards
    ns, cards = U.pd_sparse()
    yield "pd_sparse", ns, cards
    ns, cards = U.hmm_small()
    yield "hmm_small", ns, cards


def _tables(pc):
    out = []
    for g in pc.inner_layer_groups:
        for layer in g:
            for name in ("partitioned_nids", "partitioned_cids", "partitioned_pids", "partitioned_pfids",
                         "partitioned_chids", "partitioned_parids", "partitioned_parpids",
                         "partitioned_u_cids"):
                for t in getattr(layer, name, []):
                    out.append(t.numpy().copy())
    for layer in pc.input_layer_group:
        for name in ("vids", "s_pids", "s_pfids"):
            out.append(getattr(layer, name).numpy().copy())
        out.append(layer.params.detach().numpy().copy())
    out.append(pc.params.detach().numpy().copy())
    return out


@pytest.mark.parametrize("case", list(_circuits()), ids=lambda c: c[0])
def test_round_trip(case, tmp_path):
    name, ns, cards = case
    ns.init_parameters()      # parameters on every vector (compiling draws any missing ones at random)
    f = str(tmp_path / f"{name}.jpc")
    juice.io.save(f, ns)
    ns2 = juice.io.load(f)
    a, b = _tables(TensorCircuit(ns)), _tables(TensorCircuit(ns2))
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert x.dtype == y.dtype and np.array_equal(x, y)
    X = U.random_data(cards, 16, seed=2)
    np.testing.assert_array_equal(dag_eval.forward(ns, X)[ns][0], dag_eval.forward(ns2, X)[ns2][0])


def test_compiled_circuit_saves_its_parameters(tmp_path):
    ns, x, cards = U.hclt_categorical(num_vars=6, K=5, L=32)
    from pyjuice_amd.model.backend import normalize_parameters
    pc = TensorCircuit(ns)
    with torch.no_grad():                 # normalized parameters that differ from the nodes'
        pc.params.mul_(torch.rand_like(pc.params) + 0.5)
        pc.params.copy_(normalize_parameters(pc.params, pc.par_update_kwargs))
    f = str(tmp_path / "pc.jpc")
    juice.io.save(f, pc)
    pc2 = TensorCircuit(juice.io.load(f))
    torch.testing.assert_close(pc2.params, pc.params, rtol=1e-6, atol=1e-9)


def test_record_layout_is_the_reference_format(tmp_path):
    ns = kat_circuits.simple_model_dag()
    f = str(tmp_path / "m.jpc")
    juice.io.save(f, ns)
    records = juice.io.serialize_nodes(ns)
    assert records[-1]["type"] == "Sum" and records[0]["type"] == "Input"
    keys = {"type", "num_node_blocks", "block_size", "chs"}
    for rec in records:
        assert keys <= set(rec)
        if rec["type"] == "Input":
            assert {"scope", "dist", "params"} <= set(rec)
            assert b"cpyjuice.nodes.distributions.categorical\nCategorical\n" in rec["dist"]
        else:
            assert isinstance(rec["edge_ids"], np.ndarray)
    with open(f, "rb") as fh:
        raw = fh.read()
    # the file is a pickle of that list whose only globals are numpy's and the distribution's
    assert b"pyjuice_amd" not in raw


def test_loader_refuses_other_globals(tmp_path):
    # GLOBAL builtins.print, a tuple argument, REDUCE: a call the loader must never make
    evil = b"\x80\x02cbuiltins\nprint\nX\x05\x00\x00\x00hello\x85R."
    f = str(tmp_path / "evil.jpc")
    with open(f, "wb") as fh:
        fh.write(evil)
    with pytest.raises(pickle.UnpicklingError):
        juice.io.load(f)
    with pytest.raises(ValueError):
        juice.io.load(str(tmp_path / "x.pkl"))
