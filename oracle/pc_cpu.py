"""Driver of ``pc_cpu.c``: forward, flow backward and mini-batch EM of a compiled circuit on the
CPU, on the compiler's own index tables.  TEST INFRASTRUCTURE ONLY (tests, ``smoke()`` and the
CPU-baseline leg of ``bench.py``).

The orchestration follows ``TensorCircuit.forward / backward / mini_batch_em`` of the reference
(``pyjuice/model/tensorcircuit.py:207-342, 344-538, 549-590``): input layers, then the inner
layer groups in order; flows top-down with unit root flows; EM = input layers (after fusing tied
flows), then tied sum-flow fusing, then the sum parameters.  Node flows stay unmodified (the
reference's ``allow_modify_flows=False`` form).
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "_bin")

_p, _i, _l, _d = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
_SIGS = {
    "pcc_input_forward": [_i, _p, _p, _p, _p, _p, _p, _l, _l, _l],
    "pcc_input_backward": [_i, _p, _p, _p, _p, _p, _p, _l, _l, _l],
    "pcc_input_em": [_i, _p, _p, _p, _p, _p, _p, _l, _d, _d, _i],
    "pcc_prod": [_p, _p, _p, _p, _l, _l, _l, _l, _i],
    "pcc_sum_forward": [_p, _p, _p, _p, _p, _p, _p, _l, _l, _l, _l],
    "pcc_sum_ele": [_p, _p, _p, _p, _p, _p, _p, _p, _p, _l, _l, _l, _l, _l, _i],
    "pcc_sum_par": [_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _l, _l, _l, _l],
    "pcc_sum_em": [_p, _p, _p, _l, _d, _d, _i],
    "pcc_pflow_fuse": [_p, _l, _l, _p, _l],
}
_libs = dict()


def lib(dtype=np.float64):
    name = "libpc_cpu64.so" if dtype == np.float64 else "libpc_cpu32.so"
    if name not in _libs:
        path = os.path.join(BIN, name)
        if not os.path.exists(path):
            from oracle import build_cpu
            build_cpu.build()
        L = ctypes.CDLL(path)
        for fn, args in _SIGS.items():
            getattr(L, fn).argtypes = args
            getattr(L, fn).restype = None
        _libs[name] = L
    return _libs[name]


def _P(a: np.ndarray, off_rows: int = 0, row_len: int = 0):
    """Pointer to ``a`` (contiguous) advanced by ``off_rows`` rows of ``row_len`` elements."""
    return ctypes.c_void_p(a.ctypes.data + off_rows * row_len * a.itemsize)


def _i64(t) -> np.ndarray:
    return np.ascontiguousarray(t.detach().cpu().numpy().astype(np.int64))


class CompiledCPU:
    """A compiled ``TensorCircuit``'s tables and parameters, evaluated by ``pc_cpu.c``."""

    def __init__(self, pc, dtype=np.float64):
        self.dtype = dtype
        self.L = lib(dtype)
        self.num_nodes = pc.num_nodes
        self.num_elements = pc.num_elements
        self.root = pc._root_node_range
        self.elem_offsets = list(pc._elem_offsets)
        self.num_element_rows = pc.num_element_rows
        self.num_dummy_eles = pc.num_dummy_eles
        self.params = np.ascontiguousarray(pc.params.detach().cpu().numpy().astype(dtype))
        self.num_param_flows = pc.num_param_flows
        self.inputs = []
        for layer in pc.input_layer_group:
            self.inputs.append(dict(
                kind=layer.nodes[0].dist.native_kind,
                params=np.ascontiguousarray(layer.params.detach().cpu().numpy().astype(dtype)),
                vids=_i64(layer.vids[:, 0]), s_pids=_i64(layer.s_pids), s_pfids=_i64(layer.s_pfids),
                ncats=_i64(layer._ncats), src=_i64(layer.source_nids), off=layer._output_ind_range[0],
                n=layer.num_nodes, npf=layer.num_param_flows,
                tied=[(int(a), int(n), _i64(c)) for a, n, c in layer.tied2source_nids]))
        self.groups = []
        for gid, g in enumerate(pc.inner_layer_groups):
            layers = []
            for layer in g:
                if g.is_prod():
                    fw = [(_i64(n), _i64(c)) for n, c in zip(layer.partitioned_nids, layer.partitioned_cids)
                          if n.numel() > 0]
                    bk = [(_i64(n), _i64(c)) for n, c in zip(layer.partitioned_u_cids, layer.partitioned_parids)
                          if n.numel() > 0]
                    layers.append(dict(bs=layer.block_size, fw=fw, bk=bk))
                else:
                    fw = [(_i64(n), _i64(c), _i64(p), _i64(f), _i64(k)) for n, c, p, f, k in
                          zip(layer.partitioned_nids, layer.partitioned_cids, layer.partitioned_pids,
                              layer.partitioned_pfids, layer._fw_n_chs) if n.numel() > 0]
                    bk = [(_i64(c), _i64(p), _i64(pp), _i64(k), int(cbs)) for c, p, pp, k, cbs in
                          zip(layer.partitioned_chids, layer.partitioned_parids, layer.partitioned_parpids,
                              layer._bk_n_pars, layer.cs_block_sizes) if c.numel() > 0]
                    layers.append(dict(bs=layer.block_size, fw=fw, bk=bk))
            self.groups.append(("prod" if g.is_prod() else "sum", gid // 2, layers))
        self.node_blocks = _i64(pc.par_update_kwargs[7]["node_blocks"])
        self.fuse = [(int(t), int(n), np.asarray(ch, dtype=np.int64)) for t, n, ch in pc.parflow_fusing_kwargs]
        self.param_flows = np.zeros([self.num_param_flows], dtype=dtype)
        for d in self.inputs:
            d["pflows"] = np.zeros([max(d["npf"], 1)], dtype=dtype)

    # ------------------------------------------------------------------ passes

    def forward(self, x: np.ndarray) -> np.ndarray:
        """Log-likelihoods ``[B]`` (first root node) of integer data ``x [B, num_vars]``."""
        L, dt = self.L, self.dtype
        B = x.shape[0]
        self.B = B
        self.data = np.ascontiguousarray(x.T.astype(np.int32))
        self.node_mars = np.zeros([self.num_nodes, B], dtype=dt)
        self.element_mars = np.empty([self.num_element_rows, B], dtype=dt)
        for off in self.elem_offsets:
            self.element_mars[off:off + self.num_dummy_eles] = -np.inf
        for d in self.inputs:
            L.pcc_input_forward(d["kind"], _P(d["params"]), _P(self.node_mars), _P(self.data), _P(d["vids"]),
                                _P(d["s_pids"]), _P(d["ncats"]), d["n"], d["off"], B)
        for kind, slab, layers in self.groups:
            E = _P(self.element_mars, self.elem_offsets[slab], B)
            for ly in layers:
                for part in ly["fw"]:
                    if kind == "prod":
                        nids, cids = part
                        L.pcc_prod(E, _P(self.node_mars), _P(nids), _P(cids), nids.size, cids.shape[1], ly["bs"],
                                   B, 0)
                    else:
                        nids, cids, pids, _, nch = part
                        L.pcc_sum_forward(_P(self.node_mars), E, _P(self.params), _P(nids), _P(cids), _P(pids),
                                          _P(nch), nids.size, cids.shape[1], ly["bs"], B)
        return self.node_mars[self.root[0]:self.root[1]].copy()

    def backward(self, root_flows: Optional[np.ndarray] = None):
        """Node flows and accumulated parameter flows (sum and input) for unit root flows."""
        L, dt, B = self.L, self.dtype, self.B
        self.node_flows = np.zeros([self.num_nodes, B], dtype=dt)
        self.node_flows[self.root[0]:self.root[1]] = 1.0 if root_flows is None else root_flows
        eflows = np.zeros([self.num_elements, B], dtype=dt)
        for kind, slab, layers in reversed(self.groups):
            E = _P(self.element_mars, self.elem_offsets[slab], B)
            for ly in layers:
                if kind == "prod":
                    for u, parids in ly["bk"]:
                        L.pcc_prod(_P(self.node_flows), _P(eflows), _P(u), _P(parids), u.size, parids.shape[1],
                                   ly["bs"], B, 1)
                    continue
                eflows[:] = 0.0
                for chids, parids, parpids, npar, cbs in ly["bk"]:
                    L.pcc_sum_ele(_P(eflows), _P(self.node_flows), _P(self.node_mars), E, _P(self.params),
                                  _P(chids), _P(parids), _P(parpids), _P(npar), chids.size, parids.shape[1], cbs,
                                  ly["bs"], B, 1)
                for nids, cids, pids, pfids, nch in ly["fw"]:
                    L.pcc_sum_par(_P(self.param_flows), _P(self.node_flows), _P(self.node_mars), E,
                                  _P(self.params), _P(nids), _P(cids), _P(pids), _P(pfids), _P(nch), nids.size,
                                  cids.shape[1], ly["bs"], B)
        for d in self.inputs:
            L.pcc_input_backward(d["kind"], _P(d["pflows"]), _P(self.node_flows), _P(self.data), _P(d["vids"]),
                                 _P(d["s_pfids"]), _P(d["ncats"]), d["n"], d["off"], B)

    def zero_param_flows(self):
        self.param_flows[:] = 0.0
        for d in self.inputs:
            d["pflows"][:] = 0.0

    def mini_batch_em(self, step_size: float, pseudocount: float = 0.0, keep_zero_params: bool = False):
        L = self.L
        for d in self.inputs:
            for t, n, ch in d["tied"]:
                L.pcc_pflow_fuse(_P(d["pflows"]), t, n, _P(ch), ch.size)
            L.pcc_input_em(d["kind"], _P(d["params"]), _P(d["pflows"]), _P(d["src"]), _P(d["s_pids"]),
                           _P(d["s_pfids"]), _P(d["ncats"]), d["src"].size, step_size, pseudocount,
                           int(keep_zero_params and d["kind"] == 0))
        for t, n, ch in self.fuse:
            L.pcc_pflow_fuse(_P(self.param_flows), t, n, _P(ch), ch.size)
        if self.node_blocks.size > 0:
            L.pcc_sum_em(_P(self.params), _P(self.param_flows), _P(self.node_blocks), self.node_blocks.shape[0],
                         step_size, pseudocount, int(keep_zero_params))
