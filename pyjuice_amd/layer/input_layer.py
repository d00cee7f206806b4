"""Input layer: compile-time bookkeeping (``pyjuice/layer/input_layer.py:33-255``) and the
device-side forward / flow backward / mini-batch EM / sampling through the HIP library.
"""
from __future__ import annotations

from copy import deepcopy
from functools import reduce
from typing import Dict, Optional, Sequence

import numpy as np
import torch
import torch.nn as nn

from pyjuice_amd import _native as N
from pyjuice_amd.utils import BitSet
from .layer import Layer

__all__ = ["InputLayer"]


def _apply_missing_mask(data_vb, missing_mask, batch_first=True):
    from pyjuice_amd.model.tensorcircuit import apply_missing_mask
    apply_missing_mask(data_vb, missing_mask, batch_first)


class InputLayer(Layer, nn.Module):
    def __init__(self, nodes: Sequence, cum_nodes: int = 0, pc_num_vars: int = 0,
                 max_tied_ns_per_parflow_block: int = 4):
        assert len(nodes) == len(set(nodes)), "a vector appears twice in one layer"
        nodes = self._reorder_nodes(nodes)
        nn.Module.__init__(self)
        Layer.__init__(self, nodes, disable_block_size_check=True)

        self.pc_num_vars = pc_num_vars
        scope = BitSet()
        for ns in self.nodes:
            scope = scope | ns.scope
        self.layer_scope = scope
        self.layer_var_list = torch.tensor(scope.to_list()).long()

        dist0 = self.nodes[0].dist
        sig = dist0.get_signature()
        layer_num_nodes = 0
        cum_params = cum_pflows = cum_source = 0
        tied = dict()
        for ns in self.nodes:
            assert ns.dist.get_signature() == sig, "an input layer holds one distribution type"
            ns._output_ind_range = (cum_nodes, cum_nodes + ns.num_nodes)
            cum_nodes += ns.num_nodes
            layer_num_nodes += ns.num_nodes
            npar = ns.num_nodes * ns.dist.num_parameters()
            npf = ns.num_nodes * ns.dist.num_param_flows()
            if not ns.is_tied():
                ns._param_range = (cum_params, cum_params + npar)
                cum_params += npar
                ns._param_flow_range = (cum_pflows, cum_pflows + npf)
                cum_pflows += npf
                cum_source += ns.num_nodes
            else:
                # input_layer.py:86-102: a tied vector shares its source's flow range until
                # `max_tied_ns_per_parflow_block` vectors use it, then opens a new range.
                src = ns.get_source_ns()
                ns._param_range = deepcopy(src._param_range)
                if src not in tied:
                    tied[src] = [[src], 1, src._param_flow_range]
                if tied[src][1] >= max_tied_ns_per_parflow_block:
                    ns._param_flow_range = (cum_pflows, cum_pflows + npf)
                    cum_pflows += npf
                    tied[src][2] = ns._param_flow_range
                    tied[src][0].append(ns)
                    tied[src][1] = 1
                else:
                    ns._param_flow_range = deepcopy(tied[src][2])
                    tied[src][1] += 1

        self._output_ind_range = (cum_nodes - layer_num_nodes, cum_nodes)
        self.num_parameters = cum_params
        self.num_param_flows = cum_pflows
        self.num_nodes = layer_num_nodes
        self.dist_signature = sig

        nvars = len(self.nodes[0].scope)
        vids = np.empty([layer_num_nodes, nvars], dtype=np.int64)
        s_pids = np.empty(layer_num_nodes, dtype=np.int64)
        s_pfids = np.empty(layer_num_nodes, dtype=np.int64)
        s_mids = np.empty(layer_num_nodes, dtype=np.int64)
        lnids = np.zeros(layer_num_nodes, dtype=np.int64)
        source_nids = np.empty(cum_source, dtype=np.int64)
        node_id2source_id = np.empty(layer_num_nodes, dtype=np.int64)
        metadata = []
        ncats = np.zeros(layer_num_nodes, dtype=np.int64)
        s, ss = 0, 0
        var2count = dict()
        ns2range = dict()
        for ns in self.nodes:
            e = s + ns.num_nodes
            ns2range[ns] = (s, e)
            vl = ns.scope.to_list()
            assert len(vl) == nvars
            vids[s:e, :] = np.asarray(vl, dtype=np.int64)[None, :]
            npar, npf = ns.dist.num_parameters(), ns.dist.num_param_flows()
            s_pids[s:e] = ns._param_range[0] + np.arange(ns.num_nodes) * npar if npar > 0 else 0
            s_pfids[s:e] = ns._param_flow_range[0] + np.arange(ns.num_nodes) * npf if npf > 0 else 0
            if not ns.is_tied():
                source_nids[ss:ss + ns.num_nodes] = np.arange(s, e)
                ss += ns.num_nodes
                node_id2source_id[s:e] = -1
            else:
                ts, te = ns2range[ns._source_node]
                node_id2source_id[s:e] = np.arange(ts, te)
            s_mids[s:e] = len(metadata)
            md = ns.dist.get_metadata()
            metadata.extend(md)
            ncats[s:e] = md[0] if len(md) > 0 else 2
            if nvars == 1:
                v = vl[0]
                c = var2count.get(v, 0)
                lnids[s:e] = np.arange(ns.num_nodes) + c
                var2count[v] = c + ns.num_nodes
            s = e

        self.vids = torch.from_numpy(vids)
        self.s_pids = torch.from_numpy(s_pids)
        self.s_pfids = torch.from_numpy(s_pfids)
        self.metadata = torch.tensor(metadata, dtype=torch.float32)
        self.s_mids = torch.from_numpy(s_mids)
        self.nids = torch.from_numpy(lnids)
        self.source_nids = torch.from_numpy(source_nids)
        self.node_id2source_id = torch.from_numpy(node_id2source_id)
        self._ncats = torch.from_numpy(ncats)

        # input_layer.py:219-231: [target flow start, #flows, start of every copy]
        self.tied2source_nids = []
        for src, item in tied.items():
            if len(item[0]) > 1:
                n = src._param_flow_range[1] - src._param_flow_range[0]
                ch = torch.tensor([c._param_flow_range[0] for c in item[0]], dtype=torch.long)
                self.tied2source_nids.append([src._param_flow_range[0], n, ch])

        self.params = nn.Parameter(torch.empty([max(self.num_parameters, 1)], dtype=torch.float32),
                                   requires_grad=False)
        self.num_vars_per_node = nvars
        self._pf, self._pf_fresh = None, False
        self.pflow_dtype = torch.float32
        self._used_external_params = False
        self._param_batch_size = 1
        self.n_block_size = self._get_num_nodes_gcd()
        self._dev = None

    # ------------------------------------------------------------------ bookkeeping

    def _reorder_nodes(self, nodes):
        """Untied vectors first (in order, pulling in sources of tied vectors), then tied ones
        (``input_layer.py:950-975``)."""
        node_set = set(nodes)
        untied, tied_l, added = [], [], set()
        for ns in nodes:
            if ns in added:
                continue
            if not ns.is_tied():
                untied.append(ns)
                added.add(ns)
            else:
                src = ns.get_source_ns()
                if src in added:
                    tied_l.append(ns)
                    added.add(ns)
                elif src in node_set:
                    untied.append(src)
                    tied_l.append(ns)
                    added.add(ns)
                    added.add(src)
                else:
                    raise ValueError("a tied input vector must share a layer with its source")
        out = untied + tied_l
        assert len(out) == len(nodes), "reordering the input vectors changed the node count"
        return out

    def _get_num_nodes_gcd(self):
        v = self.vids.numpy()
        if v.shape[0] == 0:
            return 0
        key = (v * (int(v.max()) + 1) ** np.arange(v.shape[1])[None, :]).sum(axis=1)
        change = np.nonzero(np.diff(key))[0]
        bounds = np.concatenate([[0], change + 1, [key.shape[0]]])
        counts = np.diff(bounds)
        return int(reduce(np.gcd, counts.tolist()))

    def _init_parameters(self, perturbation):
        """Copy DAG parameters (or draw them) into the flat vector (``input_layer.py:977-990``)."""
        p0 = 0
        with torch.no_grad():
            for ns in self.nodes:
                if ns.is_tied():
                    continue
                p1 = p0 + ns.num_nodes * ns.dist.num_parameters()
                if ns.has_params():
                    self.params[p0:p1] = ns._params.to(self.params.device)
                else:
                    self.params[p0:p1] = ns.init_parameters(ret_params=True).to(self.params.device)
                p0 = p1

    def get_dist(self):
        return self.nodes[0].dist

    def get_data_dtype(self):
        return self.nodes[0].get_data_dtype()

    def is_input(self):
        return True

    def update_parameters(self):
        for ns in self.nodes:
            if ns.is_tied():
                continue
            a, b = ns._param_range
            ns._params = self.params.data[a:b].detach().cpu().clone()

    def __repr__(self):
        return f"InputLayer[{self._output_ind_range[0]}:{self._output_ind_range[1]}, {len(self.layer_scope)} vars]"

    # ------------------------------------------------------------------ device state

    def to(self, device):
        nn.Module.to(self, device)
        device = torch.device(device)
        for name in ("vids", "s_pids", "s_pfids", "metadata", "s_mids", "nids", "source_nids",
                     "node_id2source_id", "layer_var_list", "_ncats"):
            setattr(self, name, getattr(self, name).to(device))
        for item in self.tied2source_nids:
            item[2] = item[2].to(device)
        self.device = device
        self._dev = None
        if device.type == "cuda":
            self._build_device_tables()
        return self

    def _build_device_tables(self):
        d = self.device
        src = self.source_nids.long()
        self._dev = {
            "vids": self.vids[:, 0].to(torch.int32).contiguous(),
            "s_pids": self.s_pids.to(torch.int64).contiguous(),
            "s_pfids": self.s_pfids.to(torch.int64).contiguous(),
            "ncats": self._ncats.to(torch.int32).contiguous(),
            "src": src.to(torch.int32).contiguous(),
            "tied": [(int(a), int(n), c.to(torch.int64).contiguous()) for a, n, c in self.tied2source_nids],
            "kmax": int(self._ncats.max().item()) if self.num_nodes > 0 else 1,
            # Several vectors accumulate into one flow range when tied vectors share it.
            "shared_pf": any(ns.is_tied() for ns in self.nodes),
            "sorted_tiles": self._uniform_tiles(8),
            "tied_groups": self._tied_groups(),
        }
        assert self.num_vars_per_node == 1, "multivariate input distributions are not supported by the HIP layers"

    # largest category count of the tied-range backward kernel (its LDS histogram)
    tied_max_cats = 16384

    def _tied_groups(self):
        """Groups of nodes sharing one parameter-flow range, for the tied-range backward
        (``pjamd_input_backward_tied``): ``(offsets, nodes, n_groups, covers)`` as int32 device
        tensors, ``covers`` True when the ranges tile the whole accumulator (then a zeroed
        accumulator may be stored into).  None unless the layer is Categorical with tied vectors
        and every group has one category count within ``tied_max_cats``."""
        if self.num_nodes == 0 or self._kind() != 0 or not any(ns.is_tied() for ns in self.nodes):
            return None
        s = self.s_pfids.detach().cpu().numpy().astype(np.int64)
        k = self._ncats.detach().cpu().numpy().astype(np.int64)
        order = np.argsort(s, kind="stable")
        bases, first, counts = np.unique(s[order], return_index=True, return_counts=True)
        ks = k[order]
        if int(k.max()) > self.tied_max_cats or np.any(ks != np.repeat(ks[first], counts)):
            return None
        if np.any(bases[1:] < bases[:-1] + ks[first][:-1]):
            return None   # ranges overlap: not a partition of the slots
        offsets = np.concatenate([first, [len(s)]]).astype(np.int32)
        covers = int(ks[first].sum()) == self.num_param_flows
        d = self.device
        return (torch.from_numpy(offsets).to(d), torch.from_numpy(order.astype(np.int32)).to(d), len(bases), covers)

    def _uniform_tiles(self, tile: int) -> bool:
        """True if every tile of ``tile`` consecutive nodes (from node 0) has one variable and one
        category count (the backward kernel then sorts each tile's batch once)."""
        v = self.vids[:, 0].detach().cpu().numpy()
        k = self._ncats.detach().cpu().numpy()
        n = v.shape[0]
        if n == 0:
            return False
        pad = (-n) % tile
        vp = np.concatenate([v, np.repeat(v[-1:], pad)]).reshape(-1, tile)
        kp = np.concatenate([k, np.repeat(k[-1:], pad)]).reshape(-1, tile)
        return bool(np.all(vp == vp[:, :1]) and np.all(kp == kp[:, :1]))

    def _kind(self):
        return self.nodes[0].dist.native_kind

    @property
    def param_flows(self):
        """Parameter-flow accumulator of the layer.  After ``init_param_flows(0.0)`` it is only
        marked zero (the next backward stores instead of adding); reading it first clears it."""
        if self._pf_fresh and self._pf is not None:
            self._pf.zero_()
            self._pf_fresh = False
        return self._pf

    @param_flows.setter
    def param_flows(self, value):
        self._pf = value
        self._pf_fresh = False

    def init_param_flows(self, flows_memory: float = 1.0, dtype: Optional[torch.dtype] = None):
        dtype = dtype if dtype is not None else self.pflow_dtype
        if self._pf is None or self._pf.dtype != dtype or self._pf.device != self.params.device:
            self.param_flows = torch.zeros([self.num_param_flows], device=self.params.device, dtype=dtype)
            self.pflow_dtype = dtype
        elif flows_memory < 1.0:
            if flows_memory == 0.0:
                self._pf_fresh = True     # cleared lazily (see the param_flows property)
            else:
                self.param_flows.mul_(flows_memory)
        return None

    # ------------------------------------------------------------------ passes

    def forward(self, data: torch.Tensor, node_mars: torch.Tensor, params: Optional[Dict] = None,
                missing_mask: Optional[torch.Tensor] = None, **kwargs):
        """``node_mars[off + i, b] = log p_i(x[v_i, b])`` (``input_layer.py:283-404``).
        ``data`` is ``[num_vars, B]`` int32 (see ``TensorCircuit.forward``)."""
        self._used_external_params = params is not None
        p = self.params if params is None else params["params"]
        assert node_mars.is_cuda, "InputLayer.forward runs on the HIP device only"
        if missing_mask is not None:
            # marginalized entries (log p = 0) are marked -1 in a copy of the data
            data = data.clone()
            _apply_missing_mask(data, missing_mask, kwargs.get("_batch_first", True))
        t = self._dev
        B = node_mars.size(1)
        N.call("pjamd_input_forward", self._kind(), N.ptr(p), N.ptr(node_mars), N.ptr(data), N.ptr(t["vids"]),
               N.ptr(t["s_pids"]), N.ptr(t["ncats"]), self.num_nodes, self._output_ind_range[0], B, t["kmax"],
               N.stream(), algo=(4 * (self.num_nodes + len(self.layer_scope)) * B + 4 * self.num_parameters, 0))

    def backward(self, data: torch.Tensor, node_flows: torch.Tensor, node_mars: torch.Tensor,
                 params: Optional[Dict] = None, logspace_flows: bool = False,
                 missing_mask: Optional[torch.Tensor] = None, marginalized: bool = False,
                 fused_em: Optional[tuple] = None, **kwargs):
        """``param_flows[s_pfid_i + x] += node_flows[off + i, b]`` (``input_layer.py:406-571``).
        Marginalized entries (``missing_mask``, or -1 in ``data`` with ``marginalized``) add their
        flows to every category in proportion to ``theta`` (``categorical.py:103-134``).
        ``fused_em = (step_size, pseudocount, keep_zero_params)``: a training step's EM follows;
        where the layer allows (``fused_em_ok``) the flows are stored and the parameters updated in
        the same pass (``pjamd_input_backward_em``), and the next ``mini_batch_em`` skips the layer."""
        assert not logspace_flows, "logspace flows are not supported by the HIP input layers"
        if missing_mask is not None:
            data = data.clone()
            _apply_missing_mask(data, missing_mask, kwargs.get("_batch_first", True))
            marginalized = True
        if self._pf is None:
            self.init_param_flows(0.0)
        t = self._dev
        B = node_flows.size(1)
        # store instead of add into a buffer marked zero when every slot has one owner
        tg = t["tied_groups"]
        overwrite = self._pf_fresh and (not t["shared_pf"] or (tg is not None and tg[3]))
        pf = self._pf if overwrite else self.param_flows
        pf64 = 1 if pf.dtype == torch.float64 else 0
        if (fused_em is not None and overwrite and not marginalized and params is None
                and not self._used_external_params and self.fused_em_ok(B, self._output_ind_range[0])):
            step, pcount, keep_zero = fused_em
            self._pf_fresh = False
            # node flows read, slots stored, parameters read and written
            algo = (4 * (self.num_nodes + len(self.layer_scope)) * B
                    + pf.element_size() * self.num_param_flows + 8 * self.num_parameters, 0)
            N.call("pjamd_input_backward_em", N.ptr(pf), N.ptr(node_flows), pf64, N.ptr(data), N.ptr(t["vids"]),
                   N.ptr(t["s_pfids"]), N.ptr(t["ncats"]), self.num_nodes, self._output_ind_range[0], B, t["kmax"],
                   N.ptr(self.params), N.ptr(t["s_pids"]), float(step), float(pcount), 1 if keep_zero else 0,
                   N.stream(), algo=algo)
            self._em_done = True
            return
        self._pf_fresh = False
        algo = (4 * (self.num_nodes + len(self.layer_scope)) * B
                + (1 if overwrite else 2) * pf.element_size() * self.num_param_flows, 0)
        if tg is not None:
            N.call("pjamd_input_backward_tied", self._kind(), N.ptr(pf), N.ptr(node_flows), pf64, N.ptr(data),
                   N.ptr(t["vids"]), N.ptr(t["s_pfids"]), N.ptr(t["ncats"]), N.ptr(tg[0]), N.ptr(tg[1]), tg[2],
                   t["kmax"], self._output_ind_range[0], B, 1 if overwrite else 0, N.stream(), algo=algo)
        else:
            flags = (t["kmax"] << 3) | (4 if overwrite else 0) | (2 if t["sorted_tiles"] else 0) | \
                (1 if t["shared_pf"] else 0)
            N.call("pjamd_input_backward", self._kind(), N.ptr(pf), N.ptr(node_flows), pf64, N.ptr(data),
                   N.ptr(t["vids"]), N.ptr(t["s_pfids"]), N.ptr(t["ncats"]), self.num_nodes,
                   self._output_ind_range[0], B, flags, N.stream(), algo=algo)
        if marginalized:
            N.call("pjamd_input_missing_flows", self._kind(), N.ptr(pf), pf64, N.ptr(node_flows), N.ptr(self.params),
                   N.ptr(data), N.ptr(t["vids"]), N.ptr(t["s_pids"]), N.ptr(t["s_pfids"]), N.ptr(t["ncats"]),
                   self.num_nodes, self._output_ind_range[0], B, N.stream(),
                   algo=(8 * self.num_nodes * B + 2 * pf.element_size() * self.num_param_flows, 0))

    def fused_em_ok(self, B: int, node_offset: int) -> bool:
        """Whether ``pjamd_input_backward_em`` takes this layer at batch ``B``: Categorical, every
        node its own EM source with its own slots, tiles of 8 nodes on one variable and category
        count, B % 4 == 0 and <= 2048, at most 1024 categories."""
        t = self._dev
        return (self.num_nodes > 0 and self._kind() == 0 and not t["shared_pf"] and t["tied_groups"] is None
                and not t["tied"] and t["src"].numel() == self.num_nodes and t["sorted_tiles"]
                and B % 4 == 0 and B <= 2048 and (node_offset * B) % 4 == 0 and t["kmax"] <= 1024)

    def mini_batch_em(self, step_size: float, pseudocount: float = 0.0, keep_zero_params: bool = False):
        """Tied-flow fusing then the closed-form EM update (``input_layer.py:663-724``); nothing
        for a layer whose update ran inside the backward (``fused_em``)."""
        if self.__dict__.pop("_em_done", False):
            return None
        if self._used_external_params:
            return None
        if self.param_flows is None:
            self.init_param_flows(0.0)
        t = self._dev
        pf64 = 1 if self.param_flows.dtype == torch.float64 else 0
        with torch.no_grad():
            for (start, n, ch) in t["tied"]:
                N.call("pjamd_pflow_fuse", N.ptr(self.param_flows), pf64, start, n, N.ptr(ch), ch.numel(), N.stream())
            N.call("pjamd_input_em", self._kind(), N.ptr(self.params), N.ptr(self.param_flows), pf64, N.ptr(t["src"]),
                   N.ptr(t["s_pids"]), N.ptr(t["s_pfids"]), N.ptr(t["ncats"]), t["src"].numel(),
                   float(step_size), float(pseudocount), 1 if keep_zero_params else 0, N.stream(),
                   # the source vectors' parameters read and written, their (fused) flows read
                   algo=(8 * self.num_parameters
                         + self.param_flows.element_size() * self.num_parameters * (2 if self._kind() == 1 else 1), 0))

    def sample(self, samples: torch.Tensor, node_active: torch.Tensor, uniforms: Optional[torch.Tensor],
               seed: Optional[torch.Tensor] = None, cdf: Optional[torch.Tensor] = None, **kwargs):
        """Draw ``samples[v_i, s]`` for every active input node ``i`` (clearing its flag) with
        uniform ``uniforms[off + i, s]``, or the counter-based uniform of (off + i, s) under the
        device key ``seed`` when ``uniforms`` is None: Categorical picks the first category whose
        running sum reaches ``u`` (``categorical.py:137-161``), Bernoulli returns ``u < theta``
        (``bernoulli.py:79-88``).  ``cdf``: the parameters' running sums (``pjamd_sample_prefix``)."""
        t = self._dev
        S = samples.size(1)
        assert uniforms is not None or seed is not None
        N.call("pjamd_input_sample", self._kind(), N.ptr(samples), N.ptr(self.params),
               N.ptr(cdf) if cdf is not None else None, N.ptr(node_active),
               N.ptr(uniforms) if uniforms is not None else None, N.ptr(seed) if seed is not None else None,
               N.ptr(t["vids"]), N.ptr(t["s_pids"]), N.ptr(t["ncats"]),
               1 if samples.dtype == torch.int64 else 0, self.num_nodes, self._output_ind_range[0], S, N.stream(),
               algo=(self.num_nodes * S, 0))
