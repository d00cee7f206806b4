"""Compiled probabilistic circuit (``pyjuice/model/tensorcircuit.py``).

``TensorCircuit`` keeps the reference's module API (``forward``, ``backward``,
``mini_batch_em``, ``init_param_flows``, ``update_parameters`` ...) and buffer layout
(``node_mars`` / ``node_flows`` ``[num_nodes, B]``, ``element_mars`` / ``element_flows``
``[num_elements, B]``, flat ``params``).  Every pass is a sequence of HIP kernel launches on
the current stream; ``record_cudagraph=True`` captures the inner layers into a HIP graph.

Precision: ``precision`` selects the arithmetic of the sum-layer contractions
(``"fp32"``, ``"bf16"`` or ``"fp16"``; the latter two on MFMA with FP32 accumulation).
Parameter flows are accumulated across mini-batches in ``pflow_dtype`` (FP64 by default).
"""
from __future__ import annotations

import weakref
from functools import partial
from typing import Callable, Optional, Sequence, Tuple, Union

import numpy as np
import torch
import torch.nn as nn

from pyjuice_amd import _native as N
from pyjuice_amd.layer.sum_layer import PRECISIONS
from pyjuice_amd.layer import InputLayer, LayerGroup, ProdLayer, SumLayer
from pyjuice_amd.nodes import CircuitNodes, multiply, summate

# int32 fields per row of the batched parameter-flow descriptor table (PJAMD_PAR_DESC_FIELDS)
PAR_DESC_FIELDS = 9
from pyjuice_amd.nodes.nodes import dfs_postorder
from pyjuice_amd.utils import BitSet
from . import chain as _chain
from .backend import compile_cum_par_flows_fn, compile_par_update_fn, normalize_parameters

__all__ = ["TensorCircuit", "compile"]

_DTYPE_CODE = {torch.uint8: 0, torch.bool: 0, torch.int8: 3, torch.int32: 1, torch.int64: 2}


def _pc_model_backward_hook(grad, pc, inputs, record_cudagraph, apply_cudagraph, propagation_alg, **kwargs):
    """Autograd hook of ``lls``: runs the flow backward with ``ll_weights = grad / sum(grad) * B``
    (``tensorcircuit.py:24-37``)."""
    grad = grad.permute(1, 0)
    pc.backward(inputs=inputs, ll_weights=grad / grad.sum() * grad.size(1),
                compute_param_flows=pc._optim_hyperparams["compute_param_flows"],
                flows_memory=pc._optim_hyperparams["flows_memory"], record_cudagraph=record_cudagraph,
                apply_cudagraph=apply_cudagraph, propagation_alg=propagation_alg, **kwargs)
    return None


def layer_iterator(pc, reverse=False, ret_layer_groups=False, ignore_input_layers=False):
    groups = list(pc.inner_layer_groups)
    if reverse:
        groups = groups[::-1]
    seq = []
    if not reverse and not ignore_input_layers:
        seq.append(pc.input_layer_group)
    seq.extend(groups)
    if reverse and not ignore_input_layers:
        seq.append(pc.input_layer_group)
    for g in seq:
        if ret_layer_groups:
            yield g
        else:
            for layer in g:
                yield layer


def apply_missing_mask(data_vb: torch.Tensor, missing_mask: torch.Tensor, batch_first: bool = True):
    """Marks the entries of ``missing_mask`` in the ``[V, B]`` int32 device data as marginalized
    (-1) in place.  ``missing_mask``: ``[V]``, ``[B, V]``, or ``[V, B]`` when ``batch_first`` is
    False or the shape says so (the reference's ``_batch_first``, ``input_layer.py:378-383``)."""
    V, B = data_vb.shape
    m = missing_mask.to(device=data_vb.device, dtype=torch.uint8).contiguous()
    if m.dim() == 1:
        assert m.numel() == V, "a 1-D missing_mask has one entry per variable"
        mode = 0
    else:
        assert m.dim() == 2
        if tuple(m.shape) == (B, V) and (batch_first or (B, V) != (V, B)):
            mode = 1
        else:
            assert tuple(m.shape) == (V, B), "missing_mask must be [V], [B, V] or [V, B]"
            mode = 2
    N.call("pjamd_data_mask", N.ptr(data_vb), N.ptr(m), mode, B, V, N.stream())


class TensorCircuit(nn.Module):
    def __init__(self, root_ns: CircuitNodes, layer_sparsity_tol: float = 0.5,
                 max_num_partitions: Optional[int] = None, disable_gpu_compilation: bool = False,
                 force_gpu_compilation: bool = False, max_tied_ns_per_parflow_block: int = 8,
                 device: Optional[Union[int, torch.device]] = None, verbose: bool = True,
                 precision: str = "fp32", pflow_dtype: torch.dtype = torch.float64):
        super().__init__()
        assert isinstance(root_ns, CircuitNodes), "compile() takes the root CircuitNodes of a circuit"
        assert precision in ("fp32", "bf16", "fp16", "fp32_valu")
        self.root_ns = root_ns
        self.device = torch.device("cpu")
        self.num_vars = self._get_num_vars(root_ns)
        self.precision = precision
        self.pflow_dtype = pflow_dtype

        self.node_mars = None
        self.element_mars = None
        self.node_flows = None
        self.element_flows = None
        self._pf, self._pf_fresh, self._pf_overwrite_ok = None, False, False

        self._init_layers(layer_sparsity_tol=layer_sparsity_tol, max_num_partitions=max_num_partitions,
                          max_tied_ns_per_parflow_block=max_tied_ns_per_parflow_block, verbose=verbose)

        self._optim_hyperparams = {"compute_param_flows": True, "flows_memory": 1.0}
        self._recorded_cuda_graphs = dict()
        self.default_propagation_alg = "LL"
        self.propagation_alg_kwargs = dict()
        self._run_params = dict()
        self._cum_flow_host, self._cum_flow_dev = 0.0, None
        self._em_dev = None
        self.modify_flows_in_place = True
        self._par_batches = {"tables": dict(), "layers": set()}
        self._fused = {"tables": dict(), "prod_groups": set(), "premod": set()}
        # Runs of narrow levels in one launch per pass (pyjuice_amd.model.chain; BF16 / FP16
        # operands, B % 4 == 0): planned per batch size, with a packed copy of the runs'
        # parameters written at the start of every forward.
        self.use_chains = False
        self._chain_cache, self._chain_packed = dict(), dict()
        self._chain_packed = dict()
        # Layer-wise recomputation (the reference's own scheme, tensorcircuit.py:446-447): one
        # element buffer of the widest product group instead of one slab per group; the flow
        # backward recomputes each product group before the sum group above it, and parameter
        # flows are taken per layer inside the loop.  element_mars then takes max_g rows instead
        # of the sum over groups, at the cost of a second product pass.
        self.recompute_elements = False
        # Product layer groups computed inside the forward of the sum group above them
        # (pjamd_sum_forward_prod) where the layout allows it; BF16 / FP16 operands, B % 4 == 0,
        # sum launches below fuse_products_max_tiles workgroups.  Off by default: on HCLT-784
        # (256 latents, batch 512) the node block's row tiles gathering the product children
        # again cost more than the product launches they save (3.05 vs 3.02 ms per step).
        self.fuse_products = False
        self.fuse_products_max_tiles = 256
        self._prod_fusion = {"groups": set(), "tables": dict(), "tiles": dict()}
        self._data_vb, self._data_marginalized = None, False
        self._data_src = None   # (weak reference to the forward's input tensor, its version)

    # ------------------------------------------------------------------ parameter-flow buffer

    @property
    def _cum_flow(self) -> float:
        """Total root flow added to the parameter flows since the last rescaled EM step
        (``tensorcircuit.py:417-421``).  Weighted backward passes add their weight sums on the
        device; they are read back only here, so a backward never waits for the device."""
        if self._cum_flow_dev is not None:
            self._cum_flow_host += float(self._cum_flow_dev.item())
            self._cum_flow_dev = None
        return self._cum_flow_host

    @_cum_flow.setter
    def _cum_flow(self, value: float):
        self._cum_flow_host, self._cum_flow_dev = float(value), None

    @property
    def param_flows(self):
        """Sum parameter-flow accumulator.  After ``init_param_flows(0.0)`` the buffer is only
        marked zero (the next flow backward stores into the slots it owns instead of adding);
        reading it before that clears it."""
        if self._pf_fresh and self._pf is not None:
            self._pf.zero_()
            self._pf_fresh = False
        return self._pf

    @param_flows.setter
    def param_flows(self, value):
        self._pf = value
        self._pf_fresh = False

    # ------------------------------------------------------------------ device placement

    def to(self, device):
        if isinstance(device, int):
            device = torch.device(f"cuda:{device}")
        device = torch.device(device)
        if device.type == "cuda" and device.index is None:
            device = torch.device(f"cuda:{torch.cuda.current_device()}")
        super().to(device)
        self.input_layer_group.to(device)
        for g in self.inner_layer_groups:
            g.to(device)
        self.device = device
        self.node_mars = self.element_mars = self.node_flows = self.element_flows = None
        self.param_flows = None
        self._recorded_cuda_graphs = dict()
        if device.type == "cuda":
            N.lib()  # fail loudly now if the HIP library is missing
            meta = self.par_update_kwargs[7]
            blocks = meta["node_blocks"]
            assert blocks.numel() == 0 or int(blocks[:, :2].max()) + int((blocks[:, 3] * blocks[:, 4]).max()) < 2 ** 31
            self._em_dev = {
                "blocks": blocks.to(torch.int32).to(device).contiguous(),
                "max_bs": int(blocks[:, 3].max()) if blocks.numel() > 0 else 1,
                "fuse": [(int(t), int(n), torch.tensor(ch, dtype=torch.int64, device=device))
                         for t, n, ch in self.parflow_fusing_kwargs],
            }
            self._build_launch_plans(device)
        self._chain_cache, self._chain_packed = dict(), dict()
        return self

    def _build_launch_plans(self, device):
        """Launch tables of the whole circuit on ``device`` (built on the HIP device by ``to``; on the
        CPU only to inspect them): per-layer int32 tables, masked product backward tables, batched
        parameter-flow descriptors, fused product backward tables."""
        for g in self.inner_layer_groups:
            for layer in g:
                if layer._dev is None:
                    layer.build_device_tables()
        self._mask_unreached_products()
        self._par_batches = self._build_par_batches(device)
        self._fused = self._build_fused_backward(device)
        self._prod_fusion = self._build_prod_fusion(device)
        self._pf_overwrite_ok = self._sum_flows_overwritable()
        self._chain_cache, self._chain_packed = dict(), dict()

    def _chain_plan(self, precision: str, B: int):
        """Chain runs (``pyjuice_amd.model.chain.plan``) for batches of ``B``, cached per batch
        size; None where chains do not apply."""
        if not self.use_chains or self.recompute_elements or precision not in ("bf16", "fp16") or B % 4 != 0:
            return None
        plan = self._chain_cache.get(B)
        if plan is None:
            plan = _chain.plan(self, self.device, B)
            self._chain_cache[B] = plan
        return plan if plan.runs else None

    def _chain_pack(self, plan, precision: str):
        """The packed operand copy of the runs' parameters (``pjamd_chain_pack``), refreshed from
        ``params``; returns the buffer."""
        buf = self._chain_packed.get((precision, id(plan)))
        if buf is None:
            self._chain_packed = {k: v for k, v in self._chain_packed.items() if k[1] == id(plan)}
            dt = torch.bfloat16 if precision == "bf16" else torch.float16
            buf = torch.empty([plan.pack_units * 8], dtype=dt, device=self.device)
            self._chain_packed[(precision, id(plan))] = buf
        n_wg = plan.pack_wg.size(0)
        N.call("pjamd_chain_pack", N.ptr(buf), N.ptr(self.params), N.ptr(plan.pack_desc), N.ptr(plan.pack_pbase),
               N.ptr(plan.pack_wg), n_wg, PRECISIONS[precision], N.stream(),
               algo=(4 * plan.pack_units * 8 + 2 * plan.pack_units * 8, 0))
        return buf

    def _mask_unreached_products(self):
        """A product block that no sum edge of the layer above points to (sparse block
        connections, e.g. PD's) receives no flow, and the sum layer's element-flow pass does not
        write its rows of the shared element-flow buffer.  Its entries in the product layers'
        backward tables are set to padding (0), so the product backward adds nothing for it
        instead of reading rows left over from another layer."""
        groups = self.inner_layer_groups
        for gid, g in enumerate(groups):
            if not g.is_prod() or gid + 1 >= len(groups):
                continue
            covered = set()
            for sl in groups[gid + 1]:
                for chids in sl.partitioned_chids:
                    covered.update(chids.detach().cpu().numpy().astype(np.int64).tolist())
            for layer in g:
                if layer._dev is None:
                    continue
                bk = []
                for u, parids, words in layer._dev["bk"]:
                    p = parids.detach().cpu().numpy()
                    keep = np.isin(p, np.fromiter(covered, dtype=np.int64, count=len(covered))) | (p == 0)
                    if not keep.all():
                        p = np.where(keep, p, 0).astype(np.int32)
                        parids = torch.from_numpy(p).to(parids.device).contiguous()
                    bk.append((u, parids, words))
                layer._dev["bk"] = bk

    def _sum_flows_overwritable(self) -> bool:
        """True if one flow backward writes every sum parameter-flow slot exactly once with no
        atomics: no tied sharing inside or across layers, every slot owned by one (node, child)
        edge.  Then the first backward after the flows were cleared may store instead of add."""
        if any(t["desc"].view(-1, PAR_DESC_FIELDS)[:, 7].any().item() for t in self._par_batches["tables"].values()):
            return False
        cover = np.zeros(self.num_param_flows, dtype=np.int64)
        for g in self.inner_layer_groups:
            if g.is_prod():
                continue
            for layer in g:
                if layer._pf_shared:
                    return False
                for nids, cids, pids, pfids, nch, tot in layer._dev["fw"]:
                    f = pfids.detach().cpu().numpy().astype(np.int64)
                    n = nch.detach().cpu().numpy().astype(np.int64)
                    real = np.arange(f.shape[1])[None, :] < n[:, None]
                    slots = (f[real][:, None] + np.arange(layer.block_size)[None, :]).reshape(-1)
                    np.add.at(cover, slots, 1)
        return bool(np.all(cover == 1))

    # Product children a fused forward stage reads (pjamd_sum_forward_prod's prod_k choices).
    _PROD_K = (2, 4, 8)

    def _build_prod_fusion(self, device):
        """Product layer groups whose forward runs inside the next sum group's forward
        (``pjamd_sum_forward_prod``), with each sum table's ``prod_bases``.  Checked on the index
        tables: every forward table of the sum group takes the 16-byte run form (block size a
        multiple of 64, child runs of >= 128, aligned parameters); every stage of 128 children of
        a node block lies inside one product block of at most 8 child blocks; the stages cover
        every product row (so the fused forward stores the whole element slab the backward
        reads).  Used at run time only for BF16 / FP16 operands, B % 4 == 0 and sum launches of
        fewer than ``fuse_products_max_tiles`` workgroups: there a launch is a few dependent memory
        round trips and the product launch is saved whole, while on wide layers each of a node
        block's row tiles would gather the product children again."""
        groups = self.inner_layer_groups
        fused_groups, tables, tiles = set(), dict(), dict()
        for gid in range(len(groups) - 1):
            gp, gs = groups[gid], groups[gid + 1]
            if not gp.is_prod() or gs.is_prod():
                continue
            starts, sizes, kids = [], [], []
            for layer in gp:
                for nids, cids, _ in layer._dev["fw"]:
                    c = cids.detach().cpu().numpy().astype(np.int64)
                    for e, row in zip(nids.detach().cpu().numpy().astype(np.int64).tolist(), c):
                        starts.append(e)
                        sizes.append(layer.block_size)
                        kids.append(row[row != 0])
            if not starts or max(len(k) for k in kids) > self._PROD_K[-1]:
                continue
            order = np.argsort(np.asarray(starts))
            st = np.asarray(starts, dtype=np.int64)[order]
            sz = np.asarray(sizes, dtype=np.int64)[order]
            kd = [kids[i] for i in order]
            covered = np.zeros(int((st + sz).max()), dtype=bool)
            ok, per_layer = True, dict()
            for layer in gs:
                ents = []
                for (nids, cids, pids, pfids, nch, tot), run, al in zip(layer._dev["fw"], layer._fw_child_run,
                                                                        layer._fw_aligned):
                    if not (layer.block_size % 64 == 0 and run >= 128 and run % 128 == 0 and al):
                        ok = False
                        break
                    c = cids.detach().cpu().numpy().astype(np.int64)
                    n = nch.detach().cpu().numpy().astype(np.int64)
                    nst = -(-c.shape[1] // 128)
                    rows_tab = np.zeros([c.shape[0], nst, self._PROD_K[-1]], dtype=np.int64)
                    kp_max, reads, writes = 1, 0, 0
                    for j in range(c.shape[0]):
                        for sidx, k0 in enumerate(range(0, int(n[j]), 128)):
                            e0, rows = int(c[j, k0]), min(128, int(n[j]) - k0)
                            pb = int(np.searchsorted(st, e0, side="right")) - 1
                            if pb < 0 or e0 + rows > st[pb] + sz[pb]:
                                ok = False
                                break
                            r0 = e0 - int(st[pb])
                            k = kd[pb]
                            kp_max = max(kp_max, len(k))
                            rows_tab[j, sidx, :] = r0                  # dummy rows (value 0) by default
                            rows_tab[j, sidx, :len(k)] = k + r0
                            covered[e0:e0 + rows] = True
                            reads += rows * len(k)
                            writes += rows
                        if not ok:
                            break
                    if not ok:
                        break
                    kp = next(v for v in self._PROD_K if v >= kp_max)
                    tab = torch.from_numpy(np.ascontiguousarray(rows_tab[:, :, :kp]).astype(np.int32)).to(device)
                    ents.append((tab, kp, reads, writes))
                if not ok:
                    break
                per_layer[id(layer)] = ents
            if not ok:
                continue
            if not all(bool(covered[a:a + b].all()) for a, b in zip(st.tolist(), sz.tolist())):
                continue
            fused_groups.add(gid)
            tables.update(per_layer)
            # workgroups per 64 batch columns of the group's widest forward launch
            tiles[gid] = max(int(nids.numel()) * (layer.block_size // 64)
                             for layer in gs for nids, *_ in layer._dev["fw"])
        return {"groups": fused_groups, "tables": tables, "tiles": tiles}

    def _build_fused_backward(self, device):
        """Sum layers whose element-flow kernel also performs the product backward below them
        (``pjamd_sum_backward_ele_to_children``).  Checked here on the index tables: every child
        of the product group below has exactly one parent in the whole circuit (then a child's
        flow is its parent product's flow, written once, no accumulation); the sum layer's child
        blocks (cbs >= 32) are exactly those product blocks, of that size.  A sum layer all of
        whose nodes receive their flows this way is ``premod``: with ``allow_modify_flows`` the
        fused kernel already stores its flows as log(flow) - mar, so its own modify pass is
        skipped."""
        groups = self.inner_layer_groups
        cnt = np.zeros(self.num_nodes, dtype=np.int64)
        children = dict()   # product group id -> {element block start: (child block starts, bs)}
        for gid, g in enumerate(groups):
            if not g.is_prod():
                continue
            d = dict()
            for layer in g:
                bs = layer.block_size
                for n, c in zip(layer.partitioned_nids, layer.partitioned_cids):
                    n = n.detach().cpu().numpy().astype(np.int64)
                    c = c.detach().cpu().numpy().astype(np.int64)
                    real = c[c != 0]
                    if real.size:
                        np.add.at(cnt, (real[:, None] + np.arange(bs)[None, :]).reshape(-1), 1)
                    for e, ch in zip(n.tolist(), c):
                        d[int(e)] = (ch[ch != 0], bs)
            children[gid] = d

        raw_tables, fused_prod = dict(), set()
        covered = np.zeros(self.num_nodes, dtype=bool)
        for gid, g in enumerate(groups):
            if g.is_prod() or gid == 0 or not groups[gid - 1].is_prod():
                continue
            d = children[gid - 1]
            ok = all(bool(np.all(cnt[(ch[:, None] + np.arange(bs)[None, :]).reshape(-1)] == 1))
                     for ch, bs in d.values())
            per_layer, seen = dict(), set()
            for layer in g:
                if not ok:
                    break
                tabs = []
                for chids, parids, parpids, npar, cbs, totp in layer._dev["bk"]:
                    starts = chids.detach().cpu().numpy().astype(np.int64).tolist()
                    if cbs < 32 or any(c not in d or d[c][1] != cbs for c in starts):
                        ok = False
                        break
                    seen.update(starts)
                    rows = [d[c][0] for c in starts]
                    T = np.zeros([len(rows), max(max(len(x) for x in rows), 1)], dtype=np.int64)
                    for i, x in enumerate(rows):
                        T[i, :len(x)] = x
                    tabs.append(T)
                per_layer[id(layer)] = tabs
            if not ok or seen != set(d.keys()):
                continue
            raw_tables.update(per_layer)
            fused_prod.add(gid - 1)
            for ch, bs in d.values():
                covered[(ch[:, None] + np.arange(bs)[None, :]).reshape(-1)] = True

        premod, premod_row = set(), np.zeros(self.num_nodes, dtype=bool)
        for gid, g in enumerate(groups):
            if g.is_prod():
                continue
            for layer in g:
                a, b = layer.nid_range
                if covered[a:b].all():
                    premod.add(id(layer))
                    premod_row[a:b] = True
        # Node-flow rows the backward must clear first: everything except the rows the fused
        # kernels assign (every row of every child block they cover) and the root rows, as
        # contiguous ranges (the whole buffer when that would take many launches).
        need = np.ones(self.num_nodes, dtype=bool)
        need[covered] = False
        r0, r1 = self._root_node_range
        need[r0:r1] = False
        edges = np.flatnonzero(np.diff(np.concatenate([[0], need.astype(np.int8), [0]])))
        zero_ranges = [(int(a), int(b)) for a, b in zip(edges[0::2], edges[1::2])]
        if len(zero_ranges) > 64:
            zero_ranges = [(0, self.num_nodes)]
        tables = dict()
        for lid, tabs in raw_tables.items():
            out = []
            for T in tabs:
                code = 2 * T + (premod_row[T] & (T != 0)).astype(np.int64)
                assert code.max() < 2 ** 31
                # (table, child blocks written, of which stored as log(flow) - mar)
                out.append((torch.from_numpy(code).to(torch.int32).to(device).contiguous(),
                            int((code != 0).sum()), int((code & 1).sum())))
            tables[lid] = out
        return {"tables": tables, "prod_groups": fused_prod, "premod": premod, "zero_ranges": zero_ranges}

    def _build_par_batches(self, device):
        """Descriptor tables of the batched parameter-flow launch: every sum layer with block size
        >= 32 gets its parameter flows from one launch per tile shape after the whole flow
        backward (they feed only EM), instead of one latency-bound launch per layer.  Rows follow
        ``PJAMD_PAR_DESC_FIELDS``."""
        sum_layers = [(gid, layer) for gid, g in enumerate(self.inner_layer_groups) if not g.is_prod() for layer in g]
        # Tiles of different layers run concurrently here, so a layer whose parameter-flow range
        # meets another layer's (tied nodes share slots with their source) adds atomically, as
        # does a layer with tied nodes inside it.
        def pf_span(layer):
            lo, hi = 2 ** 62, -1
            for nids, cids, pids, pfids, nch, tot in layer._dev["fw"]:
                real = torch.arange(pfids.size(1), device=pfids.device).unsqueeze(0) < nch.unsqueeze(1)
                if bool(real.any()):
                    v = pfids[real]
                    lo, hi = min(lo, int(v.min())), max(hi, int(v.max()) + layer.block_size)
            return (lo, hi) if hi >= 0 else (0, 0)

        ranges = sorted((pf_span(layer), id(layer)) for _, layer in sum_layers)
        overlapping = set()
        hi, hid = -1, None
        for (r, i) in ranges:
            if r[0] < r[1] and r[0] < hi:
                overlapping.update((i, hid))
            if r[1] > hi:
                hi, hid = r[1], i
        # Tile shapes (rows x children): 128 x 128 where a partition has >= 128 rows and children
        # per block (half the L2 traffic and exp work per parameter of 64 x 64; off with
        # ``par_tile_rows = 64``), else 64 x 64, or 32 x 128 for block size 32.
        acc = {128: [], 64: [], 32: []}
        for gid, layer in sum_layers:
            bs = layer.block_size
            if bs < 32:
                continue
            erow = self._elem_offsets[gid // 2]
            atomic = layer._pf_shared or id(layer) in overlapping
            for e, (nids, cids, pids, pfids, nch, tot) in enumerate(layer._dev["fw"]):
                wide = self.par_tile_rows >= 128 and bs >= 128 and cids.size(1) >= 128
                tm, tn = (128, 128) if wide else (64, 64) if bs >= 64 else (32, 128)
                acc[tm].append((nids, cids, pids, pfids, nch, tot, bs, tn, erow, atomic, layer._fw_uniq[e]))
        batches = dict()
        for tm, parts in acc.items():
            if not parts:
                continue
            units = self._tied_par_units(parts, tm)
            parts = [parts[i] for u in units for i in u]
            desc, tile_desc = [], []
            n_tiles = blk_off = idx_off = 0
            tot_rows = tot_edges = tot_edges_bs = 0
            d = 0
            for u in units:
                for c in range(len(u)):
                    nids, cids, pids, pfids, nch, tot, bs, tn, erow, shared, uniq = parts[d]
                    nb, K = nids.numel(), cids.size(1)
                    mt, kt = bs // tm, (K + tn - 1) // tn
                    tiles = nb * mt * kt if c == 0 else 0
                    desc.append([n_tiles, K, mt, kt, blk_off, idx_off, erow, 1 if shared else 0,
                                 len(u) if c == 0 else 1])
                    tile_desc.append(torch.full([tiles], d, dtype=torch.int32))
                    n_tiles += tiles
                    blk_off += nb
                    idx_off += nb * K
                    tot_rows += nb * bs
                    tot_edges += uniq        # distinct element rows of the table
                    tot_edges_bs += tot * bs
                    d += 1
            assert idx_off < 2 ** 31 and n_tiles < 2 ** 31 // 256
            cat = lambda k: torch.cat([p[k].reshape(-1) for p in parts]).to(torch.int32).to(device).contiguous()
            batches[tm] = {
                "nids": cat(0), "cids": cat(1), "pids": cat(2), "pfids": cat(3), "nch": cat(4),
                "desc": torch.tensor(desc, dtype=torch.int32).reshape(-1).to(device),
                "tile_desc": torch.cat(tile_desc).to(device), "n_tiles": n_tiles,
                "work": (tot_rows, tot_edges, tot_edges_bs),
            }
        batched_layers = set()
        for gid, g in enumerate(self.inner_layer_groups):
            if not g.is_prod():
                batched_layers.update(id(layer) for layer in g if layer.block_size >= 32)
        return {"tables": batches, "layers": batched_layers}

    # Tiles a group of tied partitions should keep in one batched launch: fewer, longer tiles
    # mean fewer atomic adds into the shared slots, more keep the chip busy.
    par_tied_tiles = 1024
    # Largest parameter-flow tile height (128 or 64 rows).
    par_tile_rows = 128

    def _tied_par_units(self, parts, tm):
        """Order the batched parameter-flow partitions into units: a unit is a run of partitions
        whose tiles are summed before the slots are written once.  Partitions with the shared-slot
        flag and equal shapes, ``n_chs``, ``pids`` and ``pfids`` (the time steps of a homogeneous
        HMM, tied copies of one sum node set) are tied copies; a group of them is split into just
        enough units to give ``par_tied_tiles`` tiles.  Other partitions are units of one."""
        groups, order = dict(), []
        for i, (nids, cids, pids, pfids, nch, tot, bs, tn, erow, shared, _uniq) in enumerate(parts):
            if not shared:
                order.append([i])
                continue
            key = (bs, tn, nids.numel(), cids.size(1), nch.cpu().numpy().tobytes(),
                   pids.cpu().numpy().tobytes(), pfids.cpu().numpy().tobytes())
            if key not in groups:
                groups[key] = []
                order.append(groups[key])
            groups[key].append(i)
        units = []
        for g in order:
            nids, cids, bs, tn = parts[g[0]][0], parts[g[0]][1], parts[g[0]][6], parts[g[0]][7]
            tiles = nids.numel() * (bs // tm) * -(-cids.size(1) // tn)
            n_units = min(len(g), max(1, -(-self.par_tied_tiles // max(tiles, 1))))
            for u in range(n_units):
                lo, hi = u * len(g) // n_units, (u + 1) * len(g) // n_units
                units.append(g[lo:hi])
        return units

    def _run_par_batched(self, modify: bool, precision: str, overwrite: bool, negate: bool = False):
        pf = self._pf
        pf64 = pf.dtype == torch.float64
        flags = (1 if modify else 0) | (4 if negate else 0) | (8 if pf64 else 0) | (16 if overwrite else 0)
        prec = PRECISIONS[precision]
        B = self.node_flows.size(1)
        for tm, t in self._par_batches["tables"].items():
            rows, edges, edges_bs = t["work"]
            # algorithmic work: node-flow rows (and marginals unless modified) and child rows once
            # per sample, parameters read and parameter flows written once (and read, unless
            # overwritten)
            pfb = (1 if overwrite else 2) * (8 if pf64 else 4)
            algo = (4 * (rows * (1 if modify else 2) + edges) * B + (4 + pfb) * edges_bs, 2 * edges_bs * B)
            N.call("pjamd_sum_backward_par_batched", N.ptr(pf), N.ptr(self.node_flows), N.ptr(self.node_mars),
                   N.ptr(self.element_mars), N.ptr(self.params), N.ptr(t["nids"]), N.ptr(t["cids"]),
                   N.ptr(t["pids"]), N.ptr(t["pfids"]), N.ptr(t["nch"]), N.ptr(t["desc"]), N.ptr(t["tile_desc"]),
                   t["n_tiles"], tm, B, flags, prec, N.stream(), algo=algo)

    def cuda(self, device=None):
        return self.to(torch.device("cuda" if device is None else f"cuda:{device}" if isinstance(device, int) else device))

    # ------------------------------------------------------------------ buffers

    def _ensure_buffers(self, B: int, need_flows: bool = False):
        dev = self.device
        assert dev.type == "cuda", "TensorCircuit passes run on the HIP device; call `pc.to('cuda')` first"
        rows = self.num_elements if self.recompute_elements else self.num_element_rows
        if self.node_mars is None or self.node_mars.size(1) != B or self.element_mars.size(0) != rows:
            self.node_mars = torch.empty([self.num_nodes, B], device=dev, dtype=torch.float32)
            self.element_mars = torch.empty([rows, B], device=dev, dtype=torch.float32)
            self._fill(self.node_mars[:self.num_dummy_nodes], 0.0)
            for off in ([0] if self.recompute_elements else self._elem_offsets):
                self._fill(self.element_mars[off:off + self.num_dummy_eles], float("-inf"))
            self.node_flows = self.element_flows = None
        if need_flows and (self.node_flows is None or self.node_flows.size(1) != B):
            self.node_flows = torch.empty([self.num_nodes, B], device=dev, dtype=torch.float32)
            self.element_flows = torch.empty([self.num_elements, B], device=dev, dtype=torch.float32)
            self._fill(self.element_flows, 0.0)

    @staticmethod
    def _fill(t: torch.Tensor, value: float):
        if t.numel() > 0:
            N.call("pjamd_fill_f32", N.ptr(t), float(value), t.numel(), N.stream())

    def _to_vb(self, inputs: torch.Tensor, missing_mask: Optional[torch.Tensor] = None,
               batch_first: bool = True) -> torch.Tensor:
        """``[B, V]`` integer / bool data -> contiguous ``[V, B]`` int32 on the device; entries of
        ``missing_mask`` (``[V]``, ``[B, V]`` or, with ``batch_first=False``, ``[V, B]``) become -1,
        which the input layers marginalize (``input_layer.py:374-401``)."""
        assert inputs.dim() == 2 and inputs.size(1) == self.num_vars
        x = inputs if inputs.device == self.device else inputs.to(self.device)
        x = x.contiguous()
        code = _DTYPE_CODE.get(x.dtype)
        assert code is not None, f"unsupported data dtype {x.dtype}"
        B, V = x.shape
        out = torch.empty([V, B], device=self.device, dtype=torch.int32)
        N.call("pjamd_data_to_vb", N.ptr(out), code, N.ptr(x), B, V, N.stream())
        if missing_mask is not None:
            apply_missing_mask(out, missing_mask, batch_first)
        return out

    # ------------------------------------------------------------------ passes

    def set_propagation_alg(self, propagation_alg: str, **kwargs):
        """Default propagation algorithm of ``forward`` / ``backward``: "LL", "MPE", or
        "GeneralLL" with ``alpha`` (``tensorcircuit.py:192-205``)."""
        if propagation_alg in ("LL", "MPE"):
            self.default_propagation_alg = propagation_alg
            self.propagation_alg_kwargs.clear()
        elif propagation_alg == "GeneralLL":
            assert "alpha" in kwargs, "GeneralLL propagation needs an `alpha` keyword"
            self.default_propagation_alg = "GeneralLL"
            self.propagation_alg_kwargs.clear()
            self.propagation_alg_kwargs["alpha"] = kwargs["alpha"]
        else:
            raise NotImplementedError(f"propagation_alg {propagation_alg!r} is not one of LL, MPE, GeneralLL")

    def forward_ll(self, *args, **kwargs):
        return self.forward(*args, propagation_alg="LL", **kwargs)

    def forward_mpe(self, *args, **kwargs):
        return self.forward(*args, propagation_alg="MPE", **kwargs)

    def forward_general_ll(self, *args, alpha: float = 1.0, **kwargs):
        """GeneralLL forward with exponent ``alpha`` (``tensorcircuit.py:546-547``, which drops
        ``alpha`` before calling forward; here it is passed on)."""
        return self.forward(*args, propagation_alg="GeneralLL", alpha=alpha, **kwargs)

    def _group_algs(self, propagation_alg) -> list:
        """Propagation algorithm of every inner layer group: one name for all, or the reference's
        per-group list indexed like ``inner_layer_groups`` (``tensorcircuit.py:270-280, 460-470``;
        entries of product groups are ignored)."""
        if isinstance(propagation_alg, str):
            return [propagation_alg] * len(self.inner_layer_groups)
        algs = list(propagation_alg)
        assert len(algs) == len(self.inner_layer_groups), \
            f"a per-layer propagation_alg list needs {len(self.inner_layer_groups)} entries, got {len(algs)}"
        for a in algs:
            assert a in ("LL", "MPE", "GeneralLL"), f"propagation_alg {a!r} is not one of LL, MPE, GeneralLL"
        return algs

    def _run_inner_forward_alg(self, propagation_alg, alpha, precision="fp32"):
        """Layer-wise forward for MPE / GeneralLL or a per-group list of algorithms: product
        layers as for LL, sum layers through ``pjamd_sum_forward_alg`` (their "LL" groups through
        the LL kernels in ``precision``; no chains, fusion or graph capture)."""
        algs = self._group_algs(propagation_alg)
        for gid, g in enumerate(self.inner_layer_groups):
            if g.is_prod():
                g(self.node_mars, self._slab(gid))
            elif algs[gid] == "LL":
                g(self.node_mars, self._slab(gid), self.params, precision=precision)
            else:
                for layer in g:
                    layer(self.node_mars, self._slab(gid), self.params, propagation_alg=algs[gid], alpha=alpha)

    def _run_inner_backward_layerwise(self, compute_param_flows: bool, propagation_alg, alpha, precision: str,
                                      allow_modify_flows: bool, negate_pflows: bool, pre_callback=None,
                                      post_callback=None):
        """Layer-wise flow backward (``tensorcircuit.py:437-482``): MPE / GeneralLL sum groups write
        element flows (``pjamd_sum_backward_ele_alg``) and parameter flows
        (``pjamd_sum_backward_par_alg``), "LL" groups run the LL kernels without the fused product
        backward (their flows modified in place when ``allow_modify_flows``), and each product
        group spreads the element flows to its children.  ``pre_callback`` / ``post_callback`` are
        called as ``fn(layer, node_flows=..., element_flows=..., node_mars=..., element_mars=...,
        params=..., param_flows=...)`` around each sum layer's backward, like the reference's
        ``LayerGroup.callback``."""
        algs = self._group_algs(propagation_alg)
        pf = self._pf if compute_param_flows else None
        for gid in range(len(self.inner_layer_groups) - 1, -1, -1):
            g = self.inner_layer_groups[gid]
            if g.is_prod():
                g.backward(self.node_flows, self.element_flows)
                continue
            if self.recompute_elements:
                self.inner_layer_groups[gid - 1](self.node_mars, self._slab(gid - 1))
            slab = self._slab(gid)
            cb = dict(node_flows=self.node_flows, element_flows=self.element_flows, node_mars=self.node_mars,
                      element_mars=slab, params=self.params, param_flows=self._pf)
            if pre_callback is not None:
                g.callback(pre_callback, **cb)
            for layer in g:
                if algs[gid] == "LL":
                    layer.backward(self.node_flows, self.element_flows, self.node_mars, slab, self.params,
                                   param_flows=pf, allow_modify_flows=allow_modify_flows, precision=precision,
                                   negate_pflows=negate_pflows)
                else:
                    layer.backward(self.node_flows, self.element_flows, self.node_mars, slab, self.params,
                                   param_flows=pf, propagation_alg=algs[gid], alpha=alpha,
                                   negate_pflows=negate_pflows)
            if post_callback is not None:
                g.callback(post_callback, **cb)

    def _slab(self, group_id: int) -> torch.Tensor:
        """``element_mars`` rows of the product group ``group_id`` (or of the sum group right above it)."""
        off = 0 if self.recompute_elements else self._elem_offsets[group_id // 2]
        return self.element_mars.narrow(0, off, self._group_elems[group_id // 2])

    def _fuse_now(self, fz, gid: int, B: int) -> bool:
        return (fz is not None and gid in fz["groups"]
                and fz["tiles"][gid] * -(-B // 64) < self.fuse_products_max_tiles)

    def _run_inner_forward(self, force_use_bf16, force_use_fp32, precision):
        groups = self.inner_layer_groups
        B = self.node_mars.size(1)
        plan = self._chain_plan(precision, B)
        runs, packed = dict(), None
        if plan is not None:
            runs = plan.forward_runs()
            packed = self._chain_pack(plan, precision)
        fz = self._prod_fusion
        if (not self.fuse_products or runs or self.recompute_elements or precision not in ("bf16", "fp16")
                or B % 4 != 0):
            fz = None
        gid = 0
        while gid < len(groups):
            run = runs.get(gid)
            if run is not None:
                comp, ops, tab = run.fwd
                N.call("pjamd_chain_forward", N.ptr(self.node_mars), N.ptr(self.element_mars), N.ptr(packed),
                       N.ptr(comp), N.ptr(ops), N.ptr(tab), run.C, B, PRECISIONS[precision], N.stream(),
                       algo=run.algo(B, "fwd"))
                gid = run.hi
                continue
            g = groups[gid]
            if g.is_prod():
                if not self._fuse_now(fz, gid, B):   # else run inside the next sum group
                    g(self.node_mars, self._slab(gid))
            else:
                tabs = fz["tables"] if self._fuse_now(fz, gid - 1, B) else None
                g(self.node_mars, self._slab(gid), self.params, force_use_bf16=force_use_bf16,
                  force_use_fp32=force_use_fp32, precision=precision, prod_tables=tabs)
            gid += 1

    def forward(self, inputs: torch.Tensor, input_layer_fn: Optional[Union[str, Callable]] = None,
                cache: Optional[dict] = None, return_cache: bool = False, record_cudagraph: bool = False,
                apply_cudagraph: bool = True, force_use_bf16: bool = False, force_use_fp32: bool = False,
                propagation_alg: Optional[Union[str, Sequence[str]]] = None, precision: Optional[str] = None,
                **kwargs):
        """Log-likelihoods ``[B, num_roots]`` of ``inputs`` ``[B, num_vars]`` (``tensorcircuit.py:207-342``)."""
        if propagation_alg is None:
            propagation_alg = self.default_propagation_alg
            kwargs.update(self.propagation_alg_kwargs)
        alpha = kwargs.pop("alpha", None)
        prec = precision if precision is not None else ("bf16" if force_use_bf16 else self.precision)
        if force_use_fp32:
            prec = "fp32"
        origin_inputs = inputs
        missing_mask = kwargs.get("missing_mask")
        batch_first = kwargs.get("_batch_first", True)
        with torch.no_grad():
            B = inputs.size(0)
            self._ensure_buffers(B)
            if input_layer_fn is None:
                data = self._to_vb(inputs, missing_mask, batch_first)
                self._data_vb, self._data_marginalized = data, missing_mask is not None
                # the backward reuses this transpose for the same, unmodified input tensor
                self._data_src = ((weakref.ref(inputs), inputs._version) if missing_mask is None and batch_first
                                  else None)
                kw = {k: v for k, v in kwargs.items() if k not in ("missing_mask", "_batch_first")}
                for layer in self.input_layer_group:
                    layer(data, self.node_mars, **kw)
            elif isinstance(input_layer_fn, str):
                for layer in self.input_layer_group:
                    getattr(layer, input_layer_fn)(inputs, self.node_mars, **kwargs)
            else:
                for layer in self.input_layer_group:
                    ret = input_layer_fn(layer, inputs, self.node_mars, **kwargs)
                    if not ret and ret is not None:
                        layer(self._to_vb(inputs), self.node_mars, **kwargs)

            sig = (0, id(self.node_mars), id(self.element_mars), id(self.params), B, prec)
            run = lambda: self._run_inner_forward(force_use_bf16, force_use_fp32, prec)
            if not isinstance(propagation_alg, str) or propagation_alg != "LL":
                self._run_inner_forward_alg(propagation_alg, alpha, prec)
            elif record_cudagraph and sig not in self._recorded_cuda_graphs:
                run()  # this call's result; the capture below does not execute kernels
                self._recorded_cuda_graphs[sig] = self._capture(run)
            elif apply_cudagraph and sig in self._recorded_cuda_graphs:
                self._recorded_cuda_graphs[sig].replay()
            else:
                run()
            lls = self.node_mars[self._root_node_range[0]:self._root_node_range[1], :].permute(1, 0)

        if return_cache:
            cache = dict() if cache is None else cache
            cache["node_mars"] = self.node_mars.clone()
        if torch.is_grad_enabled():
            lls.requires_grad = True
            if alpha is not None:
                kwargs["alpha"] = alpha
            lls.register_hook(partial(_pc_model_backward_hook, pc=self, inputs=origin_inputs,
                                      record_cudagraph=record_cudagraph, apply_cudagraph=apply_cudagraph,
                                      propagation_alg=propagation_alg, precision=prec, **kwargs))
        out = lls.clone()
        return (out, cache) if return_cache else out

    def _capture(self, fn):
        """Record ``fn``'s kernel launches into a HIP graph (nothing is executed)."""
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            fn()
        torch.cuda.synchronize()
        return g

    def _run_inner_backward(self, compute_param_flows, allow_modify_flows, precision, overwrite=False,
                            negate_pflows=False):
        # Element marginals of every layer are still resident (one slab per product group), so no
        # product layer is recomputed here (the reference recomputes, tensorcircuit.py:446-447),
        # unless ``recompute_elements`` keeps a single shared slab.
        # ``allow_modify_flows`` permits the in-place log(flow) - mar transform of the sum nodes'
        # flows (one pass per layer, re-read by the element- and parameter-flow kernels); with
        # ``modify_flows_in_place = False`` the kernels compute it on the fly instead.
        # Parameter flows of sum layers with block size >= 32 come from one batched launch after
        # the loop (the flows and element marginals they read stay untouched until then).
        # Where the product layer below a sum layer only fans flows out to single-parent children,
        # the sum layer's element-flow kernel writes those children's node flows itself (and, for
        # sum children, already as log(flow) - mar), so that product backward and the children's
        # modify pass are not launched.
        modify = allow_modify_flows and self.modify_flows_in_place
        recompute = self.recompute_elements
        batched = self._par_batches["layers"] if (compute_param_flows and precision != "fp32_valu" and
                                                  not recompute) else set()
        fz = self._fused if precision != "fp32_valu" else {"tables": dict(), "prod_groups": set(), "premod": set()}
        B = self.node_flows.size(1)
        plan = self._chain_plan(precision, B) if not negate_pflows else None
        packed = self._chain_packed.get((precision, id(plan))) if plan is not None else None
        runs = plan.backward_runs() if packed is not None else dict()
        lid = len(self.inner_layer_groups) - 1
        while lid >= 0:
            run = runs.get(lid)
            if run is not None:
                comp, ops, tab = run.bwd[bool(modify)]
                N.call("pjamd_chain_backward", N.ptr(self.node_flows), N.ptr(self.node_mars),
                       N.ptr(self.element_mars), N.ptr(packed), N.ptr(comp), N.ptr(ops), N.ptr(tab), run.C, B,
                       PRECISIONS[precision], N.stream(), algo=run.algo(B, "bwd"))
                lid = run.lo - 1
                continue
            gid, lid = lid, lid - 1
            g = self.inner_layer_groups[gid]
            if g.is_prod():
                if gid not in fz["prod_groups"]:
                    g.backward(self.node_flows, self.element_flows)
                continue
            if recompute:
                # the element rows of the product group below, overwritten by later groups' forward
                self.inner_layer_groups[gid - 1](self.node_mars, self._slab(gid - 1))
            for layer in g:
                pf = self._pf if compute_param_flows and id(layer) not in batched else None
                layer.backward(self.node_flows, self.element_flows, self.node_mars, self._slab(gid), self.params,
                               param_flows=pf, allow_modify_flows=modify, precision=precision,
                               premodified=modify and id(layer) in fz["premod"],
                               child_tables=fz["tables"].get(id(layer)), overwrite_pflows=overwrite,
                               negate_pflows=negate_pflows)
        if batched:
            self._run_par_batched(modify, precision, overwrite, negate_pflows)

    def backward(self, inputs: Optional[torch.Tensor] = None, ll_weights: Optional[torch.Tensor] = None,
                 compute_param_flows: bool = True, flows_memory: float = 1.0,
                 input_layer_fn: Optional[Union[str, Callable]] = None, cache: Optional[dict] = None,
                 sum_layer_pre_backward_callback: Optional[Callable] = None,
                 sum_layer_post_backward_callback: Optional[Callable] = None, return_cache: bool = False,
                 record_cudagraph: bool = False, apply_cudagraph: bool = True, allow_modify_flows: bool = True,
                 propagation_alg: Union[str, Sequence[str]] = "LL", logspace_flows: bool = False,
                 negate_pflows: bool = False, precision: Optional[str] = None, force_use_fp32: bool = False,
                 **kwargs):
        """Node flows and accumulated parameter flows (``tensorcircuit.py:344-538``).  With
        ``missing_mask`` (as in ``forward``; when ``inputs`` is None, the forward's mask holds), a
        marginalized input entry's flow is spread over its categories by ``theta``."""
        missing_mask = kwargs.get("missing_mask")
        batch_first = kwargs.get("_batch_first", True)
        assert not (logspace_flows and allow_modify_flows), \
            "logspace_flows keeps node flows as log(flow): pass allow_modify_flows=False with it"
        alpha = kwargs.pop("alpha", None)
        layerwise = (not isinstance(propagation_alg, str) or propagation_alg != "LL"
                     or sum_layer_pre_backward_callback is not None or sum_layer_post_backward_callback is not None)
        assert self.node_mars is not None and self.element_mars is not None, "backward() needs the marginals of a preceding forward()"
        prec = precision if precision is not None else self.precision
        if force_use_fp32:
            prec = "fp32"
        self._run_params.update(allow_modify_flows=allow_modify_flows, propagation_alg=propagation_alg,
                                logspace_flows=logspace_flows, negate_pflows=negate_pflows, precision=prec)
        with torch.no_grad():
            B = self.node_mars.size(1)
            self._ensure_buffers(B, need_flows=True)
            if prec == "fp32_valu" or layerwise or "zero_ranges" not in self._fused:
                self._fill(self.node_flows, 0.0)
            else:
                for a, b in self._fused["zero_ranges"]:
                    self._fill(self.node_flows[a:b], 0.0)
            r0, r1 = self._root_node_range
            if ll_weights is None:
                self._fill(self.node_flows[r0:r1], 1.0)
            else:
                if ll_weights.dim() == 1:
                    ll_weights = ll_weights.unsqueeze(0)
                assert ll_weights.size(0) == self.num_root_nodes
                self.node_flows[r0:r1, :] = ll_weights.to(self.node_flows.dtype)
            if compute_param_flows:
                if ll_weights is None:
                    self._cum_flow_host += (r1 - r0) * B
                else:
                    w = ll_weights.detach().sum().to(torch.float64)
                    self._cum_flow_dev = w if self._cum_flow_dev is None else self._cum_flow_dev + w
                self.init_param_flows(flows_memory=flows_memory)

            # The first backward after init_param_flows(0.0) stores the sum parameter flows instead
            # of adding them when every slot has exactly one writer; otherwise the buffer is
            # cleared first.
            overwrite = (compute_param_flows and self._pf_fresh and self._pf_overwrite_ok and not layerwise)
            if compute_param_flows and not overwrite:
                _ = self.param_flows
            sig = (1, id(self.node_flows), id(self.element_flows), id(self.node_mars), id(self.element_mars),
                   id(self.params), id(self._pf), B, allow_modify_flows, compute_param_flows, prec, overwrite,
                   negate_pflows)
            run = lambda: self._run_inner_backward(compute_param_flows, allow_modify_flows, prec, overwrite,
                                                   negate_pflows)
            if layerwise:
                self._run_inner_backward_layerwise(compute_param_flows, propagation_alg, alpha, prec,
                                                   allow_modify_flows, negate_pflows,
                                                   sum_layer_pre_backward_callback, sum_layer_post_backward_callback)
            elif record_cudagraph and sig not in self._recorded_cuda_graphs:
                run()  # this call's result; the capture below does not execute kernels
                self._recorded_cuda_graphs[sig] = self._capture(run)
            elif apply_cudagraph and sig in self._recorded_cuda_graphs:
                self._recorded_cuda_graphs[sig].replay()
            else:
                run()
            if compute_param_flows:
                self._pf_fresh = False

            if input_layer_fn is None:
                src = self._data_src
                if inputs is None:
                    data, marg = self._data_vb, self._data_marginalized
                elif (src is not None and missing_mask is None and batch_first and src[0]() is inputs
                      and src[1] == inputs._version and self._data_vb is not None):
                    data, marg = self._data_vb, False
                else:
                    data, marg = self._to_vb(inputs, missing_mask, batch_first), missing_mask is not None
                kw = {k: v for k, v in kwargs.items() if k not in ("missing_mask", "_batch_first")}
                for layer in self.input_layer_group:
                    if compute_param_flows:
                        layer.backward(data, self.node_flows, self.node_mars, marginalized=marg, **kw)
            elif isinstance(input_layer_fn, str):
                for layer in self.input_layer_group:
                    getattr(layer, input_layer_fn)(inputs, self.node_flows, self.node_mars, **kwargs)
            else:
                for layer in self.input_layer_group:
                    ret = input_layer_fn(layer, inputs, self.node_flows, self.node_mars, **kwargs)
                    if not ret and ret is not None:
                        layer.backward(self._to_vb(inputs), self.node_flows, self.node_mars, **kwargs)
            if logspace_flows:
                # the flows were propagated (and consumed by the input layers) in linear space;
                # the buffer is handed back as log(flow), -inf where no flow arrived
                N.call("pjamd_flows_to_log", N.ptr(self.node_flows), self.node_flows.numel(), N.stream())

        if return_cache:
            cache = dict() if cache is None else cache
            cache["node_flows"] = self.node_flows.clone()
            return cache
        return None

    def mini_batch_em(self, step_size: float, pseudocount: float = 0.0, keep_zero_params: bool = False,
                      step_size_rescaling: bool = False, use_cudagraph: bool = True):
        """EM update from the accumulated parameter flows (``tensorcircuit.py:549-590``):
        input layers, then tied-flow fusing, then the sum parameters.  ``step_size_rescaling``:
        the flows become step_size / (accumulated root flow) times themselves plus (1 - step_size)
        times the current model's expected flows (theta * top-down probability), and the update
        then takes step 1 (``tensorcircuit.py:566-578``)."""
        assert not step_size_rescaling or self._cum_flow > 0.0, \
            "step-size rescaled EM needs flows from a backward() first"
        assert 0.0 < step_size <= 1.0, "EM step_size must lie in (0, 1]"
        if self.param_flows is None:
            self.init_param_flows(0.0)
        if step_size_rescaling:
            from .top_down import eval_top_down_probs
            self.init_param_flows(flows_memory=step_size / self._cum_flow)
            eval_top_down_probs(self, update_pflow=True, scale=1.0 - step_size, use_cudagraph=use_cudagraph)
            self._cum_flow = 0.0
            step_size = 1.0
        with torch.no_grad():
            for layer in self.input_layer_group:
                layer.mini_batch_em(step_size=step_size, pseudocount=pseudocount, keep_zero_params=keep_zero_params)
            pf64 = 1 if self.param_flows.dtype == torch.float64 else 0
            s = N.stream()
            for t, n, ch in self._em_dev["fuse"]:
                N.call("pjamd_pflow_fuse", N.ptr(self.param_flows), pf64, t, n, N.ptr(ch), ch.numel(), s)
            blocks = self._em_dev["blocks"]
            if blocks.numel() > 0:
                n = self.num_sum_params - self.num_dummy_params
                N.call("pjamd_sum_em", N.ptr(self.params), N.ptr(self.param_flows), pf64, N.ptr(blocks),
                       blocks.size(0), float(step_size), float(pseudocount), 1 if keep_zero_params else 0,
                       self._em_dev["max_bs"], s, algo=(n * (8 + self.param_flows.element_size()), 0))

    def cumulate_flows(self, inputs: torch.Tensor, params: Optional[torch.Tensor] = None):
        with torch.no_grad():
            self.forward(inputs)
            self.backward(inputs=inputs, compute_param_flows=True, flows_memory=1.0)

    def init_param_flows(self, flows_memory: float = 1.0, batch_size: Optional[int] = None):
        """Scale (``flows_memory`` in [0, 1]) or allocate the parameter-flow accumulators."""
        assert 0.0 <= flows_memory <= 1.0, "flows_memory must lie in [0, 1]"
        assert batch_size is None, "per-sample parameter flows are not supported"
        dev = self.device
        if self._pf is None or self._pf.device != dev or self._pf.dtype != self.pflow_dtype:
            self.param_flows = torch.zeros([self.num_param_flows], device=dev, dtype=self.pflow_dtype)
        elif flows_memory != 1.0:
            if flows_memory == 0.0:
                self._pf_fresh = True     # cleared lazily (see the param_flows property)
            else:
                self.param_flows.mul_(flows_memory)
        for layer in self.input_layer_group:
            layer.init_param_flows(flows_memory=flows_memory, dtype=self.pflow_dtype)
        return None

    def zero_param_flows(self):
        self.init_param_flows(flows_memory=0.0)

    def update_parameters(self, clone: bool = True):
        params = self.params.detach().cpu()
        for ns in self.root_ns:
            if ns.is_sum() and not ns.is_tied():
                ns.update_parameters(params, clone=clone)
        for layer in self.input_layer_group:
            layer.update_parameters()

    def update_param_flows(self, clone: bool = True, origin_ns_only: bool = True):
        pf = self.param_flows.detach().cpu()
        for ns in self.root_ns:
            if ns.is_sum() and not ns.is_tied():
                ns.update_param_flows(pf, clone=clone, origin_ns_only=origin_ns_only)

    def print_statistics(self):
        print(f"nodes:            {self.num_nodes}")
        print(f"edges:            {self.num_edges}")
        print(f"sum parameters:   {self.num_sum_params}")
        print(f"input parameters: {sum(l.params.numel() for l in self.input_layer_group)}")

    def get_node_mars(self, ns: CircuitNodes):
        assert self.node_mars is not None
        a, b = ns._output_ind_range
        if ns.is_sum() or ns.is_input():
            return self.node_mars[a:b, :].detach()
        for gid, g in enumerate(self.inner_layer_groups):
            for layer in g:
                if layer.is_prod() and any(n is ns for n in layer.nodes):
                    slab = self._slab(gid)
                    layer(self.node_mars, slab)
                    return slab[a:b, :].detach()
        raise ValueError("node not found")

    def get_node_flows(self, ns: CircuitNodes, **kwargs):
        assert self.node_flows is not None
        a, b = ns._output_ind_range
        if ns.is_sum() or ns.is_input():
            return self.node_flows[a:b, :].detach()
        for lid, g in enumerate(self.inner_layer_groups):
            for layer in g:
                if layer.is_prod() and any(n is ns for n in layer.nodes):
                    self.inner_layer_groups[lid + 1].backward(
                        self.node_flows, self.element_flows, self.node_mars, self._slab(lid), self.params,
                        param_flows=None, allow_modify_flows=False,
                        precision=self._run_params.get("precision", self.precision))
                    return self.element_flows[a:b, :].detach()
        raise ValueError("node not found")

    def layers(self, reverse: bool = False, ret_layer_groups: bool = False, ignore_input_layers: bool = False):
        return layer_iterator(self, reverse=reverse, ret_layer_groups=ret_layer_groups,
                              ignore_input_layers=ignore_input_layers)

    # ------------------------------------------------------------------ compilation

    def _get_num_vars(self, ns):
        return max(ns.scope.to_list()) + 1 if len(ns.scope) > 0 else 0

    def _init_layers(self, layer_sparsity_tol=None, max_num_partitions=None, max_tied_ns_per_parflow_block=8,
                     verbose=False):
        """Depth-layering, global id assignment and per-layer compilation (``tensorcircuit.py:920-1092``)."""
        if hasattr(self, "input_layer_group") or hasattr(self, "inner_layer_groups"):
            raise ValueError("this TensorCircuit is already compiled; "
                             "build a new one from the root nodes instead")
        self.root_ns._clear_tensor_circuit_hooks()
        pc_num_vars = len(self.root_ns.scope)
        depth2nodes, num_layers, max_node_bs, max_ele_bs = self._create_node_layers()

        self.input_layer_group = None
        self.inner_layer_groups = []
        self.num_dummy_nodes = max_ele_bs
        self.num_dummy_eles = max_node_bs
        self.num_dummy_params = max_node_bs * max_ele_bs
        num_nodes = self.num_dummy_nodes
        num_edges = 0
        num_elements = self.num_dummy_eles
        num_params = self.num_dummy_params
        num_pflows = 0
        node2tiednodes = dict()

        layer_id = 0
        self._group_elems = []
        for depth in range(num_layers):
            if depth == 0:
                sig2nodes = dict()
                for ns in depth2nodes[0]["input"]:
                    sig2nodes.setdefault(ns.dist.get_signature(), []).append(ns)
                layers = []
                for sig, nodes in sig2nodes.items():
                    layer = InputLayer(nodes=nodes, cum_nodes=num_nodes,
                                       max_tied_ns_per_parflow_block=max_tied_ns_per_parflow_block,
                                       pc_num_vars=pc_num_vars)
                    layers.append(layer)
                    num_nodes += layer.num_nodes
                self.input_layer_group = LayerGroup(layers)
                continue

            assert len(depth2nodes[depth]["prod"]) > 0 and len(depth2nodes[depth]["sum"]) > 0, \
                f"layer depth {depth} lacks product or sum vectors ({len(depth2nodes[depth]['prod'])} / {len(depth2nodes[depth]['sum'])})"
            g2p = dict()
            for ns in depth2nodes[depth]["prod"]:
                g2p.setdefault(ns.block_size, []).append(ns)
            layer_num_elements = max_node_bs
            players = []
            for gsize, nodes in g2p.items():
                pl = ProdLayer(nodes=nodes, global_nid_start=layer_num_elements, layer_sparsity_tol=layer_sparsity_tol,
                               max_num_partitions=max_num_partitions)
                layer_num_elements += pl.num_nodes
                num_edges += pl.num_edges
                players.append(pl)
            pg = LayerGroup(players)
            self.inner_layer_groups.append(pg)
            self.add_module(f"prod_layer_{layer_id}", pg)
            num_elements = max(num_elements, layer_num_elements)
            self._group_elems.append(layer_num_elements)

            g2s = dict()
            for ns in depth2nodes[depth]["sum"]:
                g2s.setdefault(ns.block_size, []).append(ns)
            slayers = []
            for gsize, nodes in g2s.items():
                sl = SumLayer(nodes=nodes, global_nid_start=num_nodes, global_pid_start=num_params,
                              global_pfid_start=num_pflows, node2tiednodes=node2tiednodes,
                              layer_sparsity_tol=layer_sparsity_tol, max_num_partitions=max_num_partitions,
                              max_tied_ns_per_parflow_block=max_tied_ns_per_parflow_block)
                num_nodes += sl.num_nodes
                num_edges += sl.num_edges
                num_params += sl.num_parameters
                num_pflows += sl.num_param_flows
                slayers.append(sl)
            sg = LayerGroup(slayers)
            self.inner_layer_groups.append(sg)
            self.add_module(f"sum_layer_{layer_id}", sg)
            layer_id += 1

        self.num_nodes = num_nodes
        self.num_edges = num_edges
        self.num_elements = num_elements
        # Element rows of every product group live in their own slab of ``element_mars`` (HBM is
        # large enough to keep all of them), so the backward pass needs no product recompute.
        offs, tot = [], 0
        for n_el in self._group_elems:
            offs.append(tot)
            tot += n_el
        self._elem_offsets = offs
        self.num_element_rows = tot
        self.num_sum_params = num_params
        self.num_param_flows = num_pflows
        self.node2tiednodes = node2tiednodes
        self.parflow_fusing_kwargs = compile_cum_par_flows_fn(node2tiednodes, MAX_NBLOCKS=2048, BLOCK_SIZE=2048)
        self.par_update_kwargs = compile_par_update_fn(self.root_ns, BLOCK_SIZE=32)
        self.num_root_nodes = self.root_ns.num_nodes
        self._root_node_range = (self.num_nodes - self.num_root_nodes, self.num_nodes)
        self._init_parameters()

    def _init_parameters(self, perturbation: float = 4.0, pseudocount: float = 0.0):
        """Random / copied initial parameters (``tensorcircuit.py:1094-1116``): same torch CPU RNG
        draws, in the same order, as the reference."""
        for ns in self.root_ns:
            if not ns.is_tied() and (ns.is_sum() or ns.is_input()) and not ns.has_params():
                ns.init_parameters(perturbation=perturbation, recursive=False)
        params = torch.exp(torch.rand([self.num_sum_params]) * -perturbation)
        params[:self.num_dummy_params] = 0.0
        for ns in self.root_ns:
            if ns.is_sum() and not ns.is_tied() and ns.has_params():
                ns.gather_parameters(params)
        normalize_parameters(params, self.par_update_kwargs, pseudocount=pseudocount)
        self.params = nn.Parameter(params, requires_grad=False)
        for layer in self.input_layer_group:
            layer._init_parameters(perturbation)

    def _create_node_layers(self):
        """Depth assignment with pass-through insertion (``tensorcircuit.py:1122-1196``).

        Inputs have depth 0, a product is one deeper than its deepest child, a sum sits at the
        depth of its deepest child.  A sum child (product) that is shallower than its parent
        gets an identity ``summate`` + ``multiply`` pair inserted, which mutates the DAG."""
        depth2nodes = dict()
        nodes2depth = dict()
        stats = {"num_layers": 1, "max_node_bs": 0, "max_ele_bs": 0}

        def finish(ns):
            if ns.is_input():
                nodes2depth[ns] = 0
                depth2nodes.setdefault(0, {"input": []})["input"].append(ns)
                return
            depth = max(nodes2depth[c] for c in ns.chs) + (1 if ns.is_prod() else 0)
            stats["num_layers"] = max(depth + 1, stats["num_layers"])
            nodes2depth[ns] = depth
            if depth not in depth2nodes:
                depth2nodes[depth] = {"sum": [], "prod": []}
            if ns.is_sum():
                for idx, cs in enumerate(ns.chs):
                    cd = nodes2depth[cs]
                    if cd < depth:
                        pass_sum = summate(cs, num_node_blocks=cs.num_node_blocks, block_size=cs.block_size,
                                           edge_ids=torch.arange(0, cs.num_node_blocks)[None, :].repeat(2, 1),
                                           params=torch.eye(cs.block_size)[None, :, :].repeat(cs.num_node_blocks, 1, 1))
                        pass_prod = multiply(pass_sum)
                        ns.chs[idx] = pass_prod
                        depth2nodes[cd]["sum"].append(pass_sum)
                        nodes2depth[pass_sum] = cd
                        nodes2depth[pass_prod] = depth
                depth2nodes[depth]["sum"].append(ns)
                stats["max_node_bs"] = max(stats["max_node_bs"], ns.block_size)
            elif ns.is_prod():
                stats["max_ele_bs"] = max(stats["max_ele_bs"], ns.block_size)
            else:
                raise NotImplementedError(f"cannot place a node vector of type {type(ns).__name__} in a layer")

        dfs_postorder(self.root_ns, finish)

        pns2layer = dict()
        for layer in range(1, len(depth2nodes)):
            for ns in depth2nodes[layer]["sum"]:
                for cs in ns.chs:
                    if cs.is_prod():
                        if cs in pns2layer:
                            assert pns2layer[cs] == layer, \
                                "a product vector would be needed by sum layers at two different depths"
                        else:
                            depth2nodes[layer]["prod"].append(cs)
                            pns2layer[cs] = layer
        return depth2nodes, stats["num_layers"], stats["max_node_bs"], stats["max_ele_bs"]


def compile(ns: CircuitNodes, layer_sparsity_tol: float = 0.5, max_num_partitions: Optional[int] = None,
            disable_gpu_compilation: bool = False, force_gpu_compilation: bool = False,
            max_tied_ns_per_parflow_block: int = 8, device: Optional[Union[int, torch.device]] = None,
            verbose: bool = True, **kwargs) -> TensorCircuit:
    """Compile a node DAG into a ``TensorCircuit`` (``tensorcircuit.py:1224-1261``)."""
    return TensorCircuit(ns, layer_sparsity_tol=layer_sparsity_tol, max_num_partitions=max_num_partitions,
                         disable_gpu_compilation=disable_gpu_compilation, force_gpu_compilation=force_gpu_compilation,
                         max_tied_ns_per_parflow_block=max_tied_ns_per_parflow_block, device=device,
                         verbose=verbose, **kwargs)
