"""Sum layer (``pyjuice/layer/sum_layer.py``): compile in ``__init__`` (``sum_layer.py:38-198``)
and the log-sum-exp forward, flow backward and parameter-flow accumulation through the
HIP library."""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np
import torch
import torch.nn as nn

from pyjuice_amd import _native as N
from .compilation import (assign_partitions, get_sum_layer_backward_stats, get_sum_layer_forward_stats,
                          sum_layer_backward_compilation, sum_layer_forward_compilation)
from .layer import Layer
from .partition import next_power_of_2, partition_nodes_by_n_edges

__all__ = ["SumLayer", "PRECISIONS"]

# Arithmetic of the intra-block contractions (FP32 accumulation and max subtraction in every
# case): FP32 / BF16 / FP16 operands on the MFMA matrix cores, or FP32 on the vector ALUs.
# Blocks smaller than 32 always use the vector-ALU FP32 kernels.
PRECISIONS = {"fp32": 0, "bf16": 1, "fp16": 2, "fp32_valu": 3}


def _unique_pow2(x: torch.Tensor) -> torch.Tensor:
    return torch.unique(torch.from_numpy(next_power_of_2(x)))


ALG_NEGATE = 16   # PJAMD_ALG_NEGATE: parameter flows of MPE / GeneralLL subtracted
LAUNCH_PRE = 1 << 28     # PJAMD_LAUNCH_PRE
LAUNCH_NOPRE = 1 << 27   # PJAMD_LAUNCH_NOPRE
LAUNCH_WIDE = 1 << 26    # PJAMD_LAUNCH_WIDE
_HINT_BITS = {"pre": LAUNCH_PRE, "nopre": LAUNCH_NOPRE, "wide": LAUNCH_WIDE}
PARAMS_ALIGNED = 1 << 30   # PJAMD_PARAMS_ALIGNED


def _params_aligned(idx: torch.Tensor, counts: torch.Tensor, bs: int) -> bool:
    """Every real entry of an index table (the first counts[j] of row j) is a multiple of 4, and
    so is the block size: parameter rows start 16-byte aligned."""
    real = torch.arange(idx.size(1), device=idx.device)[None, :] < counts.to(idx.device)[:, None]
    return bs % 4 == 0 and bool(((idx[real] % 4) == 0).all())


def _merge_rows(tables: Sequence[tuple], count_col: int) -> tuple:
    """Concatenate partition tables ``(ids [n], idx_1 [n, K_1], ..., counts [n])`` row-wise,
    zero-padding the 2-D tables to the widest K (index 0 is the padding entry of every table),
    and order the rows by decreasing count (stable)."""
    K = max(int(t[1].size(1)) for t in tables)
    out = []
    for col in range(len(tables[0])):
        parts = []
        for t in tables:
            v = t[col]
            if v.dim() == 2 and v.size(1) < K:
                v = torch.nn.functional.pad(v, (0, K - v.size(1)))
            parts.append(v)
        out.append(torch.cat(parts, dim=0))
    order = torch.sort(out[count_col], descending=True, stable=True).indices
    return tuple(v[order].contiguous() for v in out)


def _i32(t: torch.Tensor) -> torch.Tensor:
    if t.numel() > 0:
        assert int(t.max()) < 2 ** 31, "index exceeds the int32 range of the HIP kernels"
    return t.to(torch.int32).contiguous()


# Split-reduction workspaces: for each (device, stream, size) a ring of _SPLIT_RING buffers,
# handed to the layers that need that size in turn.  The layers' split launches and combines run
# in the stream's order, so layers may share a buffer; a ring of 4 measured faster on the
# homogeneous HMM than one buffer for all layers (a layer then overwrites what the previous one
# just read) and than one buffer per layer (0.5 GB there instead of 16 MB).  A layer keeps the
# buffer of its first call's stream, so circuits driven concurrently from different streams get
# different buffers.  Buffers are never freed, so HIP graphs that recorded their addresses stay
# valid.
_SPLIT_RING = 4
_SPLIT_WS = dict()


def _distinct_rows(idx: torch.Tensor, counts: torch.Tensor, rows_per_entry: int) -> int:
    """Number of distinct rows named by the real entries (the first ``counts[j]`` of row j) of an
    index table, each entry naming ``rows_per_entry`` consecutive rows."""
    I = idx.detach().cpu().numpy().astype(np.int64)
    k = counts.detach().cpu().numpy().astype(np.int64)
    real = np.arange(I.shape[1])[None, :] < k[:, None]
    return int(np.unique(I[real]).size) * rows_per_entry


def _split_workspace(numel: int, device) -> torch.Tensor:
    key = (str(device), N.stream().value, int(numel))
    ring = _SPLIT_WS.setdefault(key, [[], 0])
    bufs, turn = ring
    ring[1] = turn + 1
    if len(bufs) < _SPLIT_RING:
        bufs.append(torch.empty([numel], dtype=torch.float32, device=device))
    return bufs[turn % _SPLIT_RING]


class SumLayer(Layer, nn.Module):
    BLOCK_SPARSE = 0
    SPARSE = 1
    PYTORCH = 2
    STR2MODE = {"block_sparse": 0, "sparse": 1, "pytorch": 2}

    def __init__(self, nodes: Sequence, global_nid_start: int, global_pid_start: int, global_pfid_start: int,
                 node2tiednodes: dict, layer_sparsity_tol: Optional[float] = None,
                 max_num_partitions: Optional[int] = None, max_tied_ns_per_parflow_block: int = 8,
                 disable_gpu_compilation: bool = False, force_gpu_compilation: bool = False):
        Layer.__init__(self, nodes)
        nn.Module.__init__(self)
        assert len(self.nodes) > 0, "a sum layer needs at least one node vector"
        assert len(self.nodes) == len(set(self.nodes)), "a vector appears twice in one layer"

        nid0, pid0, pfid0 = global_nid_start, global_pid_start, global_pfid_start
        nblk, nedges, n_chs = get_sum_layer_forward_stats(self.nodes, global_nid_start)
        self.num_nodes = nblk * self.block_size
        self.num_edges = nedges

        fw_max = _unique_pow2(partition_nodes_by_n_edges(n_chs, sparsity_tolerance=layer_sparsity_tol,
                                                         max_num_partitions=max_num_partitions))
        self.num_fw_partitions = len(fw_max)
        part, loc, cnt = assign_partitions(n_chs, fw_max, start_min=1)
        nids, cids, pids, pfids, pid_end, pfid_end = sum_layer_forward_compilation(
            self.nodes, fw_max, part, loc, cnt, n_chs, global_nid_start, global_pid_start, global_pfid_start,
            node2tiednodes, max_tied_ns_per_parflow_block=max_tied_ns_per_parflow_block)
        self.partitioned_nids, self.partitioned_cids = nids, cids
        self.partitioned_pids, self.partitioned_pfids = pids, pfids
        self._fw_n_chs = [n_chs[part == p] for p in range(self.num_fw_partitions)]
        self._layer_nid_range = (nid0, nid0 + self.num_nodes)
        self._layer_pid_range = (pid0, pid_end)
        self._layer_pfid_range = (pfid0, pfid_end)

        ch_gsize2cs, ch_gsize2num_nblocks, ch_gsize2n_pargs, cs2parns = get_sum_layer_backward_stats(self.nodes)
        chids, parids, parpids, cs_bs, bk_npar = [], [], [], [], []
        for g, n_pargs in ch_gsize2n_pargs.items():
            bk_max = _unique_pow2(partition_nodes_by_n_edges(n_pargs, sparsity_tolerance=layer_sparsity_tol,
                                                             max_num_partitions=max_num_partitions))
            bpart, bloc, bcnt = assign_partitions(n_pargs, bk_max, start_min=1)
            c, pa, pp = sum_layer_backward_compilation(list(ch_gsize2cs[g]), cs2parns, bpart, bloc, bcnt, bk_max)
            chids.extend(c)
            parids.extend(pa)
            parpids.extend(pp)
            cs_bs.extend([g] * len(bk_max))
            for q in range(len(bk_max)):
                sel = (bpart == q) & (n_pargs >= 1) & (n_pargs <= int(bk_max[q]))
                bk_npar.append(n_pargs[sel])
        self.partitioned_chids, self.partitioned_parids, self.partitioned_parpids = chids, parids, parpids
        self.cs_block_sizes = cs_bs
        self._bk_n_pars = bk_npar
        self.num_bk_partitions = len(chids)
        self._pf_shared = any(ns.is_tied() for ns in self.nodes)
        self._dev = None

    @property
    def num_parameters(self):
        return self._layer_pid_range[1] - self._layer_pid_range[0]

    @property
    def num_param_flows(self):
        return self._layer_pfid_range[1] - self._layer_pfid_range[0]

    @property
    def nid_range(self):
        return self._layer_nid_range

    def is_sum(self):
        return True

    def __repr__(self):
        return f"SumLayer[{self._layer_nid_range[0]}:{self._layer_nid_range[1]}, {self.num_nodes} nodes, {self.num_edges} edges]"

    def _child_run(self, p: int) -> int:
        """Largest power of two R (<= 1024) such that every R-aligned group of real entries of
        partition p's index rows names consecutive element rows whose parameters are bs apart
        (cids[j, k0 + i] = cids[j, k0] + i, pids[j, k0 + i] = pids[j, k0] + i * bs); 0 if none.
        The compiler lays out each child block this way, so R is at least the child block size
        when all child blocks of the partition share it."""
        c = self.partitioned_cids[p].detach().cpu().numpy().astype(np.int64)
        q = self.partitioned_pids[p].detach().cpu().numpy().astype(np.int64)
        n = self._fw_n_chs[p].detach().cpu().numpy().astype(np.int64)
        K = c.shape[1]
        k = np.arange(K)
        valid = k[None, :] < n[:, None]
        for r in (1024, 512, 256, 128, 64, 32, 16, 8, 4, 2):
            if r > 2 * K and r > 2:
                continue
            base = (k // r) * r
            off = k - base
            ok = ((c - c[:, base]) == off[None, :]) & ((q - q[:, base]) == off[None, :] * self.block_size)
            if bool(np.all(ok | ~valid)):
                return r
        return 0

    def to(self, device):
        nn.Module.to(self, device)
        device = torch.device(device)
        for name in ("partitioned_nids", "partitioned_cids", "partitioned_pids", "partitioned_pfids",
                     "partitioned_chids", "partitioned_parids", "partitioned_parpids"):
            setattr(self, name, [t.to(device) for t in getattr(self, name)])
        self.device = device
        self._dev = None
        if device.type == "cuda":
            self.build_device_tables()
        return self

    # One launch per layer and pass: the partitions of a layer (and, in the backward, those of
    # one child block size) are concatenated into one table, padded to the widest partition.
    # Every kernel bounds its reduction by the block's own child / parent count, so the padding
    # costs index memory only, and blocks of uneven fan-in (pruned circuits) share one grid.
    merge_partitions = True
    # Forward launches of fewer than split_max_tiles tiles whose node blocks have at least
    # split_min_stages stages of 128 children split each reduction over workgroups
    # (pjamd_sum_forward_split; BF16 / FP16 operands, 16-byte layout).
    split_max_tiles = 128
    split_min_stages = 4

    def build_device_tables(self):
        """int32 launch tables (``_dev``) of the forward and backward partitions on the layer's
        device, with each forward table's child-run promise.  With ``merge_partitions`` a layer
        has one forward table and one backward table per child block size, rows ordered by
        decreasing fan-in so the longest tiles are dispatched first."""
        device = self.device
        self.__dict__.pop("_split_cache", None)
        # work-queue counters of pjamd_sum_forward_queue (zero between launches)
        self._queue = torch.zeros([2], dtype=torch.int32, device=device)
        fw, fw_run, fw_al = [], [], []
        for p in range(self.num_fw_partitions):
            n = self.partitioned_nids[p]
            if n.numel() == 0:
                continue
            # Real children of block j occupy columns [0, n_chs_j); the rest is padding.
            fw.append((n, self.partitioned_cids[p], self.partitioned_pids[p], self.partitioned_pfids[p],
                       self._fw_n_chs[p].to(device)))
            fw_run.append(self._child_run(p))
            fw_al.append(_params_aligned(self.partitioned_pids[p], self._fw_n_chs[p], self.block_size))
        bk, bk_al = [], []
        for q in range(self.num_bk_partitions):
            c = self.partitioned_chids[q]
            if c.numel() == 0:
                continue
            bk.append((int(self.cs_block_sizes[q]), c, self.partitioned_parids[q], self.partitioned_parpids[q],
                       self._bk_n_pars[q].to(device)))
            bk_al.append(_params_aligned(self.partitioned_parpids[q], self._bk_n_pars[q], self.block_size))
        if self.merge_partitions and len(fw) > 1:
            fw = [_merge_rows(fw, count_col=4)]
            fw_run = [min(fw_run) if all(r > 0 for r in fw_run) else 0]
            fw_al = [all(fw_al)]
        if self.merge_partitions and len(bk) > 1:
            by_cbs = dict()
            for entry, al in zip(bk, bk_al):
                by_cbs.setdefault(entry[0], []).append((entry, al))
            bk, bk_al = [], []
            for cbs, items in by_cbs.items():
                rows = [e[1:] for e, _ in items]
                bk.append((cbs,) + (_merge_rows(rows, count_col=3) if len(rows) > 1 else rows[0]))
                bk_al.append(all(al for _, al in items))
        self._fw_child_run, self._fw_aligned, self._bk_aligned = fw_run, fw_al, bk_al
        # distinct element rows read by each forward table and distinct parent node rows read by
        # each backward table: the algorithmic bytes count each row once per launch (node blocks
        # sharing children, or child blocks sharing parents, re-read them from cache)
        self._fw_uniq = [_distinct_rows(c, k, 1) for _n, c, _p, _f, k in fw]
        self._bk_uniq = [_distinct_rows(pa, k, self.block_size) for _cbs, _c, pa, _pp, k in bk]
        self._dev = {
            "fw": [(_i32(n), _i32(c), _i32(p), _i32(f), _i32(k), int(k.sum())) for n, c, p, f, k in fw],
            "bk": [(_i32(c), _i32(pa), _i32(pp), _i32(k), cbs, int(k.sum())) for cbs, c, pa, pp, k in bk],
        }

    # ------------------------------------------------------------------ passes

    def forward(self, node_mars: torch.Tensor, element_mars: torch.Tensor, params: torch.Tensor,
                force_use_bf16: bool = False, force_use_fp32: bool = False, propagation_alg: str = "LL",
                precision: Optional[str] = None, prod_tables: Optional[dict] = None, **kwargs):
        """``node_mars[n] = log sum_c theta[n, c] exp(element_mars[c])`` for every node of the
        layer (``sum_layer.py:220-262``), with per-tile max subtraction.  ``prod_tables`` (from
        ``TensorCircuit._build_prod_fusion``, keyed by layer): compute the element rows from the
        product layer's children in the same launch and store them (``pjamd_sum_forward_prod``).
        ``propagation_alg`` "MPE" / "GeneralLL" (with ``alpha``) run ``pjamd_sum_forward_alg``."""
        if propagation_alg != "LL":
            return self._forward_alg(node_mars, element_mars, params, propagation_alg, kwargs.get("alpha"))
        prec = PRECISIONS[precision] if precision is not None else (1 if force_use_bf16 else 0)
        if force_use_fp32:
            prec = 0
        B = node_mars.size(1)
        bs = self.block_size
        fused = prod_tables.get(id(self)) if prod_tables is not None else None
        # launch choice of this layer (set by pyjuice_amd.autotune.tune_launches): "pre" / "nopre"
        # force a form of the 64 x 64 16-byte-load kernels, "wide" the 128 x 128 one, "queue" runs
        # the persistent work-queue kernel,
        # "nosplit" turns the split reduction off, "split<N>" sets its split count, "valu" runs FP32
        # operands on the vector ALUs
        fh = (self.__dict__.get("launch_hints") or {}).get("fw")
        for e, ((nids, cids, pids, pfids, nch, tot), run, al) in enumerate(zip(self._dev["fw"], self._fw_child_run,
                                                                               self._fw_aligned)):
            run = run | (PARAMS_ALIGNED if al else 0)
            split = None if fh == "nosplit" else self._split_plan(e, nids.numel(), nch, run & ~PARAMS_ALIGNED, al,
                                                                     prec, B, self._forced_splits(fh))
            if split is not None and fused is None:
                nsplit, ws = split
                algo = (4 * (self._fw_uniq[e] + nids.numel() * bs) * B + 4 * tot * bs, 2 * tot * bs * B)
                N.call("pjamd_sum_forward_split", N.ptr(node_mars), N.ptr(element_mars), N.ptr(params), N.ptr(nids),
                       N.ptr(cids), N.ptr(pids), N.ptr(nch), nids.numel(), cids.size(1), bs, B, run, nsplit,
                       N.ptr(ws), prec, N.stream(), algo=algo)
                continue
            if fused is not None:
                tab, kp, reads, writes = fused[e]
                # product rows gathered and stored, node rows written, parameters read once
                algo = (4 * (reads + writes + nids.numel() * bs) * B + 4 * tot * bs, 2 * tot * bs * B)
                N.call("pjamd_sum_forward_prod", N.ptr(node_mars), N.ptr(element_mars), N.ptr(params), N.ptr(nids),
                       N.ptr(cids), N.ptr(pids), N.ptr(nch), nids.numel(), cids.size(1), bs, B, run, N.ptr(tab), kp,
                       prec, N.stream(), algo=algo)
                continue
            # algorithmic work: 2 flops per edge per sample; distinct child rows read + node rows
            # written per sample, parameters read once
            algo = (4 * (self._fw_uniq[e] + nids.numel() * bs) * B + 4 * tot * bs, 2 * tot * bs * B)
            if fh == "queue" and self.queue_ok(run, prec, B):
                N.call("pjamd_sum_forward_queue", N.ptr(node_mars), N.ptr(element_mars), N.ptr(params), N.ptr(nids),
                       N.ptr(cids), N.ptr(pids), N.ptr(nch), nids.numel(), cids.size(1), bs, B, run,
                       N.ptr(self._queue), prec, N.stream(), algo=algo)
                continue
            # "valu": FP32 operands on the vector-ALU tiles instead of the matrix cores (same arithmetic
            # type; faster for some shapes, e.g. 32-row blocks)
            N.call("pjamd_sum_forward", N.ptr(node_mars), N.ptr(element_mars), N.ptr(params), N.ptr(nids),
                   N.ptr(cids), N.ptr(pids), N.ptr(nch), nids.numel(), cids.size(1), bs, B,
                   run | _HINT_BITS.get(fh, 0), 3 if (fh == "valu" and prec == 0) else prec, N.stream(), algo=algo)

    def queue_ok(self, run: int, prec: int, B: int) -> bool:
        """Whether ``pjamd_sum_forward_queue`` takes a forward table with child-run flags ``run``."""
        r = run & ~PARAMS_ALIGNED
        bs = self.block_size
        return (prec in (1, 2) and bool(run & PARAMS_ALIGNED) and r >= 16 and r % 16 == 0 and B % 4 == 0
                and (bs == 32 or (bs >= 64 and bs % 64 == 0)))

    @staticmethod
    def _forced_splits(hint: Optional[str]) -> Optional[int]:
        """The split count a launch hint "split<N>" asks for (None for other hints)."""
        return int(hint[5:]) if isinstance(hint, str) and hint.startswith("split") else None

    def _split_plan(self, e: int, n_blocks: int, nch: torch.Tensor, run: int, aligned: bool, prec: int, B: int,
                    force: Optional[int] = None):
        """``(n_splits, workspace)`` for forward table ``e`` at batch ``B`` when the split reduction
        applies, else None; the workspace is cached per (table, batch).  ``force``: this split count
        (2 .. 16, at most one per stage) instead of the default rule's."""
        bs = self.block_size
        if not (prec in (1, 2) and aligned and run >= 128 and run % 128 == 0 and bs % 64 == 0 and B % 4 == 0):
            return None
        tiles = n_blocks * (bs // 64) * -(-B // 64)
        key = (e, B, force)
        cache = self.__dict__.setdefault("_split_cache", dict())
        if key not in cache:
            nst = -(-int(nch.max()) // 128)
            plan = None
            if tiles < self.split_max_tiles and nst >= self.split_min_stages and n_blocks * bs < 65536:
                nsplit = max(2, min(nst // 2, 256 // max(tiles, 1), 16)) if force is None else min(force, nst)
                rows = n_blocks * bs
                ws = _split_workspace(nsplit * rows * B + nsplit * (rows // 64) * B, nch.device)
                plan = (nsplit, ws)
            cache[key] = plan
        return cache[key]

    def _split_plan_bk(self, q: int, n_cblocks: int, npar: torch.Tensor, cbs: int, aligned: bool, prec: int, B: int,
                       force: Optional[int] = None):
        """``(n_splits, workspace)`` for the fused element-flow pass of backward table ``q`` at
        batch ``B`` when the split reduction applies (parent blocks of >= 128 rows), else None;
        ``force`` as for ``_split_plan``."""
        bs = self.block_size
        if not (prec in (1, 2) and aligned and bs >= 128 and cbs % 64 == 0 and B % 4 == 0):
            return None
        tiles = n_cblocks * (cbs // 64) * -(-B // 64)
        key = ("bk", q, B, force)
        cache = self.__dict__.setdefault("_split_cache", dict())
        if key not in cache:
            nst = int(npar.max()) * bs // 128
            plan = None
            rows = n_cblocks * cbs
            if tiles < self.split_max_tiles and nst >= self.split_min_stages and rows < 65536:
                nsplit = max(2, min(nst // 2, 256 // max(tiles, 1), 16)) if force is None else min(force, nst)
                ws = _split_workspace(nsplit * rows * B + nsplit * (rows // 64) * B, npar.device)
                plan = (nsplit, ws)
            cache[key] = plan
        return cache[key]

    @staticmethod
    def _alg_id(propagation_alg: str, alpha: Optional[float]):
        """(alg code, alpha) of ``Layer.propagation_alg_mapping`` (``layer.py:11-15, 89-97``)."""
        if propagation_alg == "MPE":
            return 1, 0.0
        if propagation_alg == "GeneralLL":
            if alpha is None:
                raise KeyError("alpha")
            if not alpha > 0.0:
                raise ValueError("GeneralLL needs alpha > 0")
            return 2, float(alpha)
        if propagation_alg == "LL":
            return 0, 1.0
        raise ValueError(f"propagation_alg {propagation_alg!r} is not one of LL, MPE, GeneralLL")

    def _forward_alg(self, node_mars, element_mars, params, propagation_alg, alpha):
        alg, alpha = self._alg_id(propagation_alg, alpha)
        B = node_mars.size(1)
        bs = self.block_size
        for nids, cids, pids, pfids, nch, tot in self._dev["fw"]:
            N.call("pjamd_sum_forward_alg", N.ptr(node_mars), N.ptr(element_mars), N.ptr(params), N.ptr(nids),
                   N.ptr(cids), N.ptr(pids), N.ptr(nch), nids.numel(), cids.size(1), bs, B, alg, alpha, N.stream(),
                   algo=(4 * (tot + nids.numel() * bs) * B + 4 * tot * bs, 2 * tot * bs * B))

    def _backward_alg(self, node_flows, element_flows, node_mars, element_mars, params, param_flows,
                      propagation_alg, alpha, accumulate, negate=False):
        """Element flows into ``element_flows`` (then the product layer's backward) and parameter
        flows (subtracted with ``negate``) for MPE / GeneralLL; node flows are read as stored."""
        alg, alpha = self._alg_id(propagation_alg, alpha)
        B = node_flows.size(1)
        bs = self.block_size
        s = N.stream()
        for chids, parids, parpids, npar, cbs, totp in self._dev["bk"]:
            N.call("pjamd_sum_backward_ele_alg", N.ptr(element_flows), N.ptr(node_flows), N.ptr(node_mars),
                   N.ptr(element_mars), N.ptr(params), N.ptr(chids), N.ptr(parids), N.ptr(parpids), N.ptr(npar),
                   chids.numel(), parids.size(1), cbs, bs, B, alg, alpha, 1 if accumulate else 0, s,
                   algo=(4 * (2 * totp * bs + 2 * chids.numel() * cbs) * B + 4 * totp * bs * cbs, 0))
        if param_flows is not None:
            pf64 = 1 if param_flows.dtype == torch.float64 else 0
            for nids, cids, pids, pfids, nch, tot in self._dev["fw"]:
                N.call("pjamd_sum_backward_par_alg", N.ptr(param_flows), pf64, N.ptr(node_flows), N.ptr(node_mars),
                       N.ptr(element_mars), N.ptr(params), N.ptr(nids), N.ptr(cids), N.ptr(pids), N.ptr(pfids),
                       N.ptr(nch), nids.numel(), cids.size(1), bs, B, alg | (ALG_NEGATE if negate else 0), s,
                       algo=(4 * (2 * nids.numel() * bs + tot) * B + 4 * tot * bs, 0))

    def backward(self, node_flows: torch.Tensor, element_flows: torch.Tensor, node_mars: torch.Tensor,
                 element_mars: torch.Tensor, params: torch.Tensor, param_flows: Optional[torch.Tensor] = None,
                 allow_modify_flows: bool = False, propagation_alg: str = "LL", logspace_flows: bool = False,
                 negate_pflows: bool = False, accumulate_ch_flows: bool = False, allow_neg_flows: bool = False,
                 force_use_fp32: bool = False, precision: Optional[str] = None, premodified: bool = False,
                 child_tables: Optional[Sequence[tuple]] = None, overwrite_pflows: bool = False,
                 **kwargs):
        """Element flows and (optionally) parameter flows of the layer (``sum_layer.py:264-384``).

        ``allow_modify_flows`` rewrites the layer's node flows in place as ``log(flow) - mar``
        (``sum_layer.py:1381-1392``); the element / parameter-flow kernels read either form.
        ``premodified``: the flows already arrive in that form (written by the fused kernel of the
        layer above).  ``child_tables`` (one ``(table, child blocks, of which modified)`` per
        backward partition, see ``pjamd_sum_backward_ele_to_children``): write the flows of the
        product layer's children directly into ``node_flows`` instead of element flows.
        """
        # logspace_flows is handled by TensorCircuit.backward (flows propagate in linear space and
        # the node-flow buffer is converted at the end), so a layer only sees linear flows.
        assert not logspace_flows and not allow_neg_flows, \
            "the layer kernels take linear, non-negative flows (TensorCircuit.backward converts log-space flows)"
        if propagation_alg != "LL":
            return self._backward_alg(node_flows, element_flows, node_mars, element_mars, params, param_flows,
                                      propagation_alg, kwargs.get("alpha"), accumulate_ch_flows, negate_pflows)
        assert not (accumulate_ch_flows and allow_modify_flows)
        prec = PRECISIONS[precision] if precision is not None else 0
        if force_use_fp32:
            prec = 0
        B = node_flows.size(1)
        s = N.stream()
        bs = self.block_size
        if allow_modify_flows and not premodified:
            for nids, cids, pids, pfids, nch, tot in self._dev["fw"]:
                N.call("pjamd_sum_modify_flows", N.ptr(node_flows), N.ptr(node_mars), N.ptr(nids), nids.numel(),
                       bs, B, s, algo=(12 * nids.numel() * bs * B, 0))
        mod = 1 if allow_modify_flows else 0
        for q, (chids, parids, parpids, npar, cbs, totp) in enumerate(self._dev["bk"]):
            al = 4 if self._bk_aligned[q] else 0
            if child_tables is not None:
                tab, n_blk, n_mod = child_tables[q]
                bk_choice = (self.__dict__.get("launch_hints") or {}).get("bk")
                split = None if bk_choice == "nosplit" else self._split_plan_bk(
                    q, chids.numel(), npar, cbs, bool(al), prec, B, self._forced_splits(bk_choice))
                # distinct parent rows (and marginals unless modified) + element rows read, one write
                # per child row (+ its marginal when stored as log(flow) - mar), parameters once
                n_rd = n_mod if mod else 0
                algo = (4 * (self._bk_uniq[q] * (1 if mod else 2) + chids.numel() * cbs + (n_blk + n_rd) * cbs) * B
                        + 4 * totp * bs * cbs, 2 * totp * bs * cbs * B)
                if split is not None:
                    nsplit, ws = split
                    N.call("pjamd_sum_backward_ele_split", N.ptr(node_flows), N.ptr(node_mars), N.ptr(element_mars),
                           N.ptr(params), N.ptr(chids), N.ptr(parids), N.ptr(parpids), N.ptr(npar), N.ptr(tab),
                           tab.size(1), chids.numel(), parids.size(1), cbs, bs, B,
                           mod | (2 if allow_modify_flows else 0) | al, nsplit, N.ptr(ws), prec, s, algo=algo)
                    continue
                bh = _HINT_BITS.get((self.__dict__.get("launch_hints") or {}).get("bk"), 0)
                N.call("pjamd_sum_backward_ele_to_children", N.ptr(node_flows), N.ptr(node_mars),
                       N.ptr(element_mars), N.ptr(params), N.ptr(chids), N.ptr(parids), N.ptr(parpids),
                       N.ptr(npar), N.ptr(tab), tab.size(1), chids.numel(), parids.size(1), cbs, bs, B,
                       mod | (2 if allow_modify_flows else 0) | al | bh, prec, s, algo=algo)
                continue
            algo = (4 * (self._bk_uniq[q] * (1 if mod else 2) + 2 * chids.numel() * cbs) * B + 4 * totp * bs * cbs,
                    2 * totp * bs * cbs * B)
            bk_choice = (self.__dict__.get("launch_hints") or {}).get("bk")
            N.call("pjamd_sum_backward_ele", N.ptr(element_flows), N.ptr(node_flows), N.ptr(node_mars),
                   N.ptr(element_mars), N.ptr(params), N.ptr(chids), N.ptr(parids), N.ptr(parpids), N.ptr(npar),
                   chids.numel(), parids.size(1), cbs, bs, B, mod | (2 if accumulate_ch_flows else 0) | al,
                   3 if (bk_choice == "valu" and prec == 0) else prec, s, algo=algo)
        if param_flows is not None:
            pf64 = 1 if param_flows.dtype == torch.float64 else 0
            flags = mod | (2 if self._pf_shared else 0) | (4 if negate_pflows else 0) | (8 * pf64)
            if overwrite_pflows and not self._pf_shared:
                flags |= 16
            pfw = 8 if pf64 else 4
            for e, (nids, cids, pids, pfids, nch, tot) in enumerate(self._dev["fw"]):
                algo = (4 * (nids.numel() * bs * (1 if mod else 2) + self._fw_uniq[e]) * B + (4 + 2 * pfw) * tot * bs,
                        2 * tot * bs * B)
                N.call("pjamd_sum_backward_par", N.ptr(param_flows), N.ptr(node_flows), N.ptr(node_mars),
                       N.ptr(element_mars), N.ptr(params), N.ptr(nids), N.ptr(cids), N.ptr(pids), N.ptr(pfids),
                       N.ptr(nch), nids.numel(), cids.size(1), bs, B, flags, prec, s, algo=algo)
