"""Runs of narrow inner layers executed as one launch (``pjamd_chain_forward`` /
``pjamd_chain_backward``).

A circuit pass treats every batch column independently, so a run of consecutive layers can be
split over the batch: each workgroup owns 32 columns and computes every node row of every layer
of the run in order, with only workgroup barriers in between.  This pays off for the long tails
of narrow layers that tree-shaped circuits (HCLT, HMM) end in, where one launch per layer leaves
most of the GPU idle: there a layer costs a launch plus several dependent memory round trips,
while inside a run it costs the time one workgroup needs to stream the layer's parameters.

A (product group, sum group) pair of ``TensorCircuit.inner_layer_groups`` may join a run when

* every sum layer of the pair has a power-of-two block size >= 32, and each of its partitions
  fits what a chain operation stages in LDS (``max_k`` entries: the partition's index rows, and
  the reductions of all its blocks, children per node block forward or parent rows per child
  block backward, each rounded up to 64; at most 64 blocks);
* the pair's sum parameters take at most ``max_param_bytes`` (every workgroup streams all of
  them, so wide layers are better served by one tiled launch each);
* backward only: the product group is fused into the sum layers' element-flow pass (every child
  has a single parent) and the sum layers' parameter flows come from the batched launch after
  the layer loop.

Runs are maximal sequences of at least ``min_pairs`` consecutive eligible pairs.  The reference
runs one launch per layer and group (``tensorcircuit.py:269-283``, ``:436-470``).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, List, Optional

import torch

OP_FIELDS = 16          # PJAMD_CHAIN_OP_FIELDS
MAX_K = 1024            # PJAMD_CHAIN_MAX_K (BF16 / FP16 operands; half of it with FP32 operands)
MAX_BLOCKS = 64         # blocks per sum operation
CHUNK = 64              # reductions are staged in multiples of this
KIND_PROD, KIND_SUM_FWD, KIND_SUM_ELE, KIND_MODIFY = 0, 1, 2, 3


def max_k(precision: str) -> int:
    return MAX_K // 2 if precision == "fp32" else MAX_K


@dataclass
class Run:
    """One chain launch: inner layer groups ``[lo, hi)``; ``ops`` int64 ``[n_ops, OP_FIELDS]`` and
    ``idx`` int32 on the device; ``bytes_per_col`` / ``flops_per_col`` the algorithmic work per
    batch column and ``param_bytes`` the parameters read once."""
    lo: int
    hi: int
    ops: torch.Tensor
    idx: torch.Tensor
    n_ops: int
    bytes_per_col: int
    param_bytes: int
    flops_per_col: int

    def algo(self, B: int):
        return (self.bytes_per_col * B + self.param_bytes, self.flops_per_col * B)


class _Builder:
    def __init__(self):
        self.ops: List[List[int]] = []
        self.chunks: List[torch.Tensor] = []
        self.size = 0
        self.bytes_per_col = 0
        self.param_bytes = 0
        self.flops_per_col = 0

    def table(self, t: torch.Tensor) -> int:
        off = self.size
        t = t.detach().reshape(-1).to(torch.int32)
        self.chunks.append(t)
        self.size += t.numel()
        return off

    def op(self, kind, nb, K, bs, cbs=0, eoff=0, a=0, b=0, c=0, d=0, e=0, Kc=0, flags=0):
        row = [kind, nb, K, bs, cbs, eoff, a, b, c, d, e, Kc, flags]
        self.ops.append(row + [0] * (OP_FIELDS - len(row)))

    def finish(self, lo, hi, device) -> Run:
        idx = torch.cat(self.chunks).to(device).contiguous() if self.chunks else \
            torch.zeros([1], dtype=torch.int32, device=device)
        ops = torch.tensor(self.ops, dtype=torch.int64).to(device).contiguous()
        return Run(lo, hi, ops, idx, len(self.ops), self.bytes_per_col, self.param_bytes, self.flops_per_col)


def _pow2(v: int) -> bool:
    return v > 0 and (v & (v - 1)) == 0


def _staged(lengths: torch.Tensor) -> int:
    """LDS entries of a sum operation's B operand: each block's reduction rounded up to CHUNK,
    at least one CHUNK per block."""
    return int(((lengths.to(torch.int64) + CHUNK - 1) // CHUNK).clamp(min=1).sum()) * CHUNK


def _sum_layer_ok(layer, precision: str, backward: bool) -> bool:
    bs = layer.block_size
    if bs < 32 or not _pow2(bs) or layer._dev is None:
        return False
    cap = max_k(precision)
    for nids, cids, pids, pfids, nch, tot in layer._dev["fw"]:
        if nids.numel() > MAX_BLOCKS or cids.numel() > cap or _staged(nch) > cap:
            return False
    if backward:
        for chids, parids, parpids, npar, cbs, totp in layer._dev["bk"]:
            if cbs < 32 or cbs % 32 != 0 or chids.numel() > MAX_BLOCKS or parids.numel() > cap or \
                    _staged(npar * bs) > cap:
                return False
    return True


def _prod_group_ok(group, precision: str) -> bool:
    cap = max_k(precision)
    return all(layer._dev is not None and all(cids.numel() <= cap for nids, cids, words in layer._dev["fw"])
               for layer in group)


def _pair_param_bytes(group) -> int:
    return sum(4 * tot * layer.block_size for layer in group for *_, tot in layer._dev["fw"])


def plan(pc, direction: str, precision: str, modify: bool, compute_param_flows: bool,
         max_param_bytes: int, min_pairs: int = 2) -> Dict[int, Run]:
    """Runs of ``pc`` for one pass.  Forward runs are keyed by their first (product) group id,
    backward runs by their last (top sum) group id; each covers the groups ``[lo, hi)``."""
    groups = pc.inner_layer_groups
    device = pc.device
    fz = pc._fused
    backward = direction == "bwd"
    batched = pc._par_batches["layers"]

    def eligible(p: int) -> bool:
        if p + 1 >= len(groups) or not groups[p].is_prod() or groups[p + 1].is_prod():
            return False
        g = groups[p + 1]
        if not all(_sum_layer_ok(layer, precision, backward) for layer in g):
            return False
        if not backward and not _prod_group_ok(groups[p], precision):
            return False
        if _pair_param_bytes(g) > max_param_bytes:
            return False
        if backward:
            if p not in fz["prod_groups"]:
                return False
            for layer in g:
                if fz["tables"].get(id(layer)) is None:
                    return False
                if compute_param_flows and id(layer) not in batched:
                    return False
        return True

    runs: Dict[int, Run] = dict()
    p = 0
    while p < len(groups):
        if not eligible(p):
            p += 1
            continue
        q = p
        while eligible(q):
            q += 2
        if (q - p) // 2 >= min_pairs:
            run = (_backward_run if backward else _forward_run)(pc, p, q, modify, device)
            runs[q - 1 if backward else p] = run
        p = q
    return runs


def _forward_run(pc, lo: int, hi: int, modify: bool, device) -> Run:
    bld = _Builder()
    groups = pc.inner_layer_groups
    for gid in range(lo, hi):
        eoff = pc._elem_offsets[gid // 2]
        for layer in groups[gid]:
            if groups[gid].is_prod():
                for nids, cids, words in layer._dev["fw"]:
                    bld.op(KIND_PROD, nids.numel(), cids.size(1), layer.block_size, eoff=eoff,
                           a=bld.table(nids), b=bld.table(cids))
                    bld.bytes_per_col += 4 * words
            else:
                bs = layer.block_size
                for (nids, cids, pids, pfids, nch, tot), run in zip(layer._dev["fw"], layer._fw_child_run):
                    # bit 2: aligned groups of 8 children are runs (one parameter index per 8)
                    bld.op(KIND_SUM_FWD, nids.numel(), cids.size(1), bs, eoff=eoff, a=bld.table(nids),
                           b=bld.table(cids), c=bld.table(pids), d=bld.table(nch), flags=4 if run >= 8 else 0)
                    bld.bytes_per_col += 4 * (tot + nids.numel() * bs)
                    bld.param_bytes += 4 * tot * bs
                    bld.flops_per_col += 2 * tot * bs
    return bld.finish(lo, hi, device)


def _backward_run(pc, lo: int, hi: int, modify: bool, device) -> Run:
    bld = _Builder()
    groups = pc.inner_layer_groups
    fz = pc._fused
    flags = 3 if modify else 0
    for gid in range(hi - 1, lo - 1, -1):
        if groups[gid].is_prod():
            continue   # performed by the sum layers' fused element-flow operations
        eoff = pc._elem_offsets[gid // 2]
        for layer in groups[gid]:
            bs = layer.block_size
            if modify and id(layer) not in fz["premod"]:
                for nids, cids, pids, pfids, nch, tot in layer._dev["fw"]:
                    bld.op(KIND_MODIFY, nids.numel(), 0, bs, a=bld.table(nids))
                    bld.bytes_per_col += 12 * nids.numel() * bs
            for (chids, parids, parpids, npar, cbs, totp), (tab, n_blk, n_mod) in zip(
                    layer._dev["bk"], fz["tables"][id(layer)]):
                # bit 2: every parameter run starts 16-byte aligned (vector loads of 4 parameters)
                aligned = bool(((parpids.to(torch.int64) % 4) == 0).all()) and bs % 4 == 0
                bld.op(KIND_SUM_ELE, chids.numel(), parids.size(1), bs, cbs=cbs, eoff=eoff, a=bld.table(chids),
                       b=bld.table(parids), c=bld.table(parpids), d=bld.table(npar), e=bld.table(tab),
                       Kc=tab.size(1), flags=flags | (4 if aligned else 0))
                n_rd = n_mod if modify else 0
                bld.bytes_per_col += 4 * (totp * bs * (1 if modify else 2) + chids.numel() * cbs + (n_blk + n_rd) * cbs)
                bld.param_bytes += 4 * totp * bs * cbs
                bld.flops_per_col += 2 * totp * bs * cbs
    return bld.finish(lo, hi, device)
