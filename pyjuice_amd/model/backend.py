"""Compile-time tables for mini-batch EM and tied parameter-flow fusing.

* ``compile_par_update_fn``: the EM row table of ``pyjuice/model/backend/par_update.py:15-154``
  (one row per (sum node, chunk of <= BLOCK_SIZE children), in DFS post-order of the untied
  sum vectors), plus the per-node-block table the HIP EM kernel walks
  (``[par_start, pflow_start, node_start, block_size, n_children]``).
* ``compile_cum_par_flows_fn``: groups of flow ranges of tied sum vectors that are summed
  into the source range before EM (``pyjuice/model/backend/parflow_fusing.py:11-62``).
* ``normalize_parameters``: per-node normalisation of the flat parameter vector
  (``pyjuice/model/backend/normalize.py:103-154``).
"""
from __future__ import annotations

from typing import List, Sequence

import numpy as np
import torch

__all__ = ["compile_par_update_fn", "compile_cum_par_flows_fn", "normalize_parameters", "sum_param_node_map"]


def _untied_sums(root_ns):
    return [ns for ns in root_ns if ns.is_sum() and not ns.is_tied()]


def compile_par_update_fn(root_ns, BLOCK_SIZE: int = 32):
    assert BLOCK_SIZE & (BLOCK_SIZE - 1) == 0, "the EM row width is not a power of two"
    cols = {k: [] for k in ("par", "pf", "bsz", "itv", "gnid", "nch")}
    blocks = []
    gnid = 0
    for ns in _untied_sums(root_ns):
        bs, cbs = ns.block_size, ns.ch_block_size
        par, pf = ns._param_range[0], ns._param_flow_range[0]
        ne = np.bincount(ns.edge_ids[0, :].numpy(), minlength=ns.num_node_blocks).astype(np.int64)
        for g in range(ns.num_node_blocks):
            nch = int(ne[g]) * cbs
            blocks.append((par, pf, gnid, bs, nch))
            sids = np.arange(0, nch, BLOCK_SIZE, dtype=np.int64)
            bsz = np.minimum(sids + BLOCK_SIZE, nch) - sids
            gid = np.arange(bs, dtype=np.int64)
            cols["par"].append((par + sids[:, None] * bs + gid[None, :]).reshape(-1))
            cols["pf"].append((pf + sids[:, None] * bs + gid[None, :]).reshape(-1))
            cols["bsz"].append(np.repeat(bsz, bs))
            cols["itv"].append(np.full(sids.shape[0] * bs, bs, dtype=np.int64))
            cols["gnid"].append(np.tile(gnid + gid, sids.shape[0]))
            cols["nch"].append(np.full(sids.shape[0] * bs, nch, dtype=np.int64))
            par += bs * nch
            pf += bs * nch
            gnid += bs
    cat = lambda k: torch.from_numpy(np.concatenate(cols[k]) if cols[k] else np.zeros(0, dtype=np.int64))
    par_start_ids, pflow_start_ids = cat("par"), cat("pf")
    blk_sizes, blk_intervals, global_nids, nchs = cat("bsz"), cat("itv"), cat("gnid"), cat("nch")
    tot = int(global_nids[-1].item()) + 1 if global_nids.numel() > 0 else 0
    cum_pflows = torch.zeros([tot], dtype=torch.float32)
    metadata = {"tot_num_nodes": tot, "BLOCK_SIZE": BLOCK_SIZE,
                "node_blocks": torch.tensor(blocks, dtype=torch.int64).reshape(-1, 5)}
    return [par_start_ids, pflow_start_ids, blk_sizes, blk_intervals, global_nids, nchs, cum_pflows, metadata]


def compile_cum_par_flows_fn(node2tiednodes: dict, MAX_NBLOCKS: int = 2048, BLOCK_SIZE: int = 2048):
    """List of ``[target_pfid_start, num_par_flows, [copy pfid starts]]`` for every source with
    more than one flow copy (the copies include the source's own range)."""
    groups = []
    for src, item in node2tiednodes.items():
        if len(item[0]) > 1:
            assert len(item[0]) <= MAX_NBLOCKS, \
                f"We only support fusing at most {MAX_NBLOCKS} blocks for parameter flow accumulation."
            n = src._param_flow_range[1] - src._param_flow_range[0]
            groups.append([src._param_flow_range[0], n, [c._param_flow_range[0] for c in item[0]]])
    return groups


def sum_param_node_map(node_blocks: torch.Tensor, num_sum_params: int) -> np.ndarray:
    """Global EM node index of every flat sum parameter (-1 for the dummy prefix)."""
    owner = np.full(num_sum_params, -1, dtype=np.int64)
    for par, _, gnid, bs, nch in node_blocks.tolist():
        idx = par + np.arange(nch * bs, dtype=np.int64)
        owner[idx] = gnid + (np.arange(nch * bs, dtype=np.int64) % bs)
    return owner


def normalize_parameters(params: torch.Tensor, par_update_kwargs: Sequence, pseudocount: float = 0.0):
    """``theta <- (theta + pc/nch) / (sum_node theta + pc)`` per sum node, in place (CPU)."""
    meta = par_update_kwargs[7]
    blocks = meta["node_blocks"]
    tot = meta["tot_num_nodes"]
    if blocks.numel() == 0:
        return params
    owner = torch.from_numpy(sum_param_node_map(blocks, params.numel()))
    nch_node = torch.zeros([tot], dtype=torch.float32)
    for par, _, gnid, bs, nch in blocks.tolist():
        nch_node[gnid:gnid + bs] = float(nch)
    with torch.no_grad():
        p = params.data if isinstance(params, torch.nn.Parameter) else params
        sel = owner >= 0
        o = owner[sel]
        sums = torch.zeros([tot], dtype=torch.float32)
        sums.index_add_(0, o, p[sel].float())
        p[sel] = (p[sel] + pseudocount / nch_node[o]) / (sums[o] + pseudocount)
    return params
