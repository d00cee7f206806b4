"""Seeded random pruning of sum-edge blocks (the uneven fan-in workload).

Keeps the semantics of ``pyjuice/transformations/prune.py:13-105`` for a random score: every
parent block keeps at least one edge block."""
from __future__ import annotations

import numpy as np
import torch

__all__ = ["prune_sum_blocks"]


def prune_sum_blocks(root_ns, frac: float, seed: int = 4, skip_root: bool = True):
    """Drop ``frac`` of the edge blocks of every sum vector (in DFS post-order), chosen by
    ``numpy.random.Generator(PCG64(seed))`` scores; returns the number of blocks removed."""
    assert 0.0 <= frac < 1.0
    rng = np.random.Generator(np.random.PCG64(seed))
    removed = 0
    for ns in root_ns:
        if not ns.is_sum() or ns.is_tied() or (skip_root and ns is root_ns):
            continue
        E = ns.edge_ids.numpy()
        n = E.shape[1]
        score = rng.random(n)
        ndrop = int(np.floor(frac * n))
        keep = np.ones(n, dtype=bool)
        if ndrop > 0:
            keep[np.argsort(score, kind="stable")[:ndrop]] = False
        for b in range(ns.num_node_blocks):
            sel = np.nonzero(E[0] == b)[0]
            if sel.size > 0 and not keep[sel].any():
                keep[sel[np.argmax(score[sel])]] = True
        removed += int((~keep).sum())
        old = ns._params
        ns.set_edges(torch.from_numpy(E[:, keep].copy()))
        if old is not None:
            ns._params = old[torch.from_numpy(keep)].clone()
    return removed
