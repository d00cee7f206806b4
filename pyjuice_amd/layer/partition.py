"""Fan-in partitioner: groups node blocks into <= 16 partitions by number of children.

Restates ``pyjuice/layer/backend/node_partition.py:98-227`` (the numba DP) with numpy so that
the chosen partition maxima are bit-identical, including its corner cases:

* int64 DP whose inner minimum starts at the sentinel ``2**31 - 1`` (``node_partition.py:134``);
  if no candidate beats it the back-pointer is ``-1``;
* coalescing of unique fan-ins within ``floor((max - min) / 1000)`` of the run start,
  keeping the run's last value (``node_partition.py:69-115``);
* the back-trace reads column ``target_n_partition`` at every step
  (``node_partition.py:44-49``) and resolves a ``-1`` pointer like numpy negative indexing.
"""
from __future__ import annotations

import math
from typing import Optional, Union

import numpy as np
import torch

__all__ = ["partition_nodes_by_n_edges", "next_power_of_2"]

_SENTINEL = 2 ** 31 - 1


def next_power_of_2(x: Union[np.ndarray, torch.Tensor]) -> np.ndarray:
    """Element-wise smallest power of two >= x (1 stays 1), as ``compilation.py:93-98``."""
    x = np.asarray(x.cpu().numpy() if isinstance(x, torch.Tensor) else x, dtype=np.int64)
    out = np.ones_like(x)
    big = x > 1
    out[big] = np.left_shift(1, np.array([int(v - 1).bit_length() for v in x[big]], dtype=np.int64))
    return out


def _coalesce(vals: np.ndarray):
    uvals, counts = np.unique(vals, return_counts=True)
    tol_range = int(math.floor((uvals.max() - uvals.min()) / 1000))
    keep_vals, keep_counts = [], []
    run_first, run_count = None, 0
    for v, c in zip(uvals.tolist(), counts.tolist()):
        if run_first is None:
            run_first, run_count, last = v, c, v
        elif v - run_first <= tol_range:
            run_count += c
            last = v
        else:
            keep_vals.append(last)
            keep_counts.append(run_count)
            run_first, run_count, last = v, c, v
    if run_first is not None:
        keep_vals.append(last)
        keep_counts.append(run_count)
    return np.array(keep_vals, dtype=np.int64), np.array(keep_counts, dtype=np.int64)


def _backtrace(bt: np.ndarray, T: int, n: int) -> np.ndarray:
    parts = np.zeros([T], dtype=np.int64)
    i = n - 1
    for k in range(T, 0, -1):
        parts[k - 1] = i
        i = int(bt[i, T])
    return parts


def _weighted_dp(vals: np.ndarray, counts: np.ndarray, P: int, target: int):
    n = vals.shape[0]
    cum = np.cumsum(counts).astype(np.int64)
    dp = np.zeros([n, P + 1], dtype=np.int64)
    bt = np.zeros([n, P + 1], dtype=np.int64)
    dp[:, 1] = vals * cum
    if dp[-1, 1] <= target:
        T = 1
    else:
        T = P
        for k in range(2, P + 1):
            dp[0, k] = vals[0] * cum[0]
            bt[0, k] = 0
            prev = dp[:, k - 1]
            for i in range(1, n):
                cand = prev[:i] + vals[i] * (cum[i] - cum[:i])
                j = int(np.argmin(cand))
                if cand[j] < _SENTINEL:
                    dp[i, k] = cand[j]
                    bt[i, k] = j
                else:
                    dp[i, k] = _SENTINEL
                    bt[i, k] = -1
            if dp[-1, k] <= target:
                T = k
                break
    parts = _backtrace(bt, T, n)
    return np.unique(vals[parts]), int(dp[-1, T])


def _simple_dp(vals: np.ndarray, P: int, target: int):
    n = vals.shape[0]
    dp = np.zeros([n, P + 1], dtype=np.int64)
    bt = np.zeros([n, P + 1], dtype=np.int64)
    dp[:, 1] = vals * (np.arange(n, dtype=np.int64) + 1)
    T = P
    for k in range(2, P + 1):
        dp[0, k] = vals[0]
        bt[0, k] = 0
        prev = dp[:, k - 1]
        for i in range(1, n):
            cand = prev[:i] + vals[i] * (i - np.arange(i, dtype=np.int64))
            j = int(np.argmin(cand))
            if cand[j] < _SENTINEL:
                dp[i, k] = cand[j]
                bt[i, k] = j
            else:
                dp[i, k] = _SENTINEL
                bt[i, k] = -1
        if dp[-1, k] < target:
            T = k
            break
    parts = _backtrace(bt, T, n)
    return np.unique(vals[parts]), int(dp[-1, T])


def partition_nodes_by_n_edges(node_n_edges, max_num_partitions: Optional[int] = None,
                               sparsity_tolerance: Optional[float] = None,
                               algorithm: str = "dp_with_coalesce") -> torch.Tensor:
    """Sorted partition maxima (before power-of-two rounding).  Same contract as
    ``node_partition.py:179-227``."""
    x = node_n_edges.detach().cpu().numpy() if isinstance(node_n_edges, torch.Tensor) else np.asarray(node_n_edges)
    x = x.astype(np.int64)
    if sparsity_tolerance is not None and sparsity_tolerance < 1e-6:
        sparsity_tolerance = None
        max_num_partitions = 1
    if sparsity_tolerance is not None:
        assert 1e-6 < sparsity_tolerance <= 1.0
        if max_num_partitions is None:
            max_num_partitions = max(min(np.unique(x).shape[0], 16), 1)
    elif max_num_partitions is None:
        max_num_partitions = 1
    else:
        assert max_num_partitions >= 1, "the partitioner needs at least one partition"

    target = None if sparsity_tolerance is None else int(math.ceil(x.sum() * (1.0 + sparsity_tolerance)))
    if max_num_partitions == 1:
        return torch.tensor([int(x.max())], dtype=torch.long)

    x = np.sort(x)
    if x[0] == 0:
        x = x[(x == 0).sum():]
    tgt = 0 if target is None else target
    if algorithm == "dp_simple":
        sizes, _ = _simple_dp(x, max_num_partitions, tgt)
    elif algorithm == "dp_with_coalesce":
        vals, counts = _coalesce(x)
        sizes, _ = _weighted_dp(vals, counts, max_num_partitions, tgt)
    else:
        raise ValueError(f"partitioning algorithm {algorithm!r} is not implemented")
    return torch.sort(torch.from_numpy(np.asarray(sizes, dtype=np.int64))).values
