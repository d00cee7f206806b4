"""FP64 reference evaluation on the node DAG (TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py).

Independent of the compiled layout: every vector of nodes is evaluated from its own
``edge_ids`` and parameters.  Parameters are read from the DAG (``ns._params``; a tied vector
uses its source's), so a caller first copies the compiled circuit's parameters back with
``pc.update_parameters()``.

Maths (reference citations):
* sum:   ``v[n] = log sum_c W[n, c] exp(v_c)``                       (sum_layer.py:454-568)
* prod:  ``v[n] = sum_k v[c_k(n)]``                                  (prod_layer.py:398-466)
* Categorical ``log theta[x]``, Bernoulli ``log(theta or 1 - theta)`` (categorical.py:83-91,
  bernoulli.py:65-70)
* flows: ``f_c += sum_n f_n W[n, c] exp(v_c - v_n)``; parameter flows
  ``pf[n, c] = W[n, c] sum_b f_n exp(v_c - v_n)``                    (sum_layer.py:1569-1730, 2049-2169)
* EM: ``theta' = (1 - s) theta + s (pf + pc / nch) / (sum pf + pc)`` (par_update.py:214-256,
  categorical.py:164-201, bernoulli.py:91-101); with ``keep_zero_params`` zero parameters stay
  zero and the normaliser runs over the others (par_update.py:196-205, 236-253)
* other propagation algorithms (``alg`` of forward / backward; lt = log W[n, c] + v_c):
  MPE ``v[n] = max_c lt``, flows and parameter flows (not scaled by W) where |lt - v[n]| < 1e-6;
  GeneralLL ``v[n] = log sum_c exp(alpha lt) / alpha``, flows ``f_c += f_n exp(alpha (lt - v_n))``,
  parameter flows in the LL form ``W sum_b f_n exp(v_c - v_n)``
  (sum_layer.py:454-800, 1352-1392, 1569-1730, 2049-2169).  Pinned by the tests: GeneralLL with
  alpha = 1 equals LL; GeneralLL tends to MPE as alpha grows; the log-terms along an MPE
  path add up to the MPE value; flows into each variable's inputs sum to one per sample.
  GeneralLL parameter flows with alpha != 1 are pinned only through alpha = 1 (partial).
"""
from __future__ import annotations

from typing import Dict, Optional

import numpy as np

from pyjuice_amd.nodes.nodes import node_iterator

__all__ = ["sum_dense_params", "forward", "backward", "em_update", "sample", "input_params", "philox_uniforms"]


def _np(t):
    try:
        return t.detach().cpu().numpy()
    except AttributeError:
        return np.asarray(t)


def _src(ns):
    return ns.get_source_ns() if ns.is_tied() else ns


def sum_dense_params(ns, dtype=np.float64) -> np.ndarray:
    """Dense ``[num_nodes, num_ch_nodes]`` weight matrix of a sum vector (zeros off the edges)."""
    blocks = _np(_src(ns)._params).astype(dtype)  # [E, bs, cbs]
    E = _np(ns.edge_ids)
    bs, cbs = ns.block_size, ns.ch_block_size
    W = np.zeros([ns.num_node_blocks, bs, ns.num_ch_node_blocks, cbs], dtype=dtype)
    W[E[0], :, E[1], :] = blocks
    return W.reshape(ns.num_nodes, ns.num_ch_nodes)


def input_params(ns, dtype=np.float64) -> np.ndarray:
    p = _np(_src(ns)._params).astype(dtype)
    k = ns.dist.num_parameters()
    return p.reshape(ns.num_nodes, k)


def _prod_child_index(ns, k: int) -> np.ndarray:
    """Row of child ``k`` feeding each node of product vector ``ns``."""
    E = _np(ns.edge_ids)
    if E.shape[0] == ns.num_node_blocks:
        bs = ns.block_size
        return (E[:, k][:, None] * bs + np.arange(bs)[None, :]).reshape(-1)
    return E[:, k]


def _input_ll(ns, data: np.ndarray) -> np.ndarray:
    v = ns.scope.to_list()
    assert len(v) == 1
    x = data[:, v[0]].astype(np.int64)
    th = input_params(ns)
    sig = ns.dist.get_signature()
    with np.errstate(divide="ignore"):
        if sig == "Categorical":
            K = th.shape[1]
            ok = (x >= 0) & (x < K)
            out = np.full([ns.num_nodes, x.shape[0]], -np.inf)
            out[:, ok] = np.log(th[:, x[ok]])
            return out
        if sig == "Bernoulli":
            p = th[:, 0][:, None]
            return np.log(np.where(x[None, :] > 0, p, 1.0 - p))
    raise NotImplementedError(sig)


def _lse_matmul(W: np.ndarray, child: np.ndarray) -> np.ndarray:
    m = child.max(axis=0, keepdims=True)
    mm = np.where(np.isfinite(m), m, 0.0)
    with np.errstate(divide="ignore"):
        s = W @ np.exp(child - mm)
        out = np.log(s) + mm
    out[:, ~np.isfinite(m[0])] = -np.inf
    return out


def _alg_sum(W: np.ndarray, child: np.ndarray, alg: str, alpha: float) -> np.ndarray:
    """One sum vector's values under MPE / GeneralLL (sum_layer.py:508-563)."""
    with np.errstate(divide="ignore"):
        lW = np.log(W)
    lt = lW[:, :, None] + child[None, :, :]                       # [n, C, B]
    if alg == "MPE":
        return lt.max(axis=1)
    x = alpha * lt
    m = x.max(axis=1, keepdims=True)
    mm = np.where(np.isfinite(m), m, 0.0)
    with np.errstate(divide="ignore"):
        out = (np.log(np.exp(x - mm).sum(axis=1)) + mm[:, 0, :]) / alpha
    out[~np.isfinite(m[:, 0, :])] = -np.inf
    return out


def _alg_of(alg, ns) -> str:
    """The algorithm of sum vector ``ns``: ``alg`` itself, or its entry when ``alg`` maps vectors to
    algorithms (a per-layer choice; vectors not listed take "LL")."""
    return alg.get(ns, "LL") if isinstance(alg, dict) else alg


def forward(root, data, alg="LL", alpha: float = 1.0) -> Dict:
    """Log-values of every vector: ``{ns: [num_nodes, B]}``; ``data`` is ``[B, num_vars]``.
    ``alg``: "LL", "MPE" or "GeneralLL" (with ``alpha``) for the sum vectors, or a dict from sum
    vector to one of those."""
    data = _np(data)
    vals = dict()
    for ns in node_iterator(root):
        if ns.is_input():
            vals[ns] = _input_ll(ns, data)
        elif ns.is_prod():
            acc = np.zeros([ns.num_nodes, data.shape[0]])
            for k, cs in enumerate(ns.chs):
                acc = acc + vals[cs][_prod_child_index(ns, k)]
            vals[ns] = acc
        else:
            child = np.concatenate([vals[c] for c in ns.chs], axis=0)
            a = _alg_of(alg, ns)
            if a == "LL":
                vals[ns] = _lse_matmul(sum_dense_params(ns), child)
            else:
                vals[ns] = _alg_sum(sum_dense_params(ns), child, a, alpha)
    return vals


def backward(root, data, vals: Dict, root_flows: Optional[np.ndarray] = None, alg="LL",
             alpha: float = 1.0):
    """Flows of every vector and parameter flows.

    Returns ``(flows, pflows)``: ``flows[ns]`` is ``[num_nodes, B]``; ``pflows[ns]`` is the dense
    ``[num_nodes, num_ch_nodes]`` matrix for a sum vector and ``[num_nodes, K]`` (Categorical) or
    ``[num_nodes, 2]`` (Bernoulli; slot 0 for x = 1) for an input vector."""
    data = _np(data)
    B = data.shape[0]
    flows = {root: np.ones([root.num_nodes, B]) if root_flows is None else np.asarray(root_flows, np.float64)
             .reshape(root.num_nodes, B)}
    pflows = dict()
    for ns in node_iterator(root, reverse=True):
        f = flows.get(ns)
        if f is None:
            f = np.zeros([ns.num_nodes, B])
        if ns.is_sum() and _alg_of(alg, ns) != "LL":
            a = _alg_of(alg, ns)
            W = sum_dense_params(ns)
            child = np.concatenate([vals[c] for c in ns.chs], axis=0)
            v = vals[ns]
            with np.errstate(divide="ignore", invalid="ignore"):
                lt = np.log(W)[:, :, None] + child[None, :, :]       # [n, C, B]
                if a == "MPE":
                    share = (np.abs(lt - v[:, None, :]) < 1e-6).astype(np.float64)
                    pflows[ns] = (share * f[:, None, :]).sum(axis=2)
                else:
                    share = np.where(np.isfinite(v)[:, None, :] & np.isfinite(lt),
                                     np.exp(alpha * (lt - v[:, None, :])), 0.0)
                    ratio = np.where(np.isfinite(v)[:, None, :] & np.isfinite(child)[None, :, :],
                                     np.exp(child[None, :, :] - v[:, None, :]), 0.0)
                    pflows[ns] = W * (ratio * f[:, None, :]).sum(axis=2)
            cf = (share * f[:, None, :]).sum(axis=0)
            s = 0
            for c in ns.chs:
                flows[c] = flows.get(c, np.zeros([c.num_nodes, B])) + cf[s:s + c.num_nodes]
                s += c.num_nodes
        elif ns.is_sum():
            W = sum_dense_params(ns)
            child = np.concatenate([vals[c] for c in ns.chs], axis=0)
            v = vals[ns]
            with np.errstate(divide="ignore", invalid="ignore"):
                lr = np.where(np.isfinite(v) & (f > 0), np.log(np.maximum(f, 1e-300)) - v, -np.inf)
            shift = lr.max(axis=0, keepdims=True)
            sh = np.where(np.isfinite(shift), shift, 0.0)
            R = np.exp(lr - sh)                                  # [n, B]
            C = np.exp(np.clip(child + sh, -745.0, 709.0))       # [nch, B]
            C[:, ~np.isfinite(shift[0])] = 0.0
            C[~np.isfinite(child)] = 0.0
            cf = C * (W.T @ R)
            pflows[ns] = W * (R @ C.T)
            s = 0
            for c in ns.chs:
                flows[c] = flows.get(c, np.zeros([c.num_nodes, B])) + cf[s:s + c.num_nodes]
                s += c.num_nodes
        elif ns.is_prod():
            for k, cs in enumerate(ns.chs):
                if cs not in flows:
                    flows[cs] = np.zeros([cs.num_nodes, B])
                np.add.at(flows[cs], _prod_child_index(ns, k), f)
        else:
            v = ns.scope.to_list()[0]
            x = data[:, v].astype(np.int64)
            if ns.dist.get_signature() == "Categorical":
                K = ns.dist.num_cats
                pf = np.zeros([ns.num_nodes, K])
                ok = (x >= 0) & (x < K)
                for k in range(K):
                    pf[:, k] = f[:, ok & (x == k)].sum(axis=1)
            else:
                pf = np.stack([f[:, x > 0].sum(axis=1), f[:, x <= 0].sum(axis=1)], axis=1)
            pflows[ns] = pf
        flows[ns] = f
    return flows, pflows


def em_update(root, pflows: Dict, step_size: float, pseudocount: float = 0.0,
              keep_zero_params: bool = False) -> Dict:
    """New parameters of every source vector after one EM step: sum vectors as dense
    ``[num_nodes, num_ch_nodes]`` (zeros off the edges), inputs as ``[num_nodes, K]``.

    A node with n children (sum edges, or K categories) and flows pf_c is moved towards
    ``(pf_c + pc / n) / (sum_c pf_c + pc)`` (``par_update.py:214-256``, ``categorical.py:164-201``,
    ``bernoulli.py:91-101`` with ``pc / 2``).  With ``keep_zero_params`` (sum and Categorical
    vectors; ``par_update.py:196-205, 236-253``, ``categorical.py:171-201``) parameters below
    1e-12 stay 0 and the normaliser sums ``pf_c + pc / n`` over the other children only (n still
    counts every child)."""
    tot = dict()
    for ns, pf in pflows.items():
        s = _src(ns)
        tot[s] = tot.get(s, 0.0) + pf
    out = dict()

    def _step(th, pf, mask, n):
        num = pf + pseudocount / n
        if keep_zero_params:
            live = mask & (th >= 1e-12)
            den = np.where(live, num, 0.0).sum(axis=1, keepdims=True)
            new = num / np.where(den > 0, den, 1.0)
            upd = (1.0 - step_size) * th + step_size * new
            return np.where(live, upd, 0.0)
        den = np.where(mask, pf, 0.0).sum(axis=1, keepdims=True) + pseudocount
        return np.where(mask, (1.0 - step_size) * th + step_size * num / den, 0.0)

    for s, pf in tot.items():
        if s.is_sum():
            W = sum_dense_params(s)
            E = _np(s.edge_ids)
            mb = np.zeros([s.num_node_blocks, s.block_size, s.num_ch_node_blocks, s.ch_block_size], dtype=bool)
            mb[E[0], :, E[1], :] = True
            mask = mb.reshape(W.shape)
            nch = mask.sum(axis=1, keepdims=True).astype(np.float64)
            out[s] = _step(W, np.where(mask, pf, 0.0), mask, nch)
        elif s.dist.get_signature() == "Categorical":
            th = input_params(s)
            out[s] = _step(th, pf, np.ones_like(th, dtype=bool), float(th.shape[1]))
        else:
            th = input_params(s)[:, 0]
            new = (pf[:, 0] + 0.5 * pseudocount) / (pf[:, 0] + pf[:, 1] + pseudocount)
            out[s] = ((1.0 - step_size) * th + step_size * new)[:, None]
    return out


def philox4x32_10(ctr, key):
    """Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
    SC 2011) on arrays: ``ctr`` four uint32 words, ``key`` two; returns the four output words."""
    M0, M1, W0, W1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57), np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
    mask = np.uint64(0xFFFFFFFF)
    x0, x1, x2, x3 = (np.asarray(c, dtype=np.uint64) & mask for c in ctr)
    k0, k1 = (np.asarray(k, dtype=np.uint64) & mask for k in key)
    for _ in range(10):
        p0, p1 = M0 * x0, M1 * x2
        hi0, lo0 = p0 >> np.uint64(32), p0 & mask
        hi1, lo1 = p1 >> np.uint64(32), p1 & mask
        x0, x1, x2, x3 = hi1 ^ x1 ^ k0, lo1, hi0 ^ x3 ^ k1, lo0
        k0, k1 = (k0 + W0) & mask, (k1 + W1) & mask
    return x0, x1, x2, x3


def philox_uniforms(seed: int, rows, cols) -> np.ndarray:
    """Counter-based uniforms of (node row, sample column) pairs: word 0 of Philox4x32-10 with key
    (seed & 0xffffffff, seed >> 32) and counter (col, row, 0, 0), as (x >> 8) * 2^-24 in
    float32.  The device kernels draw the same numbers."""
    rows = np.asarray(rows, dtype=np.uint64)
    cols = np.asarray(cols, dtype=np.uint64)
    z = np.zeros_like(cols)
    x0 = philox4x32_10((cols, rows, z, z), (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF))[0]
    return (x0 >> np.uint64(8)).astype(np.float32) * np.float32(2.0 ** -24)


def _chunked_pick(w: np.ndarray, u: np.ndarray, chunk: int = 512) -> np.ndarray:
    """Child picks of one sum node for uniforms ``u``: children in chunks of ``chunk``; within a
    chunk the first child whose FP32 running sum (from 0, left to right) exceeds u; a u that
    reaches the chunk's total continues in the next chunk with the total subtracted (in FP32);
    past the last chunk the last child (``queries/sample.py:115-131``)."""
    n = w.shape[0]
    pick = np.full(u.shape, -1, dtype=np.int64)
    u = u.astype(np.float32).copy()
    for c0 in range(0, n, chunk):
        cum = np.cumsum(w[c0:c0 + chunk], dtype=np.float32)
        todo = pick < 0
        k = np.searchsorted(cum, u[todo], side="right")
        hit = k < cum.shape[0]
        idx = np.nonzero(todo)[0]
        pick[idx[hit]] = c0 + k[hit]
        u[idx[~hit]] = u[idx[~hit]] - cum[-1]
    pick[pick < 0] = n - 1
    return pick


def sample(root, uniforms, num_vars: int) -> np.ndarray:
    """Ancestral samples ``[S, num_vars]``.  ``uniforms`` is either ``[num_nodes, S]`` (the uniform
    of node ``id``, its compiled row ``ns._output_ind_range[0] + i``, for sample ``s`` at
    ``uniforms[id, s]``) or a tuple ``(fn, S)`` with ``fn(rows, cols)`` returning the uniforms of
    those (row, sample) pairs (e.g. ``philox_uniforms`` with a seed).

    A sum node's children are ordered as its edge list (flattened per child node) and picked by
    ``_chunked_pick`` (FP32 running sums in chunks of 512); a Categorical node picks the first
    category whose FP32 running sum reaches u (0 if none); Bernoulli gives ``u < theta``."""
    if isinstance(uniforms, tuple):
        ufn, S = uniforms
    else:
        U = np.asarray(uniforms, dtype=np.float32)
        S = U.shape[1]
        ufn = lambda r, c: U[r, c]
    act = {root: np.ones([root.num_nodes, S], dtype=bool)}
    out = np.zeros([S, num_vars], dtype=np.int64)
    for ns in node_iterator(root, reverse=True):
        a = act.get(ns)
        if a is None or not a.any():
            continue
        off = ns._output_ind_range[0]
        if ns.is_sum():
            blocks = _np(_src(ns)._params).astype(np.float32)   # [E, bs, cbs]
            E = _np(ns.edge_ids)
            bs, cbs = ns.block_size, ns.ch_block_size
            chosen = np.full([ns.num_nodes, S], -1, dtype=np.int64)  # concatenated child node id
            for blk in range(ns.num_node_blocks):
                es = np.nonzero(E[0] == blk)[0]
                ch_nodes = (E[1][es][:, None] * cbs + np.arange(cbs)[None, :]).reshape(-1)
                for i in range(bs):
                    n = blk * bs + i
                    cols = np.nonzero(a[n])[0]
                    if cols.size == 0:
                        continue
                    w = blocks[es][:, i, :].reshape(-1)
                    u = ufn(np.full(cols.shape, off + n), cols)
                    chosen[n, cols] = ch_nodes[_chunked_pick(w, u)]
            s0 = 0
            for c in ns.chs:
                sel = (chosen >= s0) & (chosen < s0 + c.num_nodes)
                if sel.any():
                    ca = act.setdefault(c, np.zeros([c.num_nodes, S], dtype=bool))
                    nn_, ss = np.nonzero(sel)
                    ca[chosen[nn_, ss] - s0, ss] = True
                s0 += c.num_nodes
        elif ns.is_prod():
            for k, cs in enumerate(ns.chs):
                ca = act.setdefault(cs, np.zeros([cs.num_nodes, S], dtype=bool))
                idx = _prod_child_index(ns, k)
                nn_, ss = np.nonzero(a)
                ca[idx[nn_], ss] = True
        else:
            v = ns.scope.to_list()[0]
            th = input_params(ns).astype(np.float32)
            nn_, ss = np.nonzero(a)
            u = ufn(off + nn_, ss)
            if ns.dist.get_signature() == "Categorical":
                cum = np.cumsum(th, axis=1, dtype=np.float32)
                hit = cum[nn_] >= u[:, None]
                pick = np.where(hit.any(axis=1), hit.argmax(axis=1), 0)
            else:
                pick = (u < th[nn_, 0]).astype(np.int64)
            out[ss, v] = pick
    return out
