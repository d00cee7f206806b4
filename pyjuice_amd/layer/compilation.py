"""Layer compiler: turns vectors of sum / product nodes into partitioned index tensors.

Outputs are bit-identical to the reference compiler (``pyjuice/layer/compilation.py``):
the per-edge index assignment of the CPU branches (``compilation.py:475-532``,
``:659-858``, ``:1166-1183``) restated as vectorised numpy (stable sorts give the same
"order of appearance" offsets the reference computes with sequential counters).

Index conventions (all int64):

* sum forward, per partition ``p``: ``nids[p][j]`` first node id of node block ``j``;
  ``cids[p][j, k]`` child element id; ``pids[p][j, k]`` / ``pfids[p][j, k]`` start of the
  ``block_size``-long parameter / parameter-flow run of child ``k`` (parent index fastest);
  padding is 0 (dummy element, dummy parameter).
* sum backward, per partition: ``chids[p][j]`` first element id of a child block;
  ``parids[p][j, k]`` first node id of a parent block; ``parpids[p][j, k]`` start of the
  ``[cbs][bs]`` parameter block of that edge.
* product forward: ``nids[p][j]`` element id, ``cids[p][j, k]`` child node id;
  product backward: ``u_cids[p][j]`` child node id, ``parids[p][j, k]`` parent element id.
"""
from __future__ import annotations

from collections import OrderedDict
from copy import deepcopy
from typing import Dict, List, Sequence

import numpy as np
import torch

__all__ = [
    "get_sum_layer_forward_stats", "sum_layer_forward_compilation", "get_sum_layer_backward_stats",
    "sum_layer_backward_compilation", "get_prod_layer_stats", "prod_layer_forward_compilation",
    "flatten_c_ids", "get_prod_layer_parstats", "prod_layer_backward_compilation", "assign_partitions",
]


def _np(x) -> np.ndarray:
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy().astype(np.int64)
    return np.asarray(x, dtype=np.int64)


def _rank_within_groups(keys: np.ndarray, num_groups: int):
    """For each position, its rank among positions with the same key (in order of appearance);
    also the per-key counts."""
    counts = np.bincount(keys, minlength=num_groups).astype(np.int64)
    order = np.argsort(keys, kind="stable")
    starts = np.zeros(num_groups, dtype=np.int64)
    if num_groups > 1:
        starts[1:] = np.cumsum(counts)[:-1]
    rank = np.empty(keys.shape[0], dtype=np.int64)
    rank[order] = np.arange(keys.shape[0], dtype=np.int64) - starts[keys[order]]
    return rank, counts


def _excl_cumsum(x: np.ndarray) -> np.ndarray:
    out = np.zeros_like(x)
    if x.shape[0] > 1:
        out[1:] = np.cumsum(x)[:-1]
    return out


def assign_partitions(counts: torch.Tensor, partition_max: torch.Tensor, start_min: int):
    """Partition membership by ranges of ``counts``: ``[start_min, max_0], [max_0+1, max_1], ...``
    (``sum_layer.py:85-94``, ``:159-168``; ``prod_layer.py:136-146``).  Elements outside every
    range keep partition 0 / local id 0 and are not counted."""
    c = _np(counts)
    maxes = _np(partition_max)
    part = np.zeros(c.shape[0], dtype=np.int64)
    local = np.zeros(c.shape[0], dtype=np.int64)
    num = np.zeros(maxes.shape[0], dtype=np.int64)
    lo = start_min
    for p, hi in enumerate(maxes.tolist()):
        crit = (c >= lo) & (c <= hi)
        k = int(crit.sum())
        part[crit] = p
        local[crit] = np.arange(k, dtype=np.int64)
        num[p] = k
        lo = hi + 1
    return torch.from_numpy(part), torch.from_numpy(local), torch.from_numpy(num)


# --------------------------------------------------------------------------- sum layers


def get_sum_layer_forward_stats(nodes, global_nid_start: int):
    """Node ranges and per-block child-node counts (``compilation.py:107-126``)."""
    layer_num_nblocks = sum(ns.num_node_blocks for ns in nodes)
    layer_num_edges = 0
    n_chs = np.zeros(layer_num_nblocks, dtype=np.int64)
    s = 0
    for ns in nodes:
        e = s + ns.num_node_blocks
        n_chs[s:e] = np.bincount(_np(ns.edge_ids[0, :]), minlength=ns.num_node_blocks) * ns.ch_block_size
        ns._output_ind_range = (global_nid_start, global_nid_start + ns.num_nodes)
        global_nid_start += ns.num_nodes
        layer_num_edges += ns.num_edges
        s = e
    return layer_num_nblocks, layer_num_edges, torch.from_numpy(n_chs)


def _assign_sum_param_ranges(ns, state, node2tiednodes, node2tiedcounts, max_tied):
    """Parameter / flow range bookkeeping incl. tying (``compilation.py:319-391``)."""
    if not ns.is_tied():
        if not ns.provided("_param_range"):
            ns._param_range = (state["pid"], state["pid"] + ns.num_edges)
            state["pid"] += ns.num_edges
            ns._param_flow_range = (state["pfid"], state["pfid"] + ns.num_edges)
            state["pfid"] += ns.num_edges
            add_p = add_pf = True
        else:
            assert ns.provided("_param_flow_range")
            add_p = add_pf = False
        return ns._param_range[0], ns._param_flow_range[0], add_p, add_pf, True

    src = ns.get_source_ns()
    if not src.provided("_param_range"):
        pr = (state["pid"], state["pid"] + ns.num_edges)
        pfr = (state["pfid"], state["pfid"] + ns.num_edges)
        ns._param_range, ns._param_flow_range = pr, pfr
        src._param_range, src._param_flow_range = pr, pfr
        state["pid"] += ns.num_edges
        state["pfid"] += ns.num_edges
        add_p = add_pf = True
    else:
        ns._param_range = deepcopy(src._param_range)
        add_p = add_pf = False

    if src not in node2tiednodes:
        node2tiednodes[src] = [[src], [src._param_flow_range]]
        node2tiedcounts[src] = [1]
    elif src not in node2tiedcounts:
        node2tiedcounts[src] = [0 for _ in range(len(node2tiednodes[src][0]))]

    if not ns.provided("_param_flow_range"):
        counts = node2tiedcounts[src]
        if all(c >= max_tied for c in counts):
            ns._param_flow_range = (state["pfid"], state["pfid"] + ns.num_edges)
            state["pfid"] += ns.num_edges
            node2tiednodes[src][1].append(ns._param_flow_range)
            node2tiednodes[src][0].append(ns)
            counts.append(1)
            add_pf = True
        else:
            t = min(range(len(counts)), key=lambda i: counts[i])
            ns._param_flow_range = deepcopy(node2tiednodes[src][1][t])
            counts[t] += 1
            add_pf = False
    return src._param_range[0], ns._param_flow_range[0], add_p, add_pf, False


def _flat_sum_edges(ns):
    """Block edges expanded to one edge per child node, in edge order
    (``compilation.py:393-401``).  Returns parent block ids and local child node ids."""
    E = _np(ns.edge_ids)
    cbs = ns.ch_block_size
    par = np.repeat(E[0], cbs)
    ch = (E[1][:, None] * cbs + np.arange(cbs, dtype=np.int64)[None, :]).reshape(-1)
    return par, ch


def _global_child_ids(ns, local_ch: np.ndarray) -> np.ndarray:
    sizes = np.array([c.num_nodes for c in ns.chs], dtype=np.int64)
    ends = np.cumsum(sizes)
    g = np.searchsorted(ends, local_ch, side="right")
    base = ends - sizes
    starts = np.array([c._output_ind_range[0] for c in ns.chs], dtype=np.int64)
    return starts[g] + local_ch - base[g]


@torch.no_grad()
def sum_layer_forward_compilation(nodes, fw_partition_max_chs, n_partition_ids, n_id_in_partition,
                                  num_ngs_in_partition, n_chs, global_nid_start: int, global_pid_start: int,
                                  global_pfid_start: int, node2tiednodes: dict, max_tied_ns_per_parflow_block: int = 4,
                                  **unused):
    maxch = _np(fw_partition_max_chs)
    num = _np(num_ngs_in_partition)
    part_of = _np(n_partition_ids)
    loc_of = _np(n_id_in_partition)
    P = maxch.shape[0]
    nids = [np.zeros(num[p], dtype=np.int64) for p in range(P)]
    cids = [np.zeros((num[p], maxch[p]), dtype=np.int64) for p in range(P)]
    pids = [np.zeros((num[p], maxch[p]), dtype=np.int64) for p in range(P)]
    pfids = [np.zeros((num[p], maxch[p]), dtype=np.int64) for p in range(P)]

    all_param_ids, all_pflow_ids, original = dict(), dict(), []
    state = {"pid": global_pid_start, "pfid": global_pfid_start}
    node2tiedcounts = dict()
    nblock_start = 0
    for ns_idx, ns in enumerate(nodes):
        pid0, pfid0, add_p, add_pf, is_orig = _assign_sum_param_ranges(
            ns, state, node2tiednodes, node2tiedcounts, max_tied_ns_per_parflow_block)
        if is_orig:
            original.append(ns_idx)

        nb, bs = ns.num_node_blocks, ns.block_size
        par, ch = _flat_sum_edges(ns)
        offs, cnt = _rank_within_groups(par, nb)
        gcid = _global_child_ids(ns, ch)
        run = (_excl_cumsum(cnt)[par] + offs) * bs
        gpid = pid0 + run
        gpfid = pfid0 + run

        blk = nblock_start + np.arange(nb, dtype=np.int64)
        bp, bl = part_of[blk], loc_of[blk]
        node_first = ns._output_ind_range[0] + np.arange(nb, dtype=np.int64) * bs
        ep, el = part_of[nblock_start + par], loc_of[nblock_start + par]
        for p in range(P):
            m = bp == p
            if m.any():
                nids[p][bl[m]] = node_first[m]
            m = ep == p
            if m.any():
                cids[p][el[m], offs[m]] = gcid[m]
                pids[p][el[m], offs[m]] = gpid[m]
                pfids[p][el[m], offs[m]] = gpfid[m]
        if add_p:
            all_param_ids[ns_idx] = gpid
        if add_pf:
            all_pflow_ids[ns_idx] = gpfid
        nblock_start += nb

    for ns_idx, ids in all_param_ids.items():
        ns = nodes[ns_idx]
        if ns.is_tied():
            ns = ns.get_source_ns()
        ns._param_ids = torch.from_numpy(ids[0::ns.ch_block_size].copy())
    for ns_idx, ids in all_pflow_ids.items():
        ns = nodes[ns_idx]
        ns._param_flow_ids = torch.from_numpy(ids[0::ns.ch_block_size].copy())
    for ns_idx in original:
        ns = nodes[ns_idx]
        ns._inverse_param_ids = torch.from_numpy(np.argsort(ns._param_ids.numpy(), kind="stable"))

    t = lambda lst: [torch.from_numpy(a) for a in lst]
    return t(nids), t(cids), t(pids), t(pfids), state["pid"], state["pfid"]


class _OrderedSet:
    def __init__(self):
        self._set, self._list = set(), []

    def append(self, x):
        if x not in self._set:
            self._set.add(x)
            self._list.append(x)

    def __contains__(self, x):
        return x in self._set

    def __iter__(self):
        return iter(self._list)

    def __len__(self):
        return len(self._list)


def _edges_into(ns, cs):
    """Block edges of ``ns`` whose child block lies in ``cs`` (``compilation.py:148-154``):
    returns (mask over ns edges, parent block ids, local child block ids)."""
    cs_id = next(i for i, c in enumerate(ns.chs) if c is cs)
    sid = sum(c.num_node_blocks for c in ns.chs[:cs_id])
    eid = sid + cs.num_node_blocks
    E = _np(ns.edge_ids)
    crit = (E[1] >= sid) & (E[1] < eid)
    return crit, E[0][crit], E[1][crit] - sid


def get_sum_layer_backward_stats(nodes):
    """Children grouped by block size and their parent-edge counts (``compilation.py:129-174``)."""
    ch_gsize2cs = OrderedDict()
    ch_gsize2num_nblocks = OrderedDict()
    cs2parns = dict()
    for ns in nodes:
        for cs in ns.chs:
            g = cs.block_size
            if g not in ch_gsize2cs:
                ch_gsize2cs[g] = _OrderedSet()
                ch_gsize2num_nblocks[g] = 0
            if cs not in ch_gsize2cs[g]:
                ch_gsize2cs[g].append(cs)
                ch_gsize2num_nblocks[g] += cs.num_node_blocks
            if cs not in cs2parns:
                cs2parns[cs] = _OrderedSet()
            cs2parns[cs].append(ns)

    ch_gsize2n_pargs = OrderedDict()
    for g, chs in ch_gsize2cs.items():
        n_pargs = np.zeros(ch_gsize2num_nblocks[g], dtype=np.int64)
        s = 0
        for cs in chs:
            e = s + cs.num_node_blocks
            acc = np.zeros(cs.num_node_blocks, dtype=np.int64)
            for ns in cs2parns[cs]:
                _, _, lc = _edges_into(ns, cs)
                acc += np.bincount(lc, minlength=cs.num_node_blocks)
            n_pargs[s:e] = acc
            s = e
        ch_gsize2n_pargs[g] = torch.from_numpy(n_pargs)
    return ch_gsize2cs, ch_gsize2num_nblocks, ch_gsize2n_pargs, cs2parns


@torch.no_grad()
def sum_layer_backward_compilation(nodes, cs2parns, n_partition_ids, n_id_in_partition, num_ngs_in_partition,
                                   partition_max_pars, **unused):
    maxp = _np(partition_max_pars)
    num = _np(num_ngs_in_partition)
    part_of = _np(n_partition_ids)
    loc_of = _np(n_id_in_partition)
    P = maxp.shape[0]
    chids = [np.zeros(num[p], dtype=np.int64) for p in range(P)]
    parids = [np.zeros((num[p], maxp[p]), dtype=np.int64) for p in range(P)]
    parpids = [np.zeros((num[p], maxp[p]), dtype=np.int64) for p in range(P)]

    ns_cache = dict()
    cs_nblock_start = 0
    for cs in nodes:
        seg_nid, seg_cid, seg_parid, seg_parpid = [], [], [], []
        for ns in cs2parns[cs]:
            crit, par_blk, ch_blk = _edges_into(ns, cs)
            if ns not in ns_cache:
                E = _np(ns.edge_ids)
                off, cnt = _rank_within_groups(E[0], ns.num_node_blocks)
                ns_cache[ns] = (off, _excl_cumsum(cnt))
            off, cum = ns_cache[ns]
            pid0 = ns.get_source_ns()._param_range[0] if ns.is_tied() else ns._param_range[0]
            seg_cid.append(ch_blk)
            seg_parid.append(ns._output_ind_range[0] + par_blk * ns.block_size)
            seg_parpid.append(pid0 + (cum[par_blk] + off[crit]) * ns.block_size * cs.block_size)
        cid = np.concatenate(seg_cid)
        parid = np.concatenate(seg_parid)
        parpid = np.concatenate(seg_parpid)
        pars_off, _ = _rank_within_groups(cid, cs.num_node_blocks)
        gblk = cs_nblock_start + cid
        ep, el = part_of[gblk], loc_of[gblk]
        chid = cs._output_ind_range[0] + cid * cs.block_size
        for p in range(P):
            m = ep == p
            if m.any():
                chids[p][el[m]] = chid[m]
                parids[p][el[m], pars_off[m]] = parid[m]
                parpids[p][el[m], pars_off[m]] = parpid[m]
        cs_nblock_start += cs.num_node_blocks

    t = lambda lst: [torch.from_numpy(a) for a in lst]
    return t(chids), t(parids), t(parpids)


# --------------------------------------------------------------------------- product layers


def get_prod_layer_stats(nodes, block_size: int, global_nid_start: int, use_block_sparse_edges: bool):
    """Element ranges and fan-ins (``compilation.py:863-900``)."""
    counts = [ns.num_node_blocks if use_block_sparse_edges else ns.num_nodes for ns in nodes]
    n_chgs = np.zeros(sum(counts), dtype=np.int64)
    layer_num_edges = 0
    s = 0
    for ns, c in zip(nodes, counts):
        n_chgs[s:s + c] = ns.num_chs
        layer_num_edges += ns.num_nodes * ns.num_chs
        ns._output_ind_range = (global_nid_start, global_nid_start + ns.num_nodes)
        global_nid_start += ns.num_nodes
        s += c
    return sum(counts), layer_num_edges, torch.from_numpy(n_chgs)


@torch.no_grad()
def prod_layer_forward_compilation(nodes, fw_partition_max_chs, n_partition_ids, n_id_in_partition,
                                   num_ngs_in_partition, block_size: int, use_block_sparse_edges: bool, **unused):
    """``compilation.py:903-956``."""
    if not use_block_sparse_edges:
        assert block_size == 1
    maxch = _np(fw_partition_max_chs)
    num = _np(num_ngs_in_partition)
    nids = [np.zeros(num[p], dtype=np.int64) for p in range(maxch.shape[0])]
    cids = [np.zeros((num[p], maxch[p]), dtype=np.int64) for p in range(maxch.shape[0])]
    for i, ns in enumerate(nodes):
        p = int(n_partition_ids[i])
        ls = int(n_id_in_partition[i])
        E = _np(ns.edge_ids)
        n0 = ns._output_ind_range[0]
        if use_block_sparse_edges:
            le = ls + ns.num_node_blocks
            nids[p][ls:le] = np.arange(0, ns.num_nodes, block_size, dtype=np.int64) + n0
            for k, cs in enumerate(ns.chs):
                cids[p][ls:le, k] = E[:, k] * block_size + cs._output_ind_range[0]
        else:
            le = ls + ns.num_nodes
            nids[p][ls:le] = np.arange(ns.num_nodes, dtype=np.int64) + n0
            if ns.is_sparse():
                for k, cs in enumerate(ns.chs):
                    cids[p][ls:le, k] = E[:, k] + cs._output_ind_range[0]
            else:
                bs = ns.block_size
                ex = (E[:, None, :] * bs + np.arange(bs, dtype=np.int64)[None, :, None]).reshape(-1, E.shape[1])
                for k, cs in enumerate(ns.chs):
                    cids[p][ls:le, k] = ex[:, k] + cs._output_ind_range[0]
    t = lambda lst: [torch.from_numpy(a) for a in lst]
    return t(nids), t(cids)


def flatten_c_ids(nids: Sequence[torch.Tensor], cids: Sequence[torch.Tensor]):
    """``compilation.py:959-978``."""
    flat_c = [c.reshape(-1) for c in cids]
    flat_n = [n.unsqueeze(1).repeat(1, c.size(1)).reshape(-1) for n, c in zip(nids, cids)]
    if len(flat_c) == 0:
        return torch.zeros([0], dtype=torch.long), torch.zeros([0], dtype=torch.long)
    return torch.cat(flat_c), torch.cat(flat_n)


def get_prod_layer_parstats(flat_cids: torch.Tensor, global_nid_start: int):
    """Unique children with parent counts; ids below ``global_nid_start`` dropped
    (``compilation.py:981-993``)."""
    u, counts = torch.unique(flat_cids, sorted=True, return_counts=True)
    keep = u >= global_nid_start
    # The reference drops the prefix up to the last id < global_nid_start; ids are sorted,
    # so this equals dropping every id below it.
    return u[keep], counts[keep]


@torch.no_grad()
def prod_layer_backward_compilation(flat_u_cids, flat_cids, flat_cid2nid, bk_partition_max_pars, n_partition_ids,
                                    n_id_in_partition, num_ns_in_partition, **unused):
    """``compilation.py:1079-1185`` (parents listed in order of appearance in ``flat_cids``)."""
    maxp = _np(bk_partition_max_pars)
    num = _np(num_ns_in_partition)
    P = maxp.shape[0]
    u_cids = [np.zeros(num[p], dtype=np.int64) for p in range(P)]
    parids = [np.zeros((num[p], maxp[p]), dtype=np.int64) for p in range(P)]
    u = _np(flat_u_cids)
    fc = _np(flat_cids)
    fn = _np(flat_cid2nid)
    part_of = _np(n_partition_ids)
    loc_of = _np(n_id_in_partition)
    if u.shape[0] > 0:
        pos = np.searchsorted(u, fc)
        pos_c = np.minimum(pos, u.shape[0] - 1)
        valid = u[pos_c] == fc
        idx = pos_c[valid]
        par = fn[valid]
        rank, _ = _rank_within_groups(idx, u.shape[0])
        for p in range(P):
            m = part_of == p
            u_cids[p][loc_of[m]] = u[m]
            m = part_of[idx] == p
            parids[p][loc_of[idx][m], rank[m]] = par[m]
    t = lambda lst: [torch.from_numpy(a) for a in lst]
    return t(u_cids), t(parids)
