"""Runs of narrow inner layers as one launch per pass (``pjamd_chain_pack`` /
``pjamd_chain_forward`` / ``pjamd_chain_backward``).

Tree-shaped circuits (HCLT, and any circuit built from a hidden tree) end in long stretches of
layers with a few node blocks each.  Launched one kernel per layer, such a layer costs a launch
plus the dependent memory round trips of its kernel (about 20 us forward and backward for a few
megabytes of work on HCLT-784), so the narrow tail of the circuit takes a third of a training
step.  A run executes a stretch of such levels -- a level being a product layer group and the
sum layer group above it -- in one launch:

* **Groups.**  A sum layer's node blocks and the product blocks they read form a bipartite
  graph; each connected component whose edges are all present is a *group*: R parent rows
  (its node blocks) densely connected to K child element rows (its product blocks).  In an
  HCLT a group is one tree node: L latent sums over the L products of that tree node.
* **Components.**  A group depends on the groups (lower in the run) that produce the node rows
  its products read; connected groups form a *component*, a subtree of the circuit.  Components
  never exchange data inside the run.
* **Workgroups.**  Workgroup w owns component ``w % C`` and batch columns
  ``[16 (w // C), 16 (w // C) + 16)`` and walks its component's groups level by level, handing
  results between consecutive operations through LDS: no workgroup ever waits for another.
* **Operands.**  Every workgroup of a component streams the component's parameters, so they are
  read as BF16 / FP16 in the matrix cores' own lane order: ``pjamd_chain_pack`` writes, once
  per forward, a packed copy of each group's parameters for the forward (A = theta, rows =
  parents, reduction over children) and for the backward (A = theta^T, rows = children,
  reduction over parents), 16 contiguous bytes per lane per 16x16x32 instruction.

A level may join a run when its sum layer group is a single layer, every one of its groups is
dense with R and K at most 256 (K a multiple of 32, block sizes powers of two >= 8 / >= 32),
every product block of the level feeds one of its groups with at most 4 children, the product
group is fused into the sum layer's element-flow pass (every child has a single parent) and the
layer's parameter flows come from the batched launch after the layer loop.  Among the eligible
stretches, runs are chosen by a dynamic programme over a cost model (launch cost of a run,
per-operation latency and per-workgroup parameter streaming against a fixed cost per level of
the layer-wise path).  The reference runs one launch per layer and partition
(``tensorcircuit.py:269-283``, ``:436-470``).
"""
from __future__ import annotations

import bisect
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

OPW = 12                 # PJAMD_CHAIN_OP_INTS: int32 fields per operation
(F_R, F_K, F_LBS, F_LCBS, F_NPB, F_NCB, F_KC, F_NID, F_CB, F_A, F_FLAGS) = range(11)
COMPW = 4                # PJAMD_CHAIN_COMP_INTS: (first op, ops, first table entry, table entries)
DESCW = 10               # PJAMD_CHAIN_DESC_INTS: int64 fields of a pack descriptor
MAX_OPS = 64             # PJAMD_CHAIN_MAX_OPS: operations of one component staged in LDS
TAB_CAP = 2048           # PJAMD_CHAIN_TAB_INTS: table entries of one component staged in LDS
MAX_R = 256              # PJAMD_CHAIN_MAX_ROWS: parent rows of a group
MAX_K = 256              # PJAMD_CHAIN_MAX_ROWS: child rows of a group
MAX_KC = 4               # PJAMD_CHAIN_MAX_KC: children of a product block
COLS = 16                # PJAMD_CHAIN_COLS: batch columns per workgroup
PACK_WG = 128            # units (16 bytes each) packed per workgroup of pjamd_chain_pack

# flags of a backward operation
PAR_LFM = 1      # the parents' node flows already hold log(flow) - mar
PAR_WRITE = 2    # the parents' node flows are raw: store log(flow) - mar back (modify in place)
CH_MOD = 4       # children marked in their code (bit 0) receive log(flow) - mar
SYNC = 8         # planner only: the operation would read rows of its component from global memory

# cost model (microseconds per training step, forward + backward) used to choose the runs,
# calibrated on HCLT-784 (256 latents, batch 512) on one MI355X
COST_LAUNCH = 8.0         # the two launches of a run
COST_OP = 13.0            # one group operation of up to 256 x 256, both passes (latency-bound)
COST_LEVEL = 24.0         # one narrow level launched layer by layer (sum forward, product, fused backward)
NUM_CUS = 256


@dataclass
class Group:
    """Dense block of a sum layer: parent blocks (first node rows ``nid``, block size ``bs``)
    fully connected to child blocks (first element rows ``erow`` of the whole ``element_mars``
    buffer, block size ``cbs``); ``ch[b]`` the node-row starts of child block b's product
    children; ``pbase[a, b]`` the first parameter of edge block (a, b) (parameter of parent row
    i and child row r at ``pbase + r * bs + i``)."""
    level: int
    bs: int
    cbs: int
    nid: List[int]
    erow: List[int]
    ch: List[List[int]]
    pbase: np.ndarray
    index: int = -1      # position among the packed groups
    a_fwd: int = 0       # packed operand offsets (16-byte units)
    a_bwd: int = 0

    @property
    def R(self) -> int:
        return len(self.nid) * self.bs

    @property
    def K(self) -> int:
        return len(self.erow) * self.cbs

    @property
    def KC(self) -> int:
        return max([len(c) for c in self.ch] + [1])


@dataclass
class Level:
    """An eligible (product group, sum group) pair: ``pgid`` / ``sgid`` their group ids."""
    pgid: int
    sgid: int
    layer: object
    groups: List[Group]
    premod: bool


@dataclass
class Run:
    """One launch per pass over the levels ``levels`` (inner layer groups ``[lo, hi)``).
    ``comp_*`` int32 [C, COMPW], ``ops_*`` int32 [n, OPW], ``tab_*`` int32 on the device; the
    backward tables per value of ``modify``."""
    lo: int
    hi: int
    C: int
    comps: List[List[Group]]
    fwd: Tuple[torch.Tensor, torch.Tensor, torch.Tensor]
    bwd: Dict[bool, Tuple[torch.Tensor, torch.Tensor, torch.Tensor]]
    bytes_per_col: Tuple[int, int]
    flops_per_col: int
    param_bytes: int

    def algo(self, B: int, direction: str):
        """Algorithmic (bytes, flops) of one launch: node / element rows once per column and the
        packed parameters once."""
        per_col = self.bytes_per_col[0 if direction == "fwd" else 1]
        return per_col * B + self.param_bytes, self.flops_per_col * B


@dataclass
class Plan:
    runs: List[Run] = field(default_factory=list)
    groups: List[Group] = field(default_factory=list)
    pack_desc: Optional[torch.Tensor] = None      # int64 [n_groups, DESCW]
    pack_pbase: Optional[torch.Tensor] = None     # int64
    pack_wg: Optional[torch.Tensor] = None        # int32 [n_wg, 3]: group, direction, first unit
    pack_units: int = 0                           # 16-byte units of the packed buffer

    def forward_runs(self) -> Dict[int, Run]:
        return {r.lo: r for r in self.runs}

    def backward_runs(self) -> Dict[int, Run]:
        return {r.hi - 1: r for r in self.runs}


def _np(t) -> np.ndarray:
    return t.detach().cpu().numpy().astype(np.int64)


def _pow2(v: int) -> bool:
    return v > 0 and (v & (v - 1)) == 0


def _premod_ranges(pc) -> List[Tuple[int, int]]:
    out = []
    for g in pc.inner_layer_groups:
        if g.is_prod():
            continue
        for layer in g:
            if id(layer) in pc._fused["premod"]:
                out.append(tuple(layer.nid_range))
    return sorted(out)


def _level(pc, p: int) -> Optional[Level]:
    """The level (product group p, sum group p + 1) as dense groups, or None if it cannot join
    a run."""
    groups = pc.inner_layer_groups
    if p + 1 >= len(groups) or not groups[p].is_prod() or groups[p + 1].is_prod():
        return None
    layers = list(groups[p + 1])
    if len(layers) != 1:
        return None
    layer = layers[0]
    bs = layer.block_size
    if not _pow2(bs) or bs < 8 or p not in pc._fused["prod_groups"] or id(layer) not in pc._par_batches["layers"]:
        return None
    eoff = pc._elem_offsets[p // 2]
    # product blocks of the level: local first element row -> (children node rows, block size)
    pblk: Dict[int, Tuple[List[int], int]] = dict()
    for pl in groups[p]:
        pbs = pl.block_size
        for n, c in zip(pl.partitioned_nids, pl.partitioned_cids):
            for e, ch in zip(_np(n).tolist(), _np(c).tolist()):
                pblk[int(e)] = ([int(x) for x in ch if x != 0], pbs)
    if not pblk:
        return None
    # node blocks: first node row -> [(child block start, parameter base)]
    blocks: List[Tuple[int, List[Tuple[int, int]]]] = []
    for n, c, q, k in zip(layer.partitioned_nids, layer.partitioned_cids, layer.partitioned_pids, layer._fw_n_chs):
        n, c, q, k = _np(n), _np(c), _np(q), _np(k)
        for j in range(n.size):
            cj, qj = c[j, :k[j]], q[j, :k[j]]
            kids, i = [], 0
            while i < cj.size:
                e0 = int(cj[i])
                if e0 not in pblk:
                    return None
                pbs = pblk[e0][1]
                seg = np.arange(pbs)
                if i + pbs > cj.size or not np.array_equal(cj[i:i + pbs], e0 + seg) or \
                        not np.array_equal(qj[i:i + pbs], qj[i] + seg * bs):
                    return None
                kids.append((e0, int(qj[i])))
                i += pbs
            blocks.append((int(n[j]), kids))
    # connected components of the bipartite graph (node blocks, product blocks)
    parent = dict()

    def find(x):
        while parent.setdefault(x, x) != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for nid, kids in blocks:
        for e0, _ in kids:
            a, b = find(("n", nid)), find(("e", e0))
            if a != b:
                parent[a] = b
    members: Dict[object, Tuple[List, set]] = dict()
    for nid, kids in blocks:
        r = find(("n", nid))
        members.setdefault(r, ([], set()))[0].append((nid, kids))
        members[r][1].update(e0 for e0, _ in kids)
    covered = set()
    out = []
    level_idx = p // 2
    for r, (nbl, eset) in members.items():
        eb = sorted(eset)
        pbs = {pblk[e][1] for e in eb}
        if len(pbs) != 1:
            return None
        cbs = pbs.pop()
        if not _pow2(cbs) or cbs < 32:
            return None
        pos = {e: i for i, e in enumerate(eb)}
        nbl = sorted(nbl)
        pbase = np.zeros([len(nbl), len(eb)], dtype=np.int64)
        for a, (nid, kids) in enumerate(nbl):
            if len(kids) != len(eb) or {e for e, _ in kids} != eset:
                return None          # not dense
            for e0, q0 in kids:
                pbase[a, pos[e0]] = q0
        g = Group(level=level_idx, bs=bs, cbs=cbs, nid=[nid for nid, _ in nbl], erow=[eoff + e for e in eb],
                  ch=[pblk[e][0] for e in eb], pbase=pbase)
        if g.R > MAX_R or g.K > MAX_K or g.R % 32 != 0 or g.K % 32 != 0 or g.KC > MAX_KC:
            return None
        covered.update(eb)
        out.append(g)
    if covered != set(pblk.keys()):
        return None
    return Level(pgid=p, sgid=p + 1, layer=layer, groups=out, premod=id(layer) in pc._fused["premod"])


def _components(levels: List[Level], block_of) -> List[List[Group]]:
    """Components of the groups of consecutive ``levels`` (lowest first): groups joined when one
    reads node rows another produces.  ``block_of(row)`` names the group producing a node row
    (or None)."""
    inside = {id(g) for lv in levels for g in lv.groups}
    parent = {id(g): id(g) for lv in levels for g in lv.groups}

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for lv in levels:
        for g in lv.groups:
            for ch in g.ch:
                for c in ch:
                    src = block_of(c)
                    if src is not None and id(src) in inside:
                        a, b = find(id(g)), find(id(src))
                        if a != b:
                            parent[a] = b
    comps: Dict[int, List[Group]] = dict()
    for lv in levels:
        for g in lv.groups:
            comps.setdefault(find(id(g)), []).append(g)
    return list(comps.values())


def _comp_cost(comp: List[Group]) -> float:
    # latency-bound: smaller groups save little
    return sum(COST_OP * max(0.6, g.R * g.K / (MAX_R * MAX_K)) for g in comp)


def _run_cost(comps: List[List[Group]], slices: int) -> float:
    """Launch, then the slowest component, stretched when the run's workgroups (components x
    batch slices, one per compute unit) outnumber the compute units."""
    return COST_LAUNCH + max(_comp_cost(c) for c in comps) * max(1.0, len(comps) * slices / NUM_CUS)


def _fits(comps: List[List[Group]]) -> bool:
    for c in comps:
        if len(c) > MAX_OPS:
            return False
        if sum(2 * len(g.nid) + len(g.erow) * (1 + g.KC) for g in c) > TAB_CAP:
            return False
    return True


def plan(pc, device, B: int, min_levels: int = 2) -> Plan:
    """Runs of ``pc`` for batches of ``B`` columns (both passes share them) and the packing
    tables of their groups."""
    slices = -(-B // COLS)
    groups = pc.inner_layer_groups
    levels: List[Optional[Level]] = [None] * ((len(groups) + 1) // 2)
    for p in range(0, len(groups) - 1, 2):
        levels[p // 2] = _level(pc, p)
    # node row -> producing group (sum node blocks of eligible levels)
    starts, owners = [], []
    for lv in levels:
        if lv is None:
            continue
        for g in lv.groups:
            for nid in g.nid:
                starts.append(nid)
                owners.append((nid, g))
    order = np.argsort(np.asarray(starts, dtype=np.int64), kind="stable")
    starts = [starts[i] for i in order]
    owners = [owners[i] for i in order]

    def block_of(row: int):
        i = bisect.bisect_right(starts, row) - 1
        if i < 0:
            return None
        nid, g = owners[i]
        return g if row < nid + g.bs else None

    # dynamic programme over each maximal stretch of eligible levels
    chosen: List[Tuple[int, int, List[List[Group]]]] = []
    i = 0
    while i < len(levels):
        if levels[i] is None:
            i += 1
            continue
        j = i
        while j < len(levels) and levels[j] is not None:
            j += 1
        n = j - i
        best = [0.0] + [float("inf")] * n
        back: List[Optional[Tuple[int, Optional[List[List[Group]]]]]] = [None] * (n + 1)
        for b in range(1, n + 1):
            cand = best[b - 1] + COST_LEVEL
            if cand < best[b]:
                best[b], back[b] = cand, (b - 1, None)
            for a in range(b - min_levels, -1, -1):
                comps = _components(levels[i + a:i + b], block_of)
                if not _fits(comps):
                    break
                cand = best[a] + _run_cost(comps, slices)
                if cand < best[b]:
                    best[b], back[b] = cand, (a, comps)
        b = n
        picked = []
        while b > 0:
            a, comps = back[b]
            if comps is not None:
                picked.append((i + a, i + b, comps))
            b = a
        chosen.extend(reversed(picked))
        i = j

    # runs whose components would read their own rows back from global memory stay layer-wise
    chosen = [(a, b, comps) for a, b, comps in chosen if not _needs_sync(comps)]
    out = Plan()
    unit = 0
    for a, b, comps in chosen:
        for c in comps:
            for g in c:
                g.index = len(out.groups)
                g.a_fwd = unit
                unit += g.R * g.K // 8
                g.a_bwd = unit
                unit += g.R * g.K // 8
                out.groups.append(g)
    out.pack_units = unit
    for a, b, comps in chosen:
        lo, hi = levels[a].pgid, levels[b - 1].sgid + 1
        out.runs.append(_build_run(pc, levels[a:b], lo, hi, comps, device))
    if out.groups:
        out.pack_desc, out.pack_pbase, out.pack_wg = _pack_tables(out.groups, device)
    return out


def _needs_sync(comps: List[List[Group]]) -> bool:
    for direction in ("fwd", "bwd"):
        _c, ops, _t = _tables(comps, direction, lambda g: 0, lambda x: 4 * x, torch.device("cpu"))
        if bool(((ops[:, F_FLAGS] & SYNC) != 0).any()):
            return True
    return False


def _i32(rows, width: int, device) -> torch.Tensor:
    a = np.asarray(rows, dtype=np.int64).reshape(-1, width) if rows else np.zeros([0, width], dtype=np.int64)
    assert a.size == 0 or (a.min() >= -2 ** 31 and a.max() < 2 ** 31)
    return torch.from_numpy(a.astype(np.int32)).to(device).contiguous()


def _tables(comps: List[List[Group]], direction: str, flags_of, code_of, device):
    """(comp, ops, tab) int32 device tables of one pass: per component its groups in pass order,
    every table offset relative to the component's first table entry.

    Hand-over through LDS: the kernels keep the outputs of a component's last two operations in
    a ring (slot = operation index & 1).  Forward, a product child whose rows an operation
    o - 1 or o - 2 produced, as consecutive output positions, is named -(1 + 256 slot + first
    position) instead of its node row.  Backward, a parent block (bs >= 32) whose flows are the
    element flows of a child block of operation o - 1 or o - 2 gets the ring row ``256 slot +
    first row`` in its source entry (after the parent blocks' node rows; -1 = node flows in
    global memory); a child code is ``node row << 2 | marked | 2 if every row of the child is
    read through the ring`` (its producer then stores the flows raw and the consumer stores them
    back as log(flow) - mar).  An operation that would have to read in-run rows from global
    memory gets the flag SYNC, which the kernels do not implement: ``plan`` drops such runs."""
    comp_rows, op_rows, tab = [], [], []
    for c in comps:
        seq = sorted(c, key=lambda g: g.level, reverse=(direction == "bwd"))
        op0, tab0 = len(op_rows) // OPW, len(tab)
        # produced node-row ranges of this component: [first row, rows, op index, first
        # position, table index of the code (backward), rows read through the ring]
        made: List[list] = []
        for o, g in enumerate(seq):
            KC = g.KC
            flags = flags_of(g)
            made.sort(key=lambda m: m[0])
            starts = [m[0] for m in made]

            def source(row: int, n: int):
                """(made entry, position of ``row``) if rows [row, row + n) are consecutive
                positions of one in-run output, (entry, None) if produced in-run otherwise,
                None if produced outside the run."""
                i = bisect.bisect_right(starts, row) - 1
                if i < 0 or row >= made[i][0] + made[i][1]:
                    return None
                m = made[i]
                if row + n <= m[0] + m[1]:
                    return m, m[3] + row - m[0]
                return m, None

            nid_off = len(tab) - tab0
            tab.extend(g.nid)
            if direction == "bwd":
                for nid in g.nid:
                    src = source(nid, g.bs)
                    if src is not None and src[1] is not None and o - src[0][2] in (1, 2) and g.bs >= 32:
                        tab.append(((src[0][2] & 1) << 8) + src[1])
                        src[0][5] += g.bs
                    else:
                        if src is not None:
                            flags |= SYNC
                        tab.append(-1)
            cb_off = len(tab) - tab0
            for b, (e, ch) in enumerate(zip(g.erow, g.ch)):
                tab.append(e)
                if direction == "fwd":
                    # global-memory children first (the kernel prefetches the first two)
                    refs = []
                    for x in ch:
                        src = source(x, g.cbs)
                        if src is not None and src[1] is not None and o - src[0][2] in (1, 2):
                            refs.append(-(1 + ((src[0][2] & 1) << 8) + src[1]))
                            continue
                        if src is not None:
                            flags |= SYNC
                        refs.append(code_of(x))
                    tab.extend(sorted(refs, key=lambda r: r < 0))
                else:
                    for x in ch:
                        made.append([x, g.cbs, o, b * g.cbs, len(tab), 0])
                        tab.append(code_of(x))
                tab.extend([0] * (KC - len(ch)))
            if direction == "fwd":
                for pb, nid in enumerate(g.nid):
                    made.append([nid, g.bs, o, pb * g.bs, -1, 0])
            op = [0] * OPW
            op[F_R], op[F_K] = g.R, g.K
            op[F_LBS], op[F_LCBS] = g.bs.bit_length() - 1, g.cbs.bit_length() - 1
            op[F_NPB], op[F_NCB], op[F_KC] = len(g.nid), len(g.erow), KC
            op[F_NID], op[F_CB] = nid_off, cb_off
            op[F_A] = g.a_fwd if direction == "fwd" else g.a_bwd
            op[F_FLAGS] = flags
            op_rows.extend(op)
        if direction == "bwd":
            for m in made:
                if m[5] == m[1]:
                    tab[m[4]] |= 2
        comp_rows.extend([op0, len(op_rows) // OPW - op0, tab0, len(tab) - tab0])
    return _i32(comp_rows, COMPW, device), _i32(op_rows, OPW, device), _i32(tab, 1, device).reshape(-1)


def _build_run(pc, levels: List[Level], lo: int, hi: int, comps: List[List[Group]], device) -> Run:
    premod_lv = {lv.pgid // 2: lv.premod for lv in levels}
    premod = _premod_ranges(pc)
    pstarts = [a for a, _ in premod]

    def is_premod(row: int) -> bool:
        i = bisect.bisect_right(pstarts, row) - 1
        return i >= 0 and row < premod[i][1]

    fwd = _tables(comps, "fwd", lambda g: 0, lambda x: x, device)
    bwd = dict()
    for modify in (False, True):
        def flags_of(g, modify=modify):
            if not modify:
                return 0
            return CH_MOD | (PAR_LFM if premod_lv[g.level] else PAR_WRITE)

        def code_of(x, modify=modify):
            return 4 * x + (1 if is_premod(x) else 0)

        bwd[modify] = _tables(comps, "bwd", flags_of, code_of, device)
    groups = [g for c in comps for g in c]
    nkids = sum(len(ch) * g.cbs for g in groups for ch in g.ch)
    # forward: product children read, element rows written, node rows written; backward: parent
    # flows (and marginals) read, element rows read, children's flows written
    fwd_col = 4 * sum(g.K + g.R for g in groups) + 4 * nkids
    bwd_col = 4 * sum(2 * g.R + g.K for g in groups) + 4 * nkids
    return Run(lo=lo, hi=hi, C=len(comps), comps=comps, fwd=fwd, bwd=bwd, bytes_per_col=(fwd_col, bwd_col),
               flops_per_col=sum(2 * g.R * g.K for g in groups), param_bytes=sum(2 * g.R * g.K for g in groups))


def _pack_tables(groups: List[Group], device):
    desc, pbase, wg = [], [], []
    for g in groups:
        assert g.R * g.K // 8 % PACK_WG == 0
        desc.append([g.R, g.K, g.bs.bit_length() - 1, g.cbs.bit_length() - 1, len(g.nid), len(g.erow),
                     g.a_fwd, g.a_bwd, len(pbase), 0])
        pbase.extend(g.pbase.reshape(-1).tolist())
        for d in (0, 1):
            for u in range(0, g.R * g.K // 8, PACK_WG):
                wg.append([g.index, d, u])
    d = torch.tensor(desc, dtype=torch.int64).to(device).contiguous()
    pb = torch.tensor(pbase, dtype=torch.int64).to(device).contiguous()
    return d, pb, _i32(wg, 3, device)
