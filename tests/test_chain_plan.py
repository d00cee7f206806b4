"""Chain runs (``pyjuice_amd.model.chain``), CPU only.

A float64 numpy restatement of what ``pjamd_chain_pack``, ``pjamd_chain_forward`` and
``pjamd_chain_backward`` document is driven by the planner's device tables (the same int32 /
int64 fields the kernels read: component rows, operation headers, child-block entries, pack
descriptors and workgroup table) and applied to the FP64 oracle's own buffers with the rows a
run writes erased first; the run must rebuild exactly the oracle's rows.  A wrong packed
operand layout, table offset, element row, parameter base, child code or flag changes those
rows.  Also checked: every index a table reaches stays inside its buffer, the limits the kernels
assume hold, and the components of a run never share node rows."""
import numpy as np
import pytest
import torch

import parity_utils as U
from oracle.pc_cpu import CompiledCPU
from pyjuice_amd.model import TensorCircuit
from pyjuice_amd.model import chain as C


def _circuit(name):
    if name == "hclt_bs32":
        ns, x, cards = U.hclt_categorical(num_vars=14, K=6, L=64, block_size=32)
    elif name == "hclt_bs64":
        ns, x, cards = U.hclt_categorical(num_vars=10, K=5, L=64)
    elif name == "hclt_L128_bs32":
        ns, x, cards = U.hclt_categorical(num_vars=9, K=4, L=128, block_size=32)
    elif name == "hmm":
        ns, cards = U.hmm_small(T=8, L=32, V=7)
    else:
        raise KeyError(name)
    pc = TensorCircuit(ns)
    pc._build_launch_plans(torch.device("cpu"))
    return pc, cards


NAMES = ["hclt_bs32", "hclt_bs64", "hclt_L128_bs32", "hmm"]


def _np(t):
    return t.detach().cpu().numpy().astype(np.int64)


def _acc_row(r, lane):
    return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)


def pack_emulate(plan, params):
    """The packed buffer [units, 8] (float64) from the pack workgroup table, unit by unit."""
    desc, pbase, wg = _np(plan.pack_desc), _np(plan.pack_pbase), _np(plan.pack_wg)
    out = np.full([plan.pack_units, 8], np.nan)
    for gi, d, u0 in wg:
        R, K, lbs, lcbs, npb, ncb, a_f, a_b, poff, _ = desc[gi]
        bs, cbs = 1 << lbs, 1 << lcbs
        u = u0 + np.arange(C.PACK_WG)
        lane, m = u & 63, u >> 6
        if d == 0:
            t, s = m // (K >> 5), m % (K >> 5)
            pr, c0 = 16 * t + (lane & 15), 32 * s + 8 * (lane >> 4)
            src = pbase[poff + (pr >> lbs) * ncb + (c0 >> lcbs)] + (c0 & (cbs - 1)) * bs + (pr & (bs - 1))
            out[a_f + u] = params[src[:, None] + np.arange(8)[None, :] * bs]
        else:
            t, s = m // (R >> 5), m % (R >> 5)
            ch, p0 = 16 * t + (lane & 15), 32 * s + 8 * (lane >> 4)
            src = pbase[poff + (p0 >> lbs) * ncb + (ch >> lcbs)] + (ch & (cbs - 1)) * bs + (p0 & (bs - 1))
            out[a_b + u] = params[src[:, None] + np.arange(8)[None, :]]
    return out


def _operand(packed, a, M, Kr):
    """The [M, Kr] matrix a wavefront's 16x16x32 A operands describe (tile t, step s, lane l)."""
    A = np.full([M, Kr], np.nan)
    lane = np.arange(64)
    for t in range(M // 16):
        for s in range(Kr // 32):
            u = a + (t * (Kr // 32) + s) * 64 + lane
            A[(16 * t + lane % 16)[:, None], (32 * s + 8 * (lane // 16))[:, None] + np.arange(8)[None, :]] = packed[u]
    return A


def _ops(tables):
    comp, ops, tab = (_np(t) for t in tables)
    for c in range(comp.shape[0]):
        op0, nops, tab0, ntab = comp[c]
        T = tab[tab0:tab0 + ntab]
        for o in range(op0, op0 + nops):
            yield c, ops[o], T


def emulate_forward(run, packed, nm, em):
    ring, last_c = dict(), None
    for c, op, T in _ops(run.fwd):
        if c != last_c:
            ring, last_c, o = dict(), c, 0
        R, K, lbs, lcbs, npb, ncb, KC, nid_off, cb_off, a, flags = op[:11]
        bs, cbs = 1 << lbs, 1 << lcbs
        e = np.zeros([K, nm.shape[1]])
        for b in range(ncb):
            ent = cb_off + b * (1 + KC)
            rows = np.zeros([cbs, nm.shape[1]])
            for u in range(KC):
                ch = T[ent + 1 + u]
                if ch < 0:                          # ring: -(1 + 256 slot + first position)
                    slot, p0 = (-ch - 1) >> 8, (-ch - 1) & 255
                    rows = rows + ring[slot][p0:p0 + cbs]
                else:
                    rows = rows + nm[ch:ch + cbs]   # child 0: the zero dummy rows
            e[b * cbs:(b + 1) * cbs] = rows
            em[T[ent]:T[ent] + cbs] = rows
        mx = e.max(axis=0)
        with np.errstate(invalid="ignore", over="ignore"):
            w = np.where(np.isfinite(mx)[None, :], np.exp(e - mx[None, :]), 0.0)
        acc = _operand(packed, a, R, K) @ w
        with np.errstate(divide="ignore"):
            out = np.where((acc > 0) & np.isfinite(mx)[None, :], np.log(np.maximum(acc, 1e-300)) + mx[None, :], -np.inf)
        for p in range(R):
            nm[T[nid_off + (p >> lbs)] + (p & (bs - 1))] = out[p]
        ring[o & 1] = out
        o += 1


def _lfdm(f, m):
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(m == -np.inf, -np.inf, np.log(f) - m)


def emulate_backward(run, modify, packed, nf, nm, em):
    ring, last_c = dict(), None
    for c, op, T in _ops(run.bwd[modify]):
        if c != last_c:
            ring, last_c, o = dict(), c, 0
        R, K, lbs, lcbs, npb, ncb, KC, nid_off, cb_off, a, flags = op[:11]
        bs, cbs = 1 << lbs, 1 << lcbs
        lfm = np.zeros([R, nf.shape[1]])
        for pb in range(npb):
            rows = T[nid_off + pb] + np.arange(bs)
            src = T[nid_off + npb + pb]
            if src >= 0:                            # ring row 256 slot + first row: raw flows
                f = ring[src >> 8][(src & 255):(src & 255) + bs]
                lfm[pb * bs:(pb + 1) * bs] = _lfdm(f, nm[rows])
                if flags & (C.PAR_WRITE | C.PAR_LFM):
                    nf[rows] = lfm[pb * bs:(pb + 1) * bs]
            elif flags & C.PAR_LFM:
                lfm[pb * bs:(pb + 1) * bs] = nf[rows]
            else:
                lfm[pb * bs:(pb + 1) * bs] = _lfdm(nf[rows], nm[rows])
                if flags & C.PAR_WRITE:
                    nf[rows] = lfm[pb * bs:(pb + 1) * bs]
        s = lfm.max(axis=0)
        with np.errstate(invalid="ignore", over="ignore"):
            w = np.where(np.isfinite(s)[None, :], np.exp(lfm - s[None, :]), 0.0)
        acc = _operand(packed, a, K, R) @ w
        vall = np.zeros([K, nf.shape[1]])
        for b in range(ncb):
            ent = cb_off + b * (1 + KC)
            e = em[T[ent]:T[ent] + cbs]
            with np.errstate(invalid="ignore", over="ignore"):
                v = np.where(np.isfinite(s)[None, :] & (e != -np.inf), acc[b * cbs:(b + 1) * cbs] * np.exp(e + s[None, :]),
                             0.0)
            vall[b * cbs:(b + 1) * cbs] = v
            for u in range(KC):
                code = T[ent + 1 + u]
                if code == 0:
                    continue
                row = code >> 2
                if (code & 1) and not (code & 2) and (flags & C.CH_MOD):
                    nf[row:row + cbs] = _lfdm(v, nm[row:row + cbs])
                else:
                    nf[row:row + cbs] = v
        ring[o & 1] = vall
        o += 1


def _close(a, b):
    fa, fb = np.isfinite(a), np.isfinite(b)
    assert np.array_equal(fa, fb) and np.array_equal(a[~fa], b[~fb])
    assert np.allclose(a[fa], b[fb], rtol=1e-9, atol=1e-12)


def _plan(pc, B=16):
    plan = C.plan(pc, torch.device("cpu"), B)
    assert plan.runs, "no run planned"
    return plan


def _check_tables(pc, plan):
    P = pc.params.numel()
    desc, pbase = _np(plan.pack_desc), _np(plan.pack_pbase)
    assert desc.shape[1] == C.DESCW and plan.pack_units == sum(2 * d[0] * d[1] // 8 for d in desc)
    for d in desc:
        R, K, lbs, lcbs, npb, ncb, a_f, a_b, poff, _ = d
        bs, cbs = 1 << lbs, 1 << lcbs
        assert R == npb * bs <= C.MAX_R and K == ncb * cbs <= C.MAX_K and R % 32 == 0 and K % 32 == 0
        assert bs >= 8 and cbs >= 32
        assert a_f + R * K // 8 <= plan.pack_units and a_b + R * K // 8 <= plan.pack_units
        pb = pbase[poff:poff + npb * ncb]
        assert pb.min() >= 0 and pb.max() + (cbs - 1) * bs + bs <= P
    for run in plan.runs:
        for tables in [run.fwd] + list(run.bwd.values()):
            comp = _np(tables[0])
            assert comp.shape == (run.C, C.COMPW)
            assert (comp[:, 1] <= C.MAX_OPS).all() and (comp[:, 3] <= C.TAB_CAP).all()
            for _c, op, T in _ops(tables):
                R, K, lbs, lcbs, npb, ncb, KC, nid_off, cb_off, a, flags = op[:11]
                bs, cbs = 1 << lbs, 1 << lcbs
                assert 1 <= KC <= C.MAX_KC and R <= C.MAX_R and K <= C.MAX_K
                assert nid_off + npb <= T.size and cb_off + ncb * (1 + KC) <= T.size
                assert T[nid_off:nid_off + npb].max() + bs <= pc.num_nodes
                if tables is not run.fwd:
                    src = T[nid_off + npb:nid_off + 2 * npb]
                    ok = src < 0
                    assert ((src[~ok] & 255) + bs <= C.MAX_K).all() and (src[~ok] >> 8 <= 1).all() and bs >= 32 or ok.all()
                ents = T[cb_off:cb_off + ncb * (1 + KC)].reshape(ncb, 1 + KC)
                assert ents[:, 0].min() >= 0 and ents[:, 0].max() + cbs <= pc.num_element_rows
                kids = ents[:, 1:]
                if tables is not run.fwd:
                    kids = kids >> 2
                else:
                    ring = kids < 0
                    pos = (-kids[ring] - 1) & 255
                    assert (pos + cbs <= C.MAX_R).all() and ((-kids[ring] - 1) >> 8 <= 1).all()
                    kids = np.where(ring, 0, kids)
                assert kids.max() + cbs <= pc.num_nodes


@pytest.mark.parametrize("name", NAMES)
def test_forward_runs_rebuild_oracle_rows(name):
    pc, cards = _circuit(name)
    plan = _plan(pc)
    _check_tables(pc, plan)
    x = U.random_data(cards, 16, seed=3)
    cpu = CompiledCPU(pc, np.float64)
    cpu.forward(x)
    packed = pack_emulate(plan, pc.params.detach().numpy().astype(np.float64))
    assert not np.isnan(packed).any()
    for run in plan.runs:
        G = pc.inner_layer_groups
        assert G[run.lo].is_prod() and not G[run.hi - 1].is_prod()
        nm, em = cpu.node_mars.copy(), cpu.element_mars.copy()
        for c in run.comps:
            for g in c:
                for n in g.nid:
                    nm[n:n + g.bs] = np.nan
                for e in g.erow:
                    em[e:e + g.cbs] = np.nan
        emulate_forward(run, packed, nm, em)
        _close(nm, cpu.node_mars)
        _close(em, cpu.element_mars)


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("modify", [True, False])
def test_backward_runs_rebuild_oracle_rows(name, modify):
    pc, cards = _circuit(name)
    plan = _plan(pc)
    x = U.random_data(cards, 16, seed=4)
    cpu = CompiledCPU(pc, np.float64)
    cpu.forward(x)
    cpu.backward()
    packed = pack_emulate(plan, pc.params.detach().numpy().astype(np.float64))
    nm, em, raw = cpu.node_mars, cpu.element_mars, cpu.node_flows
    fz = pc._fused
    # rows whose flows the layer-wise pipeline keeps as log(flow) - mar: every row of a
    # premodified sum layer
    mod_rows = np.zeros(pc.num_nodes, dtype=bool)
    if modify:
        for g in pc.inner_layer_groups:
            for layer in g:
                if not g.is_prod() and id(layer) in fz["premod"]:
                    mod_rows[layer.nid_range[0]:layer.nid_range[1]] = True
    for run in plan.runs:
        written = np.zeros(pc.num_nodes, dtype=bool)
        parents = np.zeros(pc.num_nodes, dtype=bool)
        for _c, op, T in _ops(run.bwd[modify]):
            R, K, lbs, lcbs, npb, ncb, KC, nid_off, cb_off, a, flags = op[:11]
            for n in T[nid_off:nid_off + npb]:
                parents[n:n + (1 << lbs)] = True
            ents = T[cb_off:cb_off + ncb * (1 + KC)].reshape(ncb, 1 + KC)
            for code in ents[:, 1:].reshape(-1):
                if code:
                    written[(code >> 2) + np.arange(1 << lcbs)] = True
        # parents of the run's operations end modified when the pipeline modifies
        final_mod = mod_rows | (parents if modify else np.zeros_like(parents))
        want = raw.copy()
        want[final_mod] = _lfdm(raw[final_mod], nm[final_mod])
        start = raw.copy()
        start[mod_rows] = want[mod_rows]
        start[written] = np.nan
        emulate_backward(run, modify, packed, start, nm, em)
        touched = written | parents
        assert written.any()
        _close(start[touched], want[touched])


@pytest.mark.parametrize("name", NAMES)
def test_components_are_independent(name):
    """Rows one component writes are never read or written by another component of the run,
    and each component's operations run bottom-up (forward) / top-down (backward)."""
    pc, _ = _circuit(name)
    plan = _plan(pc)
    for run in plan.runs:
        owner = dict()
        for ci, comp in enumerate(run.comps):
            for g in comp:
                for n in g.nid:
                    for r in range(n, n + g.bs):
                        assert owner.setdefault(("n", r), ci) == ci
        for ci, comp in enumerate(run.comps):
            for g in comp:
                for ch in g.ch:
                    for c in ch:
                        assert owner.get(("n", c), ci) == ci
        for tables, rev in [(run.fwd, False)] + [(t, True) for t in run.bwd.values()]:
            comp_rows = _np(tables[0])
            ops = _np(tables[1])
            for op0, nops, _t0, _nt in comp_rows:
                lv = ops[op0:op0 + nops]
                assert lv.shape[0] >= 1


def test_runs_skip_the_root_and_follow_costs():
    pc, _ = _circuit("hclt_bs32")
    plan = _plan(pc)
    G = pc.inner_layer_groups
    for r in plan.runs:
        for gid in range(r.lo + 1, r.hi, 2):
            assert len(list(G[gid])) == 1
            layer = list(G[gid])[0]
            assert layer.block_size >= 8 and id(layer) in pc._par_batches["layers"]
            assert gid - 1 in pc._fused["prod_groups"]
    # a huge launch cost makes every run lose to the layer-wise path
    old = C.COST_LAUNCH
    try:
        C.COST_LAUNCH = 1e9
        assert not C.plan(pc, torch.device("cpu"), 16).runs
    finally:
        C.COST_LAUNCH = old
