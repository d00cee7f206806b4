"""Chain runs (``pyjuice_amd.model.chain``), CPU only.

The planner's operation tables are applied, by a float64 numpy restatement of what
``pjamd_chain_forward`` / ``pjamd_chain_backward`` document for each operation kind, to the
FP64 oracle's own buffers with the rows a run writes erased first; the run must rebuild exactly
the oracle's rows.  A wrong table offset, element slab, block size, parameter base or flag
changes those rows.  Also checked: every index an operation reaches stays inside its buffer,
and the runs are where the planner's rules put them."""
import numpy as np
import pytest
import torch

import parity_utils as U
from oracle.pc_cpu import CompiledCPU
from pyjuice_amd.model import TensorCircuit
from pyjuice_amd.model import chain as C

F = C.OP_FIELDS


def _circuit(name):
    if name == "hclt_bs32":
        ns, x, cards = U.hclt_categorical(num_vars=14, K=6, L=64, block_size=32)
    elif name == "hclt_bs64":
        ns, x, cards = U.hclt_categorical(num_vars=10, K=5, L=64)
    elif name == "hmm":
        ns, cards = U.hmm_small(T=8, L=32, V=7)
    else:
        raise KeyError(name)
    pc = TensorCircuit(ns)
    pc._build_launch_plans(torch.device("cpu"))
    return pc, cards


def _ops(run):
    return run.ops.numpy().astype(np.int64), run.idx.numpy().astype(np.int64)


def _fields(op):
    kind, nb, K, bs, cbs, eoff, a, b, c, d, e, Kc, flags = (int(v) for v in op[:13])
    return kind, nb, K, bs, cbs, eoff, a, b, c, d, e, Kc, flags


def _lfdm(nf, nm):
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(nm == -np.inf, -np.inf, np.log(nf) - nm)


def _emulate_forward(run, nm, em, params):
    ops, idx = _ops(run)
    for op in ops:
        kind, nb, K, bs, cbs, eoff, a, b, c, d, e, Kc, flags = _fields(op)
        if kind == C.KIND_PROD:
            nids, cids = idx[a:a + nb], idx[b:b + nb * K].reshape(nb, K)
            for n in range(nb):
                acc = np.zeros_like(nm[:bs])
                for ch in cids[n]:
                    if ch != 0:
                        acc = acc + nm[ch:ch + bs]
                em[eoff + nids[n]:eoff + nids[n] + bs] = acc
        elif kind == C.KIND_SUM_FWD:
            nids, nch = idx[a:a + nb], idx[d:d + nb]
            cids, pids = idx[b:b + nb * K].reshape(nb, K), idx[c:c + nb * K].reshape(nb, K)
            for j in range(nb):
                n = nch[j]
                E = em[eoff + cids[j, :n]]
                mx = E.max(axis=0)
                ms = np.where(np.isfinite(mx), mx, 0.0)
                P = params[pids[j, :n][:, None] + np.arange(bs)[None, :]]
                S = P.T @ np.exp(E - ms)
                with np.errstate(divide="ignore"):
                    nm[nids[j]:nids[j] + bs] = np.where((S > 0) & np.isfinite(mx), np.log(S) + ms, -np.inf)
        else:
            raise AssertionError(f"kind {kind} in a forward run")


def _emulate_backward(run, nf, nm, em, params):
    ops, idx = _ops(run)
    for op in ops:
        kind, nb, K, bs, cbs, eoff, a, b, c, d, e, Kc, flags = _fields(op)
        if kind == C.KIND_MODIFY:
            for n in idx[a:a + nb]:
                nf[n:n + bs] = _lfdm(nf[n:n + bs], nm[n:n + bs])
        elif kind == C.KIND_SUM_ELE:
            chids, npar = idx[a:a + nb], idx[d:d + nb]
            parids, parpids = idx[b:b + nb * K].reshape(nb, K), idx[c:c + nb * K].reshape(nb, K)
            codes = idx[e:e + nb * Kc].reshape(nb, Kc)
            for q in range(nb):
                rows = np.concatenate([parids[q, p] + np.arange(bs) for p in range(npar[q])])
                lf = nf[rows] if flags & 1 else _lfdm(nf[rows], nm[rows])
                mx = lf.max(axis=0)
                ms = np.where(np.isfinite(mx), mx, 0.0)
                th = np.concatenate([params[parpids[q, p] + np.arange(cbs)[:, None] * bs + np.arange(bs)[None, :]]
                                     for p in range(npar[q])], axis=1)
                S = th @ np.exp(lf - ms)
                ev = em[eoff + chids[q] + np.arange(cbs)]
                with np.errstate(over="ignore", invalid="ignore"):
                    v = np.where(np.isfinite(mx) & np.isfinite(ev), S * np.exp(ev + ms), 0.0)
                for code in codes[q]:
                    if code == 0:
                        continue
                    dst = (code >> 1) + np.arange(cbs)
                    nf[dst] = _lfdm(v, nm[dst]) if (code & 1) and (flags & 2) else v
        else:
            raise AssertionError(f"kind {kind} in a backward run")


def _written_rows_fwd(run):
    ops, idx = _ops(run)
    node_rows, elem_rows = [], []
    for op in ops:
        kind, nb, K, bs, cbs, eoff, a, b, c, d, e, Kc, flags = _fields(op)
        for n in idx[a:a + nb]:
            (elem_rows if kind == C.KIND_PROD else node_rows).append(eoff * (kind == C.KIND_PROD) + n +
                                                                     np.arange(bs))
    return np.concatenate(node_rows), np.concatenate(elem_rows)


def _close(a, b):
    fa, fb = np.isfinite(a), np.isfinite(b)
    assert np.array_equal(fa, fb) and np.array_equal(a[~fa], b[~fb])
    assert np.allclose(a[fa], b[fb], rtol=1e-9, atol=1e-12)


def _check_bounds(pc, run):
    ops, idx = _ops(run)
    P = pc.params.numel()
    for op in ops:
        kind, nb, K, bs, cbs, eoff, a, b, c, d, e, Kc, flags = _fields(op)
        assert 0 <= a and a + nb <= idx.size
        if kind == C.KIND_PROD:
            assert nb * K <= C.MAX_K and bs & (bs - 1) == 0
            nids, cids = idx[a:a + nb], idx[b:b + nb * K]
            assert nids.min() >= 0 and eoff + nids.max() + bs <= pc.num_element_rows
            assert cids.min() >= 0 and cids.max() + bs <= pc.num_nodes
        elif kind == C.KIND_SUM_FWD:
            nch = idx[d:d + nb]
            assert bs >= 32 and bs & (bs - 1) == 0 and nb <= C.MAX_BLOCKS and nb * K <= C.MAX_K
            assert np.maximum((nch + C.CHUNK - 1) // C.CHUNK, 1).sum() * C.CHUNK <= C.MAX_K
            cids, pids = idx[b:b + nb * K].reshape(nb, K), idx[c:c + nb * K].reshape(nb, K)
            for j in range(nb):
                assert eoff + cids[j, :nch[j]].max() < pc.num_element_rows
                assert pids[j, :nch[j]].min() >= 0 and pids[j, :nch[j]].max() + bs <= P
            assert idx[a:a + nb].max() + bs <= pc.num_nodes
        elif kind == C.KIND_SUM_ELE:
            npar = idx[d:d + nb]
            assert nb <= C.MAX_BLOCKS and nb * K <= C.MAX_K and cbs % 32 == 0
            assert np.maximum((npar * bs + C.CHUNK - 1) // C.CHUNK, 1).sum() * C.CHUNK <= C.MAX_K
            parids, parpids = idx[b:b + nb * K].reshape(nb, K), idx[c:c + nb * K].reshape(nb, K)
            for q in range(nb):
                assert parids[q, :npar[q]].max() + bs <= pc.num_nodes
                assert parpids[q, :npar[q]].max() + cbs * bs <= P
            assert eoff + idx[a:a + nb].max() + cbs <= pc.num_element_rows
            codes = idx[e:e + nb * Kc]
            assert (codes[codes != 0] >> 1).max() + cbs <= pc.num_nodes
        elif kind == C.KIND_MODIFY:
            assert idx[a:a + nb].max() + bs <= pc.num_nodes


@pytest.mark.parametrize("name", ["hclt_bs32", "hclt_bs64", "hmm"])
def test_forward_runs_rebuild_oracle_rows(name):
    pc, cards = _circuit(name)
    runs = C.plan(pc, "fwd", "bf16", True, True, 1 << 20)
    assert runs, "no run planned"
    x = U.random_data(cards, 16, seed=3)
    cpu = CompiledCPU(pc, np.float64)
    cpu.forward(x)
    params = pc.params.detach().numpy().astype(np.float64)
    for run in runs.values():
        _check_bounds(pc, run)
        assert pc.inner_layer_groups[run.lo].is_prod() and not pc.inner_layer_groups[run.hi - 1].is_prod()
        nm, em = cpu.node_mars.copy(), cpu.element_mars.copy()
        rn, re = _written_rows_fwd(run)
        nm[rn] = np.nan
        em[re] = np.nan
        _emulate_forward(run, nm, em, params)
        _close(nm, cpu.node_mars)
        _close(em, cpu.element_mars)


@pytest.mark.parametrize("name", ["hclt_bs32", "hclt_bs64", "hmm"])
@pytest.mark.parametrize("modify", [True, False])
def test_backward_runs_rebuild_oracle_rows(name, modify):
    pc, cards = _circuit(name)
    runs = C.plan(pc, "bwd", "bf16", modify, True, 1 << 20)
    assert runs, "no run planned"
    x = U.random_data(cards, 16, seed=4)
    cpu = CompiledCPU(pc, np.float64)
    cpu.forward(x)
    cpu.backward()
    params = pc.params.detach().numpy().astype(np.float64)
    nm, em, raw = cpu.node_mars, cpu.element_mars, cpu.node_flows
    fz = pc._fused
    # rows whose flows the pipeline keeps as log(flow) - mar: every row of a premodified sum layer
    mod_rows = np.zeros(pc.num_nodes, dtype=bool)
    if modify:
        for g in pc.inner_layer_groups:
            for layer in g:
                if not g.is_prod() and id(layer) in fz["premod"]:
                    mod_rows[layer.nid_range[0]:layer.nid_range[1]] = True
    for key, run in runs.items():
        _check_bounds(pc, run)
        assert key == run.hi - 1 and all(g in fz["prod_groups"] for g in range(run.lo, run.hi, 2))
        ops, idx = _ops(run)
        want = raw.copy()
        start = raw.copy()
        written = np.zeros(pc.num_nodes, dtype=bool)
        modified_here = np.zeros(pc.num_nodes, dtype=bool)
        for op in ops:
            kind, nb, K, bs, cbs, eoff, a, b, c, d, e, Kc, flags = _fields(op)
            if kind == C.KIND_MODIFY:
                for n in idx[a:a + nb]:
                    modified_here[n:n + bs] = True
            else:
                for code in idx[e:e + nb * Kc]:
                    if code:
                        written[(code >> 1) + np.arange(cbs)] = True
        final_mod = mod_rows | modified_here
        want[final_mod] = _lfdm(raw[final_mod], nm[final_mod])
        start[mod_rows] = want[mod_rows]
        start[written] = np.nan
        _emulate_backward(run, start, nm, em, params)
        touched = written | modified_here
        assert touched.any()
        _close(start[touched], want[touched])


def test_run_boundaries_follow_the_rules():
    pc, _ = _circuit("hclt_bs32")
    G = pc.inner_layer_groups
    fwd = C.plan(pc, "fwd", "bf16", True, True, 1 << 20)
    # the root sum layer (block size 1) ends every run
    assert all(G[r.hi - 1].layers[0].block_size >= 32 for r in fwd.values())
    # a parameter budget below one pair's footprint plans nothing
    assert C.plan(pc, "fwd", "bf16", True, True, 1024) == {}
    # FP32 operands halve the staged reduction length
    assert C.max_k("fp32") == C.MAX_K // 2
    bwd = C.plan(pc, "bwd", "bf16", True, True, 1 << 20)
    for r in bwd.values():
        for gid in range(r.lo + 1, r.hi, 2):
            for layer in G[gid]:
                assert id(layer) in pc._par_batches["layers"]
