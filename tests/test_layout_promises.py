"""The host-side layout promises passed to the sum kernels hold on the compiled index tables
(checked entry by entry with plain loops, CPU only)."""
import numpy as np
import pytest

import parity_utils as U
from pyjuice_amd.model import TensorCircuit


def _check_run(layer, p, r):
    c = layer.partitioned_cids[p].numpy()
    q = layer.partitioned_pids[p].numpy()
    n = layer._fw_n_chs[p].numpy()
    for j in range(c.shape[0]):
        for k in range(int(n[j])):
            k0 = (k // r) * r
            assert c[j, k] == c[j, k0] + (k - k0)
            assert q[j, k] == q[j, k0] + (k - k0) * layer.block_size


@pytest.mark.parametrize("maker,min_run", [(lambda: U.hclt_categorical(L=64, block_size=32)[0], 32),
                                           (lambda: U.hclt_categorical(num_vars=6, L=128, K=10)[0], 128),
                                           (lambda: U.pd_small()[0], 0),
                                           (lambda: U.hmm_small()[0], 32)])
def test_child_run_promise(maker, min_run):
    pc = TensorCircuit(maker())
    runs = []
    for g in pc.inner_layer_groups:
        if g.is_prod():
            continue
        for layer in g:
            for p in range(layer.num_fw_partitions):
                if layer.partitioned_nids[p].numel() == 0:
                    continue
                r = layer._child_run(p)
                if r > 0:
                    _check_run(layer, p, r)
                if layer.block_size >= 32:
                    runs.append(r)
    if min_run:
        assert runs and min(runs) >= min_run


def test_unreached_products_are_masked():
    """PD's sparse block connections leave product blocks without any parent edge; the product
    backward tables must not read their (unwritten) element-flow rows."""
    import torch
    from pyjuice_amd.model import TensorCircuit
    pc = TensorCircuit(U.pd_sparse()[0])
    G = pc.inner_layer_groups
    for gid, g in enumerate(G):
        if g.is_prod():
            for layer in g:
                layer._dev = {"bk": [(u, p.to(torch.int32), 0) for u, p in
                                     zip(layer.partitioned_u_cids, layer.partitioned_parids)]}
    before = sum(int((p != 0).sum()) for g in G if g.is_prod() for l in g for _, p, _ in l._dev["bk"])
    pc._mask_unreached_products()
    after = sum(int((p != 0).sum()) for g in G if g.is_prod() for l in g for _, p, _ in l._dev["bk"])
    assert after < before        # this circuit does have unreached product blocks
    for gid in range(0, len(G) - 1, 2):
        covered = set()
        for sl in G[gid + 1]:
            for c in sl.partitioned_chids:
                covered.update(c.tolist())
        for layer in G[gid]:
            for _, p, _ in layer._dev["bk"]:
                vals = set(p.flatten().tolist()) - {0}
                assert vals <= covered


@pytest.mark.parametrize("num_vars,L", [(10, 128), (12, 64)])
def test_merged_partition_tables(num_vars, L):
    """One launch table per layer (and child block size): the rows are exactly the partitions'
    rows with zero padding on the right, merged tables ordered by decreasing child / parent
    count, and the merged child-run promise holds row by row."""
    import torch
    from pyjuice_amd.layer.sum_layer import SumLayer
    assert SumLayer.merge_partitions
    pc = TensorCircuit(U.hclt_pruned(num_vars=num_vars, L=L)[0])
    pc._build_launch_plans(torch.device("cpu"))
    merged = 0
    for g in pc.inner_layer_groups:
        if g.is_prod():
            continue
        for layer in g:
            fw = layer._dev["fw"]
            parts = [p for p in range(layer.num_fw_partitions) if layer.partitioned_nids[p].numel() > 0]
            assert len(fw) == 1 and len(layer._fw_child_run) == 1 and len(layer._fw_aligned) == 1
            nids, cids, pids, pfids, nch, tot = (t.numpy() if hasattr(t, "numpy") else t for t in fw[0])
            merged += len(parts) > 1
            want = dict()
            for p in parts:
                n = layer._fw_n_chs[p].numpy()
                for j, nid in enumerate(layer.partitioned_nids[p].numpy()):
                    k = int(n[j])
                    want[int(nid)] = (k, layer.partitioned_cids[p].numpy()[j, :k].tolist(),
                                      layer.partitioned_pids[p].numpy()[j, :k].tolist(),
                                      layer.partitioned_pfids[p].numpy()[j, :k].tolist())
            assert sorted(nids.tolist()) == sorted(want)
            assert tot == int(nch.sum()) and (len(parts) == 1 or np.all(np.diff(nch) <= 0))
            for j, nid in enumerate(nids.tolist()):
                k = int(nch[j])
                assert (k, cids[j, :k].tolist(), pids[j, :k].tolist(), pfids[j, :k].tolist()) == want[nid]
                assert not cids[j, k:].any() and not pids[j, k:].any() and not pfids[j, k:].any()
            r = layer._fw_child_run[0]
            for j in range(cids.shape[0]):
                for k in range(int(nch[j])):
                    k0 = (k // r) * r if r else k
                    assert cids[j, k] == cids[j, k0] + (k - k0)
                    assert pids[j, k] == pids[j, k0] + (k - k0) * layer.block_size
            by_cbs = dict()
            for q in range(layer.num_bk_partitions):
                if layer.partitioned_chids[q].numel() == 0:
                    continue
                npar = layer._bk_n_pars[q].numpy()
                d = by_cbs.setdefault(int(layer.cs_block_sizes[q]), dict())
                for i, c in enumerate(layer.partitioned_chids[q].numpy()):
                    k = int(npar[i])
                    d[int(c)] = (k, layer.partitioned_parids[q].numpy()[i, :k].tolist(),
                                 layer.partitioned_parpids[q].numpy()[i, :k].tolist())
            assert sorted(e[4] for e in layer._dev["bk"]) == sorted(by_cbs)
            for chids, parids, parpids, npar, cbs, totp in layer._dev["bk"]:
                chids, parids, parpids, npar = chids.numpy(), parids.numpy(), parpids.numpy(), npar.numpy()
                n_parts = sum(1 for q in range(layer.num_bk_partitions) if layer.partitioned_chids[q].numel() > 0
                              and int(layer.cs_block_sizes[q]) == cbs)
                assert totp == int(npar.sum()) and (n_parts == 1 or np.all(np.diff(npar) <= 0))
                assert sorted(chids.tolist()) == sorted(by_cbs[cbs])
                for i, c in enumerate(chids.tolist()):
                    k = int(npar[i])
                    assert (k, parids[i, :k].tolist(), parpids[i, :k].tolist()) == by_cbs[cbs][c]
                    assert not parids[i, k:].any() and not parpids[i, k:].any()
    assert merged > 0


@pytest.mark.parametrize("num_vars,L,K", [(6, 128, 10), (8, 256, 6)])
def test_product_fusion_tables(num_vars, L, K):
    """The fused forward's ``prod_bases`` rebuild the product layer: for every stage of every node
    block, the sum over the listed node rows equals the element rows the stage reads (FP64 oracle
    buffers), padding entries point into the zero dummy node rows, and the fused product groups'
    element rows are all covered."""
    import torch
    from oracle.pc_cpu import CompiledCPU
    ns, x, cards = U.hclt_categorical(num_vars=num_vars, L=L, K=K)
    pc = TensorCircuit(ns)
    pc._build_launch_plans(torch.device("cpu"))
    fz = pc._prod_fusion
    assert fz["groups"], "no product group fused"
    cpu = CompiledCPU(pc, np.float64)
    cpu.forward(U.random_data(cards, 8, seed=5))
    nm = cpu.node_mars
    G = pc.inner_layer_groups
    assert not nm[:pc.num_dummy_nodes].any()
    for gid in fz["groups"]:
        assert G[gid].is_prod() and not G[gid + 1].is_prod()
        em = cpu.element_mars[pc._elem_offsets[gid // 2]:]
        seen = set()
        for layer in G[gid + 1]:
            for (nids, cids, pids, pfids, nch, tot), (tab, kp, reads, writes) in zip(layer._dev["fw"],
                                                                                      fz["tables"][id(layer)]):
                c, n, t = cids.numpy(), nch.numpy(), tab.numpy().astype(np.int64)
                assert t.shape == (c.shape[0], -(-c.shape[1] // 128), kp) and kp in (2, 4, 8)
                for j in range(c.shape[0]):
                    for s, k0 in enumerate(range(0, int(n[j]), 128)):
                        rows = min(128, int(n[j]) - k0)
                        e0 = int(c[j, k0])
                        got = sum(nm[t[j, s, u]:t[j, s, u] + rows] for u in range(kp))
                        np.testing.assert_allclose(got, em[e0:e0 + rows], rtol=1e-12, atol=1e-12)
                        assert (t[j, s] + rows <= pc.num_nodes).all()
                        seen.update(range(e0, e0 + rows))
        for layer in G[gid]:
            for nids, cids, _ in layer._dev["fw"]:
                for e in nids.numpy().tolist():
                    assert set(range(e, e + layer.block_size)) <= seen


@pytest.mark.parametrize("T,L,V,tie", [(12, 16, 6, 8), (5, 32, 10, 2), (9, 16, 4, 3)])
def test_tied_input_groups(T, L, V, tie):
    """The tied-range input backward's groups: every node in exactly one group, a group is
    exactly the nodes with its flow range, one category count per group, and full coverage of the
    accumulator reported only when the ranges tile it."""
    import torch
    ns, cards = U.hmm_small(T=T, L=L, V=V)
    pc = TensorCircuit(ns, max_tied_ns_per_parflow_block=tie)
    layer = pc.input_layer_group[0]
    layer.device = torch.device("cpu")
    offsets, nodes, n_groups, covers = layer._tied_groups()
    off, nd = offsets.numpy(), nodes.numpy()
    s = layer.s_pfids.numpy()
    k = layer._ncats.numpy()
    assert off[0] == 0 and off[-1] == layer.num_nodes and len(off) == n_groups + 1
    assert sorted(nd.tolist()) == list(range(layer.num_nodes))
    slots = np.zeros(layer.num_param_flows, dtype=np.int64)
    for g in range(n_groups):
        members = nd[off[g]:off[g + 1]]
        assert len(members) >= 1
        assert len(set(s[members].tolist())) == 1 and len(set(k[members].tolist())) == 1
        assert set(np.flatnonzero(s == s[members[0]]).tolist()) == set(members.tolist())
        slots[s[members[0]]:s[members[0]] + k[members[0]]] += 1
    assert slots.max() == 1
    assert covers == bool(np.all(slots == 1))
    assert max(off[1:] - off[:-1]) <= tie
