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
