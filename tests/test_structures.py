"""Structure builders (CPU only), pinned by what their definitions fix rather than by the
reference's outputs (which cannot be produced here):

* the Chow-Liu tree is a maximum spanning tree of the mutual-information matrix: its weight equals
  the best of all K^(K-2) labelled trees, enumerated from Pruefer sequences;
* on data drawn from a strong Markov chain, MI estimation plus the spanning tree recover the chain;
* every builder emits a smooth, decomposable circuit (sum children share the sum's scope, product
  children have disjoint scopes that cover the product's), over every variable exactly once;
* HCLT's sum-edge count is (I - 1) L^2 + L for I tree vertices with children (the root sum has one
  node over L children, every other inner variable an L x L sum);
* the homogeneous HMM ties every transition sum and every emission vector to one source.
"""
import itertools

import numpy as np
import pytest
import torch

import parity_utils as U
from pyjuice_amd.nodes.nodes import node_iterator
from pyjuice_amd.structures.hclt import chow_liu_tree, mutual_information_chunked


def _pruefer_to_edges(seq, K):
    degree = [1] * K
    for v in seq:
        degree[v] += 1
    edges = []
    for v in seq:
        leaf = min(u for u in range(K) if degree[u] == 1)
        edges.append((leaf, v))
        degree[leaf] -= 1
        degree[v] -= 1
    u, w = [x for x in range(K) if degree[x] == 1]
    edges.append((u, w))
    return edges


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_chow_liu_is_maximum_spanning_tree(seed):
    K = 6
    g = np.random.default_rng(seed)
    a = g.random([K, K])
    mi = a + a.T
    T = chow_liu_tree(mi)
    edges = list(T.edges())
    assert len(edges) == K - 1
    got = sum(mi[u, v] for u, v in edges)
    best = max(sum(mi[u, v] for u, v in _pruefer_to_edges(seq, K))
               for seq in itertools.product(range(K), repeat=K - 2))
    assert abs(got - best) <= 1e-12


def test_hclt_recovers_markov_chain():
    K, n = 7, 4000
    g = np.random.default_rng(3)
    perm = g.permutation(K)               # the chain runs over a shuffled variable order
    x = np.zeros([n, K], dtype=np.int64)
    x[:, perm[0]] = g.integers(0, 4, size=n)
    for t in range(1, K):
        keep = g.random(n) < 0.9
        x[:, perm[t]] = np.where(keep, x[:, perm[t - 1]], g.integers(0, 4, size=n))
    mi = mutual_information_chunked(torch.from_numpy(x).float(), torch.from_numpy(x).float(), 32, 0.5 / 32,
                                     chunk_size=4).numpy()
    np.testing.assert_allclose(mi, mi.T, atol=1e-5)
    T = chow_liu_tree(mi)
    want = {frozenset((int(perm[t - 1]), int(perm[t]))) for t in range(1, K)}
    assert {frozenset(e) for e in T.edges()} == want


def _scope_set(ns):
    return frozenset(ns.scope.to_list())


def _check_smooth_decomposable(root, num_vars):
    assert _scope_set(root) == frozenset(range(num_vars))
    for ns in node_iterator(root):
        if ns.is_sum():
            for c in ns.chs:
                assert _scope_set(c) == _scope_set(ns)
        elif ns.is_prod():
            seen = set()
            for c in ns.chs:
                sc = _scope_set(c)
                assert not (seen & sc)
                seen |= sc
            assert frozenset(seen) == _scope_set(ns)
        else:
            assert len(ns.scope.to_list()) == 1


@pytest.mark.parametrize("num_vars,L", [(8, 16), (12, 32)])
def test_hclt_structure(num_vars, L):
    ns, x, cards = U.hclt_categorical(num_vars=num_vars, K=5, L=L, n=512, seed=4)
    _check_smooth_decomposable(ns, num_vars)
    mi = mutual_information_chunked(torch.from_numpy(x).float(), torch.from_numpy(x).float(), 32, 0.5 / 32,
                                     chunk_size=32).numpy()
    T = chow_liu_tree(mi)
    import networkx as nx
    root = nx.center(T)[0]
    clt = nx.bfs_tree(T, root)
    inner = sum(1 for v in clt.nodes if clt.out_degree(v) > 0)
    e_sum = sum(n.edge_ids.size(1) * n.block_size * n.ch_block_size for n in node_iterator(ns) if n.is_sum())
    assert e_sum == (inner - 1) * L * L + L


def test_pd_structure():
    ns, cards = U.pd_small()
    _check_smooth_decomposable(ns, len(cards))
    ns, cards = U.pd_sparse()
    _check_smooth_decomposable(ns, len(cards))


@pytest.mark.parametrize("T,L", [(5, 16), (9, 32)])
def test_hmm_structure(T, L):
    ns, cards = U.hmm_small(T=T, L=L, V=7)
    _check_smooth_decomposable(ns, T)
    sums = [n for n in node_iterator(ns) if n.is_sum()]
    ins = [n for n in node_iterator(ns) if n.is_input()]
    assert len(ins) == T
    srcs = {id(n.get_source_ns()) for n in ins if n.is_tied()} | {id(n) for n in ins if not n.is_tied()}
    assert len(srcs) == 1
    trans = [n for n in sums if n.num_nodes == L]
    assert len(trans) == T - 1
    tsrc = {id(n.get_source_ns()) if n.is_tied() else id(n) for n in trans}
    assert len(tsrc) == 1
