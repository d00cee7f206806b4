"""Pins the FP64 DAG oracle against what the mathematics fixes (CPU only).

* forward: probabilities of all assignments sum to one (exhaustive enumeration) and the
  closed-form per-vector values of ``tests/model/simple_model_test.py:240-277``;
* backward: parameter flows equal ``theta * d(sum_b LL_b) / d theta`` (central differences),
  flows are conserved (``sum_c pf[n, c] = sum_b flow[n, b]``), root-child flows sum to one;
* EM: closed-form update and monotone likelihood for a full step;
* sampling: empirical frequencies match the enumerated distribution.
"""
import itertools

import numpy as np
import pytest
import torch

import kat_circuits
import pyjuice_amd as juice
import pyjuice_amd.nodes.distributions as dists
from oracle import dag_eval
from pyjuice_amd.model import TensorCircuit


def _all_assignments(cards):
    return np.array(list(itertools.product(*[range(k) for k in cards])), dtype=np.int64)


def _hclt_bernoulli(num_vars=5, L=4, seed=3):
    torch.manual_seed(seed)
    g = np.random.default_rng(seed)
    x = torch.from_numpy((g.random([64, num_vars]) < 0.5).astype(np.float32))
    ns = juice.structures.HCLT(x, num_latents=L, input_dist=dists.Bernoulli(), chunk_size=8)
    return ns


def _compiled(ns, **kw):
    """Compile, copy the (FP32) parameters back to the DAG and renormalise them in FP64 so
    that sums of probabilities are exact up to FP64 rounding."""
    pc = TensorCircuit(ns, **kw)
    pc.update_parameters()
    for n in ns:
        if n.is_tied():
            continue
        if n.is_sum():
            P = n._params.double()
            E = n.edge_ids
            tot = torch.zeros([n.num_node_blocks, n.block_size], dtype=torch.float64)
            tot.index_add_(0, E[0], P.sum(dim=2))
            n._params = P / tot[E[0]].unsqueeze(2)
        elif n.is_input():
            P = n._params.double()
            if n.dist.get_signature() == "Categorical":
                P = (P.reshape(-1, n.dist.num_cats) / P.reshape(-1, n.dist.num_cats).sum(1, keepdim=True)).reshape(-1)
            n._params = P
    return pc


CASES = {
    "simple_model": (lambda: kat_circuits.simple_model_dag(), [4, 4, 6, 6]),
    "tying_block1": (lambda: kat_circuits.tying_block1_dag(), [5, 5, 5, 5]),
    "hclt_bern5": (lambda: _hclt_bernoulli(), [2] * 5),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_forward_normalises(case):
    build, cards = CASES[case]
    ns = build()
    _compiled(ns, layer_sparsity_tol=0.1, max_tied_ns_per_parflow_block=2)
    X = _all_assignments(cards)
    vals = dag_eval.forward(ns, X)
    ll = vals[ns][0]
    total = np.exp(ll).sum()
    assert abs(total - 1.0) < 1e-12, total


def test_forward_closed_form_simple_model():
    """Per-vector values as in simple_model_test.py:240-277 (products add, sums are matmuls)."""
    ns = kat_circuits.simple_model_dag()
    _compiled(ns, layer_sparsity_tol=0.1)
    g = np.random.default_rng(0)
    X = g.integers(0, 4, size=[32, 4])
    vals = dag_eval.forward(ns, X)
    np4, np5, np6 = ns.chs
    ns0, ni2, ni3 = np4.chs
    np0, np3 = ns0.chs
    ni0, ni1 = np0.chs
    th0 = ni0._params.numpy().reshape(-1, 4).astype(np.float64)
    assert np.allclose(vals[ni0], np.log(th0[:, X[:, 0]]), atol=1e-12)
    ch = np.concatenate([vals[ni0] + vals[ni1], vals[ni0] + vals[ni1]], axis=0)
    W = ns0._params.numpy().astype(np.float64).reshape(2, 4, 16, 16).transpose(0, 2, 1, 3).reshape(32, 64)
    ref = np.log(W @ np.exp(ch))
    assert np.allclose(vals[ns0], ref, atol=1e-10)
    assert np.allclose(vals[np4], vals[ns0] + vals[ni2] + vals[ni3], atol=1e-12)


def _total_ll(ns, X):
    return dag_eval.forward(ns, X)[ns].sum()


@pytest.mark.parametrize("case", sorted(CASES))
def test_param_flows_are_scaled_gradients(case):
    build, cards = CASES[case]
    ns = build()
    _compiled(ns, layer_sparsity_tol=0.1, max_tied_ns_per_parflow_block=2)
    g = np.random.default_rng(1)
    X = np.stack([g.integers(0, k, size=24) for k in cards], axis=1)
    vals = dag_eval.forward(ns, X)
    flows, pflows = dag_eval.backward(ns, X, vals)
    # tied copies accumulate into their source
    tot = dict()
    for n, pf in pflows.items():
        s = n.get_source_ns() if n.is_tied() else n
        tot[s] = tot.get(s, 0.0) + pf
    h = 1e-6
    checked = 0
    for s, pf in tot.items():
        if s.is_sum():
            P = s._params.double()
            s._params = P
            E = s.edge_ids.numpy()
            for _ in range(3):
                e = int(g.integers(0, P.shape[0]))
                i = int(g.integers(0, P.shape[1]))
                c = int(g.integers(0, P.shape[2]))
                th = float(P[e, i, c])
                P[e, i, c] = th * (1 + h)
                lp = _total_ll(ns, X)
                P[e, i, c] = th * (1 - h)
                lm = _total_ll(ns, X)
                P[e, i, c] = th
                fd = (lp - lm) / (2 * h)
                row = E[0, e] * s.block_size + i
                col = E[1, e] * s.ch_block_size + c
                assert abs(fd - pf[row, col]) < 1e-6 * max(1.0, abs(fd)), (s, fd, pf[row, col])
                checked += 1
        elif s.dist.get_signature() == "Categorical":
            P = s._params.double()
            s._params = P
            K = s.dist.num_cats
            for _ in range(3):
                i = int(g.integers(0, s.num_nodes))
                k = int(g.integers(0, K))
                th = float(P[i * K + k])
                P[i * K + k] = th * (1 + h)
                lp = _total_ll(ns, X)
                P[i * K + k] = th * (1 - h)
                lm = _total_ll(ns, X)
                P[i * K + k] = th
                fd = (lp - lm) / (2 * h)
                assert abs(fd - pf[i, k]) < 1e-6 * max(1.0, abs(fd)), (fd, pf[i, k])
                checked += 1
    assert checked > 0


@pytest.mark.parametrize("case", sorted(CASES))
def test_flow_conservation(case):
    build, cards = CASES[case]
    ns = build()
    _compiled(ns, layer_sparsity_tol=0.1, max_tied_ns_per_parflow_block=2)
    g = np.random.default_rng(2)
    X = np.stack([g.integers(0, k, size=16) for k in cards], axis=1)
    vals = dag_eval.forward(ns, X)
    flows, pflows = dag_eval.backward(ns, X, vals)
    for n, pf in pflows.items():
        if n.is_sum():
            assert np.allclose(pf.sum(axis=1), flows[n].sum(axis=1), rtol=1e-10, atol=1e-10)
        else:
            assert np.allclose(pf.sum(axis=1), flows[n].sum(axis=1), rtol=1e-10, atol=1e-10)
    # the unit flow of every root node is split exactly among the root's children
    assert np.allclose(sum(flows[c].sum(axis=0) for c in ns.chs), float(ns.num_nodes))


@pytest.mark.parametrize("case", sorted(CASES))
def test_em_full_step_monotone(case):
    build, cards = CASES[case]
    ns = build()
    _compiled(ns, layer_sparsity_tol=0.1, max_tied_ns_per_parflow_block=2)
    g = np.random.default_rng(3)
    X = np.stack([g.integers(0, k, size=64) for k in cards], axis=1)
    vals = dag_eval.forward(ns, X)
    before = vals[ns].sum()
    flows, pflows = dag_eval.backward(ns, X, vals)
    new = dag_eval.em_update(ns, pflows, step_size=1.0, pseudocount=0.0)
    for s, p in new.items():
        if s.is_sum():
            rows = p.sum(axis=1)
            assert np.allclose(rows, 1.0)
            E = s.edge_ids.numpy()
            blocks = p.reshape(s.num_node_blocks, s.block_size, s.num_ch_node_blocks, s.ch_block_size)
            s._params = torch.from_numpy(blocks[E[0], :, E[1], :].copy())
        else:
            assert np.allclose(p.sum(axis=1), 1.0) or p.shape[1] == 1
            s._params = torch.from_numpy(p.reshape(-1).copy())
    after = dag_eval.forward(ns, X)[ns].sum()
    assert after >= before - 1e-9, (before, after)


def test_sampling_matches_enumeration():
    ns = _hclt_bernoulli(num_vars=4, L=2, seed=5)
    pc = _compiled(ns)
    X = _all_assignments([2] * 4)
    p = np.exp(dag_eval.forward(ns, X)[ns][0])
    S = 40000
    U = np.random.default_rng(7).random([pc.num_nodes, S], dtype=np.float32)
    smp = dag_eval.sample(ns, U, pc.num_vars)
    idx = (smp * (2 ** np.arange(3, -1, -1))).sum(axis=1)
    freq = np.bincount(idx, minlength=16) / S
    sd = np.sqrt(p * (1 - p) / S)
    assert np.all(np.abs(freq - p) < 5 * sd + 1e-4), np.abs(freq - p) / (sd + 1e-12)


def test_philox_known_answers():
    """Philox4x32-10 against the generator's published known-answer vectors (Random123)."""
    cases = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
             ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
             ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
              (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in cases:
        got = dag_eval.philox4x32_10([np.array([c]) for c in ctr], key)
        assert tuple(int(g[0]) for g in got) == want
    u = dag_eval.philox_uniforms(7, np.arange(1000) % 13, np.arange(1000))
    assert u.dtype == np.float32 and (u >= 0).all() and (u < 1).all()
    assert np.all(u * 2 ** 24 == np.floor(u * 2 ** 24))      # 24-bit fractions, exact in FP32
    # uniform: mean and variance of 200k draws
    v = dag_eval.philox_uniforms(11, np.arange(200000) // 512, np.arange(200000) % 512).astype(np.float64)
    assert abs(v.mean() - 0.5) < 4e-3 and abs(v.var() - 1 / 12) < 4e-3


def test_chunked_pick_matches_single_scan_in_exact_arithmetic():
    """With weights and uniforms that are multiples of 2^-12 (exact FP32 sums), subtracting a
    chunk's total and scanning on is the same as one scan over all children, for any chunk size;
    a u at or past the total takes the last child; chunk sizes >= n are a single scan."""
    g = np.random.default_rng(3)
    for n in (1, 7, 100, 700, 1300):
        w = g.integers(0, 5, size=n).astype(np.float32) / 4096
        w[g.integers(0, n)] += 1 / 4096
        tot = float(w.astype(np.float64).sum())
        u = (g.integers(0, int(tot * 4096) + 3, size=300) / 4096).astype(np.float32)
        ref = np.minimum(np.searchsorted(np.cumsum(w.astype(np.float64)), u.astype(np.float64), side="right"), n - 1)
        for chunk in (1, 3, 64, 512, 4096):
            assert np.array_equal(dag_eval._chunked_pick(w, u, chunk), ref), (n, chunk)
