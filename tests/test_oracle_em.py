"""Pins the oracle's EM update (``dag_eval.em_update`` and ``pc_cpu.c``'s EM) with checks that a
dropped or misplaced pseudocount term fails (CPU only).

For a node with n children, flows pf and pseudocount pc, the reference moves theta towards
``(pf_c + pc / n) / (sum_c pf_c + pc)`` (``model/backend/par_update.py:214-256``,
``nodes/distributions/categorical.py:164-201``; Bernoulli ``bernoulli.py:91-101``).  What the
mathematics fixes about that target, independently of the formula:

* it is the maximiser over the simplex of ``sum_c (pf_c + pc / n) log theta_c`` (the MAP
  estimate under a symmetric Dirichlet prior that adds pc in total): checked by numerical
  optimisation with scipy, so a denominator without ``+ pc`` (rows summing to more than one) or
  a prior of ``pc / K`` with the wrong K (a different maximiser) fails;
* rows sum to one for any flows; zero flows give the uniform distribution 1 / n;
* a step s interpolates linearly between theta and the full-step target.

``keep_zero_params`` (``par_update.py:196-205, 236-253``, ``categorical.py:171-201``): zero
parameters stay zero and the others are the same MAP restricted to the support (n still counts
every child).  The closed form of ``tests/model/simple_model_test.py:490-528`` (step 0.25,
pseudocount 0.01, per-vector child counts 64 / 32 / 32 / 96) is checked on the reference's own
circuit.
"""
import numpy as np
import pytest
import torch
from scipy.optimize import minimize, minimize_scalar

import kat_circuits
import parity_utils as U
from oracle import dag_eval
from oracle.pc_cpu import CompiledCPU
from pyjuice_amd.model import TensorCircuit
from pyjuice_amd.nodes.nodes import node_iterator


def _edge_mask(s):
    E = s.edge_ids.numpy()
    mb = np.zeros([s.num_node_blocks, s.block_size, s.num_ch_node_blocks, s.ch_block_size], dtype=bool)
    mb[E[0], :, E[1], :] = True
    return mb.reshape(s.num_nodes, s.num_ch_nodes)


def _random_pflows(ns, g, zero=False):
    """Arbitrary non-negative flows (not from a backward pass) for every source vector."""
    out = dict()
    for n in node_iterator(ns):
        if n.is_tied() or not (n.is_sum() or n.is_input()):
            continue
        if n.is_sum():
            m = _edge_mask(n)
            out[n] = np.where(m, 0.0 if zero else g.gamma(0.7, 3.0, size=m.shape), 0.0)
        else:
            K = n.dist.num_param_flows()
            out[n] = np.zeros([n.num_nodes, K]) if zero else g.gamma(0.7, 3.0, size=[n.num_nodes, K])
    return out


def _map_on_simplex(w):
    """argmax_theta sum_c w_c log theta_c over the simplex, by BFGS on softmax logits."""
    w = np.asarray(w, np.float64)

    def f(z):
        z = z - z.max()
        lse = np.log(np.exp(z).sum())
        val = -(w * (z - lse)).sum()
        p = np.exp(z - lse)
        return val, -(w - w.sum() * p)

    r = minimize(f, np.zeros_like(w), jac=True, method="BFGS", options=dict(gtol=1e-12, maxiter=10000))
    z = r.x - r.x.max()
    return np.exp(z) / np.exp(z).sum()


def _bernoulli_map(f1, f0, pc):
    r = minimize_scalar(lambda p: -((f1 + pc / 2) * np.log(p) + (f0 + pc / 2) * np.log1p(-p)),
                        bounds=(1e-12, 1 - 1e-12), method="bounded", options=dict(xatol=1e-13))
    return r.x


CIRCUITS = {
    "simple_model": (lambda: kat_circuits.simple_model_dag(), dict(layer_sparsity_tol=0.1)),
    "tying_block1": (lambda: kat_circuits.tying_block1_dag(), dict(max_tied_ns_per_parflow_block=2)),
    "hclt_bern16": (lambda: U.hclt_bernoulli(num_vars=6, L=4)[0], dict()),
}


@pytest.mark.parametrize("name", sorted(CIRCUITS))
@pytest.mark.parametrize("pc", [0.01, 0.5, 4.0])
def test_em_target_is_the_map_estimate(name, pc):
    build, kw = CIRCUITS[name]
    ns = build()
    TensorCircuit(ns, **kw).update_parameters()
    g = np.random.default_rng(17)
    pf = _random_pflows(ns, g)
    new = dag_eval.em_update(ns, pf, step_size=1.0, pseudocount=pc)
    checked = 0
    for s, p in new.items():
        if s.is_sum():
            m = _edge_mask(s)
            assert np.allclose(p.sum(axis=1), 1.0, atol=1e-12)
            n = m.sum(axis=1)
            for r in g.choice(s.num_nodes, size=min(3, s.num_nodes), replace=False):
                w = pf[s][r][m[r]] + pc / n[r]
                assert np.abs(p[r][m[r]] - _map_on_simplex(w)).max() < 1e-6
                checked += 1
        elif s.dist.get_signature() == "Categorical":
            K = p.shape[1]
            assert np.allclose(p.sum(axis=1), 1.0, atol=1e-12)
            for r in g.choice(s.num_nodes, size=min(3, s.num_nodes), replace=False):
                assert np.abs(p[r] - _map_on_simplex(pf[s][r] + pc / K)).max() < 1e-6
                checked += 1
        else:
            for r in range(min(3, s.num_nodes)):
                assert abs(p[r, 0] - _bernoulli_map(pf[s][r, 0], pf[s][r, 1], pc)) < 1e-7
                checked += 1
    assert checked > 0


@pytest.mark.parametrize("name", sorted(CIRCUITS))
def test_em_zero_flows_give_uniform_and_steps_interpolate(name):
    build, kw = CIRCUITS[name]
    ns = build()
    TensorCircuit(ns, **kw).update_parameters()
    zero = _random_pflows(ns, np.random.default_rng(0), zero=True)
    full = dag_eval.em_update(ns, zero, step_size=1.0, pseudocount=0.3)
    for s, p in full.items():
        if s.is_sum():
            m = _edge_mask(s)
            n = m.sum(axis=1, keepdims=True)
            assert np.allclose(p, np.where(m, 1.0 / n, 0.0), atol=1e-14)
        elif s.dist.get_signature() == "Categorical":
            assert np.allclose(p, 1.0 / p.shape[1], atol=1e-14)
        else:
            assert np.allclose(p, 0.5, atol=1e-14)
    pf = _random_pflows(ns, np.random.default_rng(5))
    one = dag_eval.em_update(ns, pf, step_size=1.0, pseudocount=0.2)
    part = dag_eval.em_update(ns, pf, step_size=0.3, pseudocount=0.2)
    for s in one:
        th = dag_eval.sum_dense_params(s) if s.is_sum() else dag_eval.input_params(s)
        th = th if not s.is_input() or s.dist.get_signature() == "Categorical" else th[:, :1]
        assert np.allclose(part[s], 0.7 * th + 0.3 * one[s], atol=1e-13)


def _zero_some(ns, g, frac=0.3):
    """Sets a seeded fraction of the sum and Categorical parameters to 0 (at least one kept per
    node); returns the zero masks."""
    zmask = dict()
    for s in node_iterator(ns):
        if s.is_tied():
            continue
        if s.is_sum():
            P = s._params.double().numpy().copy()      # [E, bs, cbs]
            z = g.random(P.shape) < frac
            P[z] = 0.0
            s._params = torch.from_numpy(P)
            W = dag_eval.sum_dense_params(s)
            m = _edge_mask(s)
            dead = m & (W == 0)
            for r in np.nonzero((~dead & m).sum(axis=1) == 0)[0]:   # keep one child alive
                c = np.nonzero(m[r])[0][0]
                blk, i = divmod(r, s.block_size)
                cb, cc = divmod(c, s.ch_block_size)
                e = np.nonzero((s.edge_ids[0].numpy() == blk) & (s.edge_ids[1].numpy() == cb))[0][0]
                P[e, i, cc] = 0.5
            s._params = torch.from_numpy(P)
            zmask[s] = _edge_mask(s) & (dag_eval.sum_dense_params(s) == 0)
        elif s.is_input() and s.dist.get_signature() == "Categorical":
            K = s.dist.num_cats
            P = s._params.double().numpy().reshape(-1, K).copy()
            z = g.random(P.shape) < frac
            z[:, 0] = False
            P[z] = 0.0
            s._params = torch.from_numpy(P.reshape(-1))
            zmask[s] = z
    return zmask


@pytest.mark.parametrize("name", sorted(CIRCUITS))
def test_em_keep_zero_params(name):
    build, kw = CIRCUITS[name]
    ns = build()
    TensorCircuit(ns, **kw).update_parameters()
    g = np.random.default_rng(23)
    zmask = _zero_some(ns, g)
    pf = _random_pflows(ns, g)
    pc = 0.4
    new = dag_eval.em_update(ns, pf, step_size=1.0, pseudocount=pc, keep_zero_params=True)
    checked = 0
    for s, z in zmask.items():
        p = new[s]
        assert np.all(p[z] == 0.0)
        if s.is_sum():
            m = _edge_mask(s)
            live = m & ~z
            n = m.sum(axis=1)
        else:
            live = ~z
            n = np.full(s.num_nodes, z.shape[1])
        assert np.allclose(np.where(live, p, 0.0).sum(axis=1), 1.0, atol=1e-12)
        for r in range(min(3, s.num_nodes)):
            w = pf[s][r][live[r]] + pc / n[r]
            assert np.abs(p[r][live[r]] - _map_on_simplex(w)).max() < 1e-6
            checked += 1
    assert checked > 0
    # without the flag, zero parameters move away from zero
    moved = dag_eval.em_update(ns, pf, step_size=1.0, pseudocount=pc, keep_zero_params=False)
    assert any(np.all(moved[s][z] > 0) for s, z in zmask.items() if z.any())


def test_simple_model_em_closed_form():
    """``tests/model/simple_model_test.py:490-528``: step 0.25, pseudocount 0.01; the child counts
    of the four sum vectors are 64, 32, 32 and 96."""
    ns = kat_circuits.simple_model_dag()
    TensorCircuit(ns, layer_sparsity_tol=0.1).update_parameters()
    np4, np5, np6 = ns.chs
    ns0 = np4.chs[0]
    ns1 = np5.chs[0]
    ns2 = np6.chs[0]
    g = np.random.default_rng(0)
    X = np.stack([g.integers(0, k, size=64) for k in [4, 4, 6, 6]], axis=1)
    vals = dag_eval.forward(ns, X)
    _, pflows = dag_eval.backward(ns, X, vals)
    step, pc = 0.25, 0.01
    new = dag_eval.em_update(ns, pflows, step_size=step, pseudocount=pc)
    for s, nch in ((ns0, 64), (ns1, 32), (ns2, 32), (ns, 96)):
        f = pflows[s]
        old = dag_eval.sum_dense_params(s)
        assert int(_edge_mask(s).sum(axis=1)[0]) == nch
        target = (f + pc / nch) / (f.sum(axis=1, keepdims=True) + pc)
        assert np.abs(new[s] - ((1 - step) * old + step * target)).max() < 1e-12


@pytest.mark.parametrize("name", ["simple_model", "tying_block1", "hclt_cat_bs32"])
def test_compiled_cpu_keep_zero_matches_dag_oracle(name):
    if name == "hclt_cat_bs32":
        ns, _, cards = U.hclt_categorical(L=64, block_size=32)
        kw = dict()
    else:
        build, kw = CIRCUITS[name]
        ns = build()
        cards = [4, 4, 6, 6] if name == "simple_model" else [5, 5, 5, 5]
    pc_ = TensorCircuit(ns, **kw)
    pc_.update_parameters()
    g = np.random.default_rng(31)
    _zero_some(ns, g, frac=0.25)
    pc_ = TensorCircuit(ns, **kw)      # recompiled with the zeroed parameters
    pc_.update_parameters()
    X = U.random_data(cards, 24, seed=3)
    cpu = CompiledCPU(pc_)
    cpu.forward(X)
    cpu.backward()
    vals = dag_eval.forward(ns, X)
    _, pflows = dag_eval.backward(ns, X, vals)
    cpu.mini_batch_em(step_size=0.6, pseudocount=0.2, keep_zero_params=True)
    new = dag_eval.em_update(ns, pflows, step_size=0.6, pseudocount=0.2, keep_zero_params=True)
    pc_.params.data = torch.from_numpy(cpu.params.astype(np.float32))
    pc_.update_parameters()
    for layer, d in zip(pc_.input_layer_group, cpu.inputs):
        for n in layer.nodes:
            if not n.is_tied():
                a, b = n._param_range
                n._params = torch.from_numpy(d["params"][a:b].copy())
    for s, p in new.items():
        got = U.dense_pf_from_blocks(s, s._params.numpy()) if s.is_sum() else \
            s._params.numpy().astype(np.float64).reshape(p.shape)
        assert np.abs(got - p).max() <= 1e-6, (s, np.abs(got - p).max())
