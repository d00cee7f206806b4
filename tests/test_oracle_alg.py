"""The FP64 oracle's MPE and GeneralLL propagation (CPU only), pinned by what the mathematics fixes:

* GeneralLL with alpha = 1 is LL: values, flows and parameter flows equal the LL oracle;
* GeneralLL tends to MPE: (1/a) log sum exp(a x) lies in [max x, max x + log(n) / a];
* along the MPE path (flows of one sample from a unit root flow), the log-weights of the taken
  sum edges plus the inputs' log-likelihoods add up to the root's MPE value;
* each sum node's flow is split over its children (shares sum to one), so with a unit root flow
  the flows into each variable's input nodes sum to one per sample.
"""
import numpy as np
import pytest

import kat_circuits
import parity_utils as U
from oracle import dag_eval
from pyjuice_amd.nodes.nodes import node_iterator


def _circuit(name):
    if name == "hmm":
        ns, cards = U.hmm_small(T=4, L=8, V=5)
    elif name == "hclt":
        ns, x, cards = U.hclt_categorical(num_vars=6, K=4, L=8, n=64)
    elif name == "pd":
        ns, cards = U.pd_small()
    else:
        ns, cards = kat_circuits.simple_model_dag(), [4, 4, 6, 6]
    ns.init_parameters()
    return ns, cards


NAMES = ["hmm", "hclt", "pd", "simple"]


@pytest.mark.parametrize("name", NAMES)
def test_general_ll_alpha_one_is_ll(name):
    ns, cards = _circuit(name)
    X = U.random_data(cards, 12, seed=1)
    v0 = dag_eval.forward(ns, X)
    v1 = dag_eval.forward(ns, X, alg="GeneralLL", alpha=1.0)
    for n in node_iterator(ns):
        np.testing.assert_allclose(v1[n], v0[n], rtol=1e-10, atol=1e-10)
    f0, p0 = dag_eval.backward(ns, X, v0)
    f1, p1 = dag_eval.backward(ns, X, v1, alg="GeneralLL", alpha=1.0)
    for n in node_iterator(ns):
        np.testing.assert_allclose(f1[n], f0[n], rtol=1e-9, atol=1e-12)
        if n in p0:
            np.testing.assert_allclose(p1[n], p0[n], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("name", NAMES)
def test_general_ll_tends_to_mpe(name):
    ns, cards = _circuit(name)
    X = U.random_data(cards, 12, seed=2)
    vm = dag_eval.forward(ns, X, alg="MPE")
    a = 4000.0
    va = dag_eval.forward(ns, X, alg="GeneralLL", alpha=a)
    # each sum level adds at most log(#children) / a above the max; products add their children
    slack = 0.0
    for n in node_iterator(ns):
        if n.is_sum():
            slack = max(slack, np.log(sum(c.num_nodes for c in n.chs)) / a)
    depth = sum(1 for n in node_iterator(ns) if not n.is_input())
    for n in node_iterator(ns):
        d = va[n] - vm[n]
        fin = np.isfinite(vm[n])
        assert np.all(d[fin] >= -1e-9)
        assert np.all(d[fin] <= depth * slack + 1e-9)


@pytest.mark.parametrize("name", NAMES)
def test_mpe_path_adds_up_to_mpe_value(name):
    ns, cards = _circuit(name)
    X = U.random_data(cards, 6, seed=3)
    for b in range(X.shape[0]):
        xb = X[b:b + 1]
        v = dag_eval.forward(ns, xb, alg="MPE")
        f, pf = dag_eval.backward(ns, xb, v, alg="MPE")
        total = 0.0
        for n in node_iterator(ns):
            if n.is_sum():
                W = dag_eval.sum_dense_params(n)
                taken = pf[n] > 0
                total += float((pf[n][taken] * np.log(W[taken])).sum())
                # one sample, unique maxima: a node's flow goes to exactly one child
                share = pf[n].sum(axis=1)
                np.testing.assert_allclose(share, f[n][:, 0], rtol=1e-12, atol=1e-12)
            elif n.is_input():
                on = f[n][:, 0] > 0
                total += float((f[n][on, 0] * v[n][on, 0]).sum())
        np.testing.assert_allclose(total, v[ns][0, 0], rtol=1e-10, atol=1e-9)


@pytest.mark.parametrize("name", ["hmm", "hclt", "simple"])
@pytest.mark.parametrize("alg,alpha", [("MPE", 0.0), ("GeneralLL", 0.5), ("GeneralLL", 2.5)])
def test_flows_reach_every_variable_once(name, alg, alpha):
    ns, cards = _circuit(name)
    X = U.random_data(cards, 10, seed=4)
    v = dag_eval.forward(ns, X, alg=alg, alpha=alpha)
    f, pf = dag_eval.backward(ns, X, v, alg=alg, alpha=alpha)
    per_var = np.zeros([len(cards), X.shape[0]])
    for n in node_iterator(ns):
        if n.is_input():
            per_var[n.scope.to_list()[0]] += f[n].sum(axis=0)
    np.testing.assert_allclose(per_var, 1.0, rtol=1e-9)


def test_per_vector_algorithm_map():
    """A map from sum vectors to algorithms reduces to the single-algorithm oracle when every
    vector takes the same one, and mixes per vector otherwise (a vector's value depends only on
    its own algorithm given its children)."""
    import kat_circuits
    from pyjuice_amd.model import TensorCircuit
    from pyjuice_amd.nodes.nodes import node_iterator
    ns = kat_circuits.simple_model_dag()
    TensorCircuit(ns, layer_sparsity_tol=0.1).update_parameters()
    X = np.random.default_rng(1).integers(0, 4, size=[16, 4])
    sums = [n for n in node_iterator(ns) if n.is_sum()]
    for a in ("LL", "MPE", "GeneralLL"):
        one = dag_eval.forward(ns, X, alg=a, alpha=0.5)
        mp = dag_eval.forward(ns, X, alg={n: a for n in sums}, alpha=0.5)
        for n in sums:
            assert np.array_equal(one[n], mp[n])
        f1, p1 = dag_eval.backward(ns, X, one, alg=a, alpha=0.5)
        f2, p2 = dag_eval.backward(ns, X, mp, alg={n: a for n in sums}, alpha=0.5)
        for n in sums:
            assert np.array_equal(f1[n], f2[n]) and np.array_equal(p1[n], p2[n])
    mixed = {n: ("MPE" if i % 2 else "LL") for i, n in enumerate(sums)}
    mv = dag_eval.forward(ns, X, alg=mixed)
    ll, mpe = dag_eval.forward(ns, X), dag_eval.forward(ns, X, alg="MPE")
    for i, n in enumerate(sums):
        child = np.concatenate([mv[c] for c in n.chs], axis=0)
        W = dag_eval.sum_dense_params(n)
        ref = dag_eval._alg_sum(W, child, "MPE", 1.0) if i % 2 else dag_eval._lse_matmul(W, child)
        assert np.allclose(mv[n], ref)
