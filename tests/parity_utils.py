"""Helpers shared by the GPU parity tests: circuits, seeded data and oracle comparisons."""
import sys

import numpy as np
import torch

import kat_circuits
import pyjuice_amd as juice
import pyjuice_amd.nodes.distributions as dists
from oracle import dag_eval


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def cuda_ok():
    return torch.cuda.is_available()


def hclt_categorical(num_vars=12, K=8, L=32, block_size=None, n=256, seed=0):
    """HCLT over seeded chain-correlated categorical data."""
    g = np.random.default_rng(seed)
    x = np.zeros([n, num_vars], dtype=np.int64)
    x[:, 0] = g.integers(0, K, size=n)
    for v in range(1, num_vars):
        keep = g.random(n) < 0.7
        x[:, v] = np.where(keep, x[:, v - 1], g.integers(0, K, size=n))
    torch.manual_seed(seed)
    ns = juice.structures.HCLT(torch.from_numpy(x).float(), num_latents=L, block_size=block_size,
                               input_dist=dists.Categorical(num_cats=K), chunk_size=32)
    return ns, x, [K] * num_vars


def hclt_pruned(num_vars=10, K=8, L=128, block_size=32, frac=0.5, seed=0):
    """HCLT with a seeded fraction of its sum-edge blocks removed: uneven fan-in, so several sum
    layers split into more than one forward or backward partition."""
    ns, x, cards = hclt_categorical(num_vars=num_vars, K=K, L=L, block_size=block_size, seed=seed)
    juice.transformations.prune_sum_blocks(ns, frac, seed=4)
    return ns, x, cards


def wide_sums(block_size=64, seed=0, K=5):
    """A sum node block with more children than the EM kernel keeps in registers (320 children at
    block size 64, 640 at block size 32) under a block-size-1 root."""
    from pyjuice_amd.nodes import inputs, multiply, summate
    torch.manual_seed(seed)
    nb = 5 if block_size == 64 else 20
    with juice.set_block_size(block_size):
        a = [inputs(v, num_node_blocks=nb, dist=dists.Categorical(num_cats=K)) for v in range(2)]
        s = summate(multiply(a[0], a[1]), num_node_blocks=1)
        top = multiply(s)
    root = summate(top, num_node_blocks=1, block_size=1)
    root.init_parameters()
    return root, [K] * 2


def hclt_bernoulli(num_vars=16, L=8, n=512, seed=0):
    """The small Bernoulli HCLT of the CPU-scale workload: chain of flips with p ~ U(0.1, 0.4)."""
    g = np.random.Generator(np.random.PCG64(seed))
    p = g.uniform(0.1, 0.4, size=num_vars)
    x = np.zeros([n, num_vars], dtype=np.int64)
    x[:, 0] = g.random(n) < 0.5
    for v in range(1, num_vars):
        x[:, v] = x[:, v - 1] ^ (g.random(n) < p[v])
    torch.manual_seed(seed)
    ns = juice.structures.HCLT(torch.from_numpy(x).float(), num_latents=L, input_dist=dists.Bernoulli(), chunk_size=16)
    return ns, x, [2] * num_vars


def pd_small(seed=0):
    torch.manual_seed(seed)
    ns = juice.structures.PD(data_shape=(4, 4), num_latents=32, split_intervals=(2, 2),
                             input_dist=dists.Categorical(num_cats=6))
    return ns, [6] * 16


def pd_sparse(seed=0):
    """PD whose sparse block connections leave some product blocks without a parent edge."""
    torch.manual_seed(seed)
    ns = juice.structures.PD(data_shape=(8, 8), num_latents=16, split_intervals=(2, 2),
                             structure_type="sum_dominated", input_dist=dists.Categorical(num_cats=4))
    return ns, [4] * 64


def hmm_small(seed=0, T=6, L=32, V=10):
    torch.manual_seed(seed)
    ns = juice.structures.HMM(seq_length=T, num_latents=L, num_emits=V, homogeneous=True)
    return ns, [V] * T


def random_data(cards, B, seed):
    g = np.random.default_rng(seed)
    return np.stack([g.integers(0, k, size=B) for k in cards], axis=1).astype(np.int64)


def dense_pf_from_blocks(ns, blocks):
    """[E, bs, cbs] parameter-flow blocks -> dense [num_nodes, num_ch_nodes]."""
    E = ns.edge_ids.numpy()
    D = np.zeros([ns.num_node_blocks, ns.block_size, ns.num_ch_node_blocks, ns.ch_block_size])
    D[E[0], :, E[1], :] = blocks
    return D.reshape(ns.num_nodes, ns.num_ch_nodes)


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = max(np.abs(b).max(), 1e-30)
    return float(np.abs(a - b).max() / den)
