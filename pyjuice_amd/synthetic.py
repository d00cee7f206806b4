"""Seeded synthetic data for the benchmark / test workloads (no dataset downloads).

* ``bernoulli_chain``: 16-ish binary variables, x_0 ~ Bern(0.5), x_i = x_{i-1} XOR Bern(p_i)
  with p_i ~ U(0.1, 0.4) drawn once (``numpy PCG64(seed)``).
* ``image_categorical``: H x W pixels with K categories.  A Gaussian copula with separable
  AR(1) correlation rho^(|di| + |dj|) (Z = A eps A^T, A = chol of the AR(1) matrix) gives
  uniforms U = Phi(Z); pixel v has the marginal p_v(k) ~ 0.95 N(k; mu_v, sigma_v) + 0.05 / K
  with mu_v, sigma_v from a smooth seeded field; x_v = F_v^{-1}(U_v).
* ``hmm_sequences``: sequences from a random HMM whose initial, transition and emission rows
  are Dirichlet(alpha) draws.
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["bernoulli_chain", "image_categorical", "hmm_sequences"]


def bernoulli_chain(n: int, num_vars: int = 16, seed: int = 0) -> np.ndarray:
    g = np.random.Generator(np.random.PCG64(seed))
    p = g.uniform(0.1, 0.4, size=num_vars)
    x = np.zeros([n, num_vars], dtype=np.uint8)
    x[:, 0] = g.random(n) < 0.5
    for i in range(1, num_vars):
        x[:, i] = x[:, i - 1] ^ (g.random(n) < p[i])
    return x


def _ar1_chol(n: int, rho: float) -> np.ndarray:
    idx = np.arange(n)
    R = rho ** np.abs(idx[:, None] - idx[None, :])
    return np.linalg.cholesky(R)


def _smooth_field(g: np.random.Generator, H: int, W: int, terms: int = 6) -> np.ndarray:
    yy, xx = np.meshgrid(np.arange(H) / H, np.arange(W) / W, indexing="ij")
    f = np.zeros([H, W])
    for _ in range(terms):
        a, b = g.uniform(0.5, 2.5, size=2)
        ph = g.uniform(0, 2 * math.pi, size=2)
        f += np.sin(2 * math.pi * a * yy + ph[0]) * np.cos(2 * math.pi * b * xx + ph[1])
    f -= f.min()
    f /= max(f.max(), 1e-12)
    return f  # in [0, 1]


def _std_normal_cdf(z: np.ndarray) -> np.ndarray:
    from scipy.special import ndtr
    return ndtr(z)


def image_categorical(n: int, H: int = 28, W: int = 28, K: int = 256, rho: float = 0.6, seed: int = 0,
                      field_seed: int = 1, chunk: int = 4096) -> np.ndarray:
    g = np.random.Generator(np.random.PCG64(seed))
    gf = np.random.Generator(np.random.PCG64(field_seed))
    Ah, Aw = _ar1_chol(H, rho), _ar1_chol(W, rho)
    mu = 0.15 * K + 0.7 * K * _smooth_field(gf, H, W)
    sigma = 0.04 * K + 0.12 * K * _smooth_field(gf, H, W)
    k = np.arange(K)
    dens = np.exp(-0.5 * ((k[None, None, :] - mu[:, :, None]) / sigma[:, :, None]) ** 2)
    dens /= dens.sum(axis=2, keepdims=True)
    probs = 0.95 * dens + 0.05 / K
    cdf = np.cumsum(probs, axis=2).reshape(H * W, K)
    cdf[:, -1] = 1.0
    out = np.empty([n, H * W], dtype=np.uint8 if K <= 256 else np.int64)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        eps = g.standard_normal([e - s, H, W])
        Z = np.einsum("ij,bjk,lk->bil", Ah, eps, Aw).reshape(e - s, H * W)
        U = _std_normal_cdf(Z)
        # inverse CDF per pixel: number of cdf entries strictly below u
        for v in range(H * W):
            out[s:e, v] = np.searchsorted(cdf[v], U[:, v], side="left").clip(0, K - 1)
    return out


def hmm_sequences(n: int, seq_length: int, num_latents: int, num_emits: int, alpha: float = 0.1,
                  seed: int = 2) -> np.ndarray:
    g = np.random.Generator(np.random.PCG64(seed))
    init = g.dirichlet(np.full(num_latents, alpha))
    trans = g.dirichlet(np.full(num_latents, alpha), size=num_latents)
    emit = g.dirichlet(np.full(num_emits, alpha), size=num_latents)
    tc = np.cumsum(trans, axis=1)
    ec = np.cumsum(emit, axis=1)
    z = np.searchsorted(np.cumsum(init), g.random(n)).clip(0, num_latents - 1)
    out = np.empty([n, seq_length], dtype=np.int64)
    for t in range(seq_length):
        u = g.random(n)
        out[:, t] = (ec[z] < u[:, None]).sum(axis=1).clip(0, num_emits - 1)
        u = g.random(n)
        z = (tc[z] < u[:, None]).sum(axis=1).clip(0, num_latents - 1)
    return out
