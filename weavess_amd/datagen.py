"""Seeded synthetic workloads (the reference's real datasets are offline).

Every generator is deterministic in its seed (numpy PCG64 via SeedSequence;
large sets are generated in independent chunks so memory stays bounded).

  c1  10,000 base + 1,000 queries, dim 128: 20-cluster Gaussian mixture,
      centres U[0,255]^128, sigma 20, rounded and clipped to [0,255]
      (integer-valued, so every FP32 partial sum is exact).
  c2  1,000,000 base + 10,000 queries, dim 128: 1,000 clusters, centres
      U[-1,1]^128, sigma 0.15, cluster sizes Zipf(1.1).
  c3  1,000,000 base + 1,000 queries, dim 960, non-negative: 500 clusters,
      centre_k = |N(0,1)| * s with per-dimension s ~ U[0.1,1],
      x = |centre_k + 0.3 * N(0,1) * s|.
  c4  2,000,000 base + 200,000 queries, dim 96, unit norm: 2,000 clusters,
      centres uniform on the unit sphere, x = centre + 0.3 * N(0,I) (sigma
      0.3 per dimension before normalisation), then x / |x|; per-query
      L ~ U{16..400}, k ~ U{1,10,100}.
  c4m as c4 but each cluster has its own spread, sigma_c log-uniform in
      [0.05, 0.6]: query difficulty varies across clusters, the setting an
      adaptive (learned) stop rule is for.
All query sets come from the same mixture as their base set.
"""
from __future__ import annotations

import numpy as np

CHUNK = 1 << 17


def _rng(seed: int, *key: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(np.random.SeedSequence([seed, *key])))


def _mixture(n: int, centres: np.ndarray, probs: np.ndarray, sample, seed: int, stream: int, d=None):
    d = centres.shape[1] if d is None else d
    out = np.empty((n, d), np.float32)
    for ci, s0 in enumerate(range(0, n, CHUNK)):
        m = min(CHUNK, n - s0)
        g = _rng(seed, stream, ci)
        lab = g.choice(len(centres), size=m, p=probs)
        out[s0:s0 + m] = sample(g, centres[lab])
    return out


def gmm_c1(n_base=10_000, n_query=1_000, d=128, clusters=20, sigma=20.0, seed=1):
    g = _rng(seed, 0)
    centres = g.uniform(0.0, 255.0, size=(clusters, d))
    probs = np.full(clusters, 1.0 / clusters)

    def sample(gg, c):
        return np.clip(np.rint(c + gg.normal(0.0, sigma, size=c.shape)), 0, 255).astype(np.float32)

    return (_mixture(n_base, centres, probs, sample, seed, 1),
            _mixture(n_query, centres, probs, sample, seed, 2))


def gmm_c2(n_base=1_000_000, n_query=10_000, d=128, clusters=1_000, sigma=0.15, zipf=1.1, seed=2):
    g = _rng(seed, 0)
    centres = g.uniform(-1.0, 1.0, size=(clusters, d))
    w = 1.0 / np.arange(1, clusters + 1) ** zipf
    probs = w / w.sum()

    def sample(gg, c):
        return (c + gg.normal(0.0, sigma, size=c.shape)).astype(np.float32)

    return (_mixture(n_base, centres, probs, sample, seed, 1),
            _mixture(n_query, centres, probs, sample, seed, 2))


def gmm_c3(n_base=1_000_000, n_query=1_000, d=960, clusters=500, noise=0.3, seed=3):
    g = _rng(seed, 0)
    scale = g.uniform(0.1, 1.0, size=d)
    centres = np.abs(g.normal(0.0, 1.0, size=(clusters, d))) * scale
    probs = np.full(clusters, 1.0 / clusters)

    def sample(gg, c):
        return np.abs(c + noise * gg.normal(0.0, 1.0, size=c.shape) * scale).astype(np.float32)

    return (_mixture(n_base, centres, probs, sample, seed, 1),
            _mixture(n_query, centres, probs, sample, seed, 2))


def gmm_c4(n_base=2_000_000, n_query=200_000, d=96, clusters=2_000, sigma=0.3, seed=4):
    g = _rng(seed, 0)
    centres = g.normal(0.0, 1.0, size=(clusters, d))
    centres /= np.linalg.norm(centres, axis=1, keepdims=True)
    probs = np.full(clusters, 1.0 / clusters)

    def sample(gg, c):
        x = c + gg.normal(0.0, sigma, size=c.shape)
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        return x.astype(np.float32)

    return (_mixture(n_base, centres, probs, sample, seed, 1),
            _mixture(n_query, centres, probs, sample, seed, 2))


def gmm_c4m(n_base=2_000_000, n_query=200_000, d=96, clusters=2_000, sigma_lo=0.05, sigma_hi=0.6, seed=5):
    g = _rng(seed, 0)
    centres = g.normal(0.0, 1.0, size=(clusters, d))
    centres /= np.linalg.norm(centres, axis=1, keepdims=True)
    sig = np.exp(g.uniform(np.log(sigma_lo), np.log(sigma_hi), size=clusters))
    table = np.concatenate([centres, sig[:, None]], axis=1)  # last column: the cluster's sigma
    probs = np.full(clusters, 1.0 / clusters)

    def sample(gg, c):
        x = c[:, :d] + gg.normal(0.0, 1.0, size=(c.shape[0], d)) * c[:, d:]
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        return x.astype(np.float32)

    return (_mixture(n_base, table, probs, sample, seed, 1, d),
            _mixture(n_query, table, probs, sample, seed, 2, d))


def c4_query_budgets(n_query: int, seed=4, L_lo=16, L_hi=400, ks=(1, 10, 100)):
    """Per-query (L, k) of the batching workload; L_eff = max(L, k)."""
    g = _rng(seed, 9)
    L = g.integers(L_lo, L_hi + 1, size=n_query).astype(np.uint32)
    k = np.asarray(ks, np.uint32)[g.integers(0, len(ks), size=n_query)]
    return np.maximum(L, k).astype(np.uint32), k.astype(np.uint32)


def small(n_base=2_000, n_query=200, d=128, clusters=10, seed=7, integer=True):
    """Small C1-like set for tests (integer-valued when `integer`)."""
    g = _rng(seed, 0)
    if integer:
        centres = g.uniform(0.0, 255.0, size=(clusters, d))

        def sample(gg, c):
            return np.clip(np.rint(c + gg.normal(0.0, 20.0, size=c.shape)), 0, 255).astype(np.float32)
    else:
        centres = g.uniform(-1.0, 1.0, size=(clusters, d))

        def sample(gg, c):
            return (c + gg.normal(0.0, 0.15, size=c.shape)).astype(np.float32)
    probs = np.full(clusters, 1.0 / clusters)
    return (_mixture(n_base, centres, probs, sample, seed, 1),
            _mixture(n_query, centres, probs, sample, seed, 2))
