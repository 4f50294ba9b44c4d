"""Seeded synthetic workloads for the UMAP/densMAP fitting path (numpy, float32).

The reference's own generators (src/data/synthetic.rs) use rand 0.9 streams
that cannot be reproduced here; these generators define the workloads by their
distribution, with a fixed numpy seed:

* plain_cpu       n=10 000, d=50: 10 isotropic Gaussians, centres ~ N(0, 25 I)
* anisotropic     n=100 000, d=50: 20 Gaussians with covariance Q diag(l) Q^T,
                  l log-uniform in [0.1, 10], centres ~ N(0, 25 I), plus 5 %
                  uniform background over the bounding box
* subspace        n=1 000 000, d=64: 50 Gaussians on a random 8-D orthonormal
                  subspace (latent centres ~ N(0, 25 I_8), unit latent variance)
                  plus N(0, 0.05) ambient noise
* uneven_density  n=500 000, d=50: 30 Gaussians, per-cluster sd log-uniform in
                  [0.2, 5], centres ~ N(0, 25 I)
"""
from __future__ import annotations

import numpy as np


def plain_cpu(n: int = 10_000, d: int = 50, n_clusters: int = 10, seed: int = 42):
    rng = np.random.default_rng(seed)
    centres = rng.normal(0.0, 5.0, size=(n_clusters, d))
    labels = rng.integers(0, n_clusters, size=n)
    x = centres[labels] + rng.normal(0.0, 1.0, size=(n, d))
    return x.astype(np.float32), labels.astype(np.int32)


def anisotropic(n: int = 100_000, d: int = 50, n_clusters: int = 20, noise_frac: float = 0.05,
                seed: int = 42):
    rng = np.random.default_rng(seed)
    n_noise = int(round(n * noise_frac))
    n_clu = n - n_noise
    centres = rng.normal(0.0, 5.0, size=(n_clusters, d))
    labels = rng.integers(0, n_clusters, size=n_clu)
    x = np.empty((n, d), dtype=np.float64)
    for c in range(n_clusters):
        idx = np.nonzero(labels == c)[0]
        q, _ = np.linalg.qr(rng.normal(size=(d, d)))
        lam = np.exp(rng.uniform(np.log(0.1), np.log(10.0), size=d))
        z = rng.normal(size=(idx.size, d)) * np.sqrt(lam)
        x[idx] = centres[c] + z @ q.T
    lo, hi = x[:n_clu].min(axis=0), x[:n_clu].max(axis=0)
    x[n_clu:] = rng.uniform(lo, hi, size=(n_noise, d))
    all_labels = np.concatenate([labels, np.full(n_noise, -1)])
    perm = rng.permutation(n)
    return x[perm].astype(np.float32), all_labels[perm].astype(np.int32)


def subspace(n: int = 1_000_000, d: int = 64, n_clusters: int = 50, latent: int = 8,
             noise_var: float = 0.05, seed: int = 42):
    rng = np.random.default_rng(seed)
    basis, _ = np.linalg.qr(rng.normal(size=(d, latent)))
    centres = rng.normal(0.0, 5.0, size=(n_clusters, latent))
    labels = rng.integers(0, n_clusters, size=n)
    x = np.empty((n, d), dtype=np.float32)
    step = 200_000
    for s in range(0, n, step):
        e = min(n, s + step)
        z = centres[labels[s:e]] + rng.normal(size=(e - s, latent))
        x[s:e] = (z @ basis.T + rng.normal(0.0, np.sqrt(noise_var), size=(e - s, d))).astype(
            np.float32)
    return x, labels.astype(np.int32)


def uneven_density(n: int = 500_000, d: int = 50, n_clusters: int = 30, seed: int = 42):
    rng = np.random.default_rng(seed)
    centres = rng.normal(0.0, 5.0, size=(n_clusters, d))
    sds = np.exp(rng.uniform(np.log(0.2), np.log(5.0), size=n_clusters))
    labels = rng.integers(0, n_clusters, size=n)
    x = np.empty((n, d), dtype=np.float32)
    step = 200_000
    for s in range(0, n, step):
        e = min(n, s + step)
        lab = labels[s:e]
        x[s:e] = (centres[lab] + rng.normal(size=(e - s, d)) * sds[lab][:, None]).astype(
            np.float32)
    return x, labels.astype(np.int32)


def random_init(n: int, n_dim: int = 2, seed: int = 42, rng_range: float = 10.0):
    """Uniform init in [-range, range] (the range of random_layout, src/data/init.rs:1024-1043)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(-rng_range, rng_range, size=(n, n_dim)).astype(np.float32)
