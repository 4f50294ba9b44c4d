"""Shared test fixtures: reference test setups and small seeded graphs."""
from __future__ import annotations

import numpy as np

import mrs_amd
import pyoracle as orc


def triangle_setup():
    """triangle_setup, src/training/umap_optimiser_gpu.rs:1657-1680."""
    embd = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]], np.float32)
    graph = [[(1, 1.0), (2, 0.5)], [(0, 1.0), (2, 1.0)], [(0, 0.5), (1, 1.0)]]
    params = mrs_amd.UmapOptimParams(a=1.5, b=0.9, lr=1.0, gamma=1.0, n_epochs=10,
                                     neg_sample_rate=5, min_dist=0.1, beta1=0.9, beta2=0.999,
                                     eps=1e-7)
    return embd, graph, params


def coo_to_adjacency(n, rows, cols, vals):
    """coo_to_adjacency_list (src/data/graph.rs:709-758) for row-grouped COO."""
    rows = np.asarray(rows, np.int64)
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, rows + 1, 1)
    off = np.cumsum(off)
    order = np.argsort(rows, kind="stable")
    return off, np.asarray(cols, np.int32)[order], np.asarray(vals, np.float32)[order]


def umap_graph_from_data(x, k, n_epochs=500, mix_weight=1.0):
    """Exact kNN + smooth_knn_dist + knn_to_coo + symmetrise + filter with the oracle.

    Returns (knn_idx, knn_dist, (rows, cols, vals), (adj_off, adj_cols, adj_vals)).
    """
    n = x.shape[0]
    idx, dist = orc.knn_exact(x, k)
    sigma, rho = orc.smooth_knn_dist(dist, k)
    r, c, v = orc.knn_to_coo(idx, dist, sigma, rho)
    r, c, v = orc.symmetrise(n, r, c, v, mix_weight)
    r, c, v = orc.filter_weak_edges(r, c, v, n_epochs)
    adj = coo_to_adjacency(n, r, c, v)
    return idx, dist, (r, c, v), adj


def blobs(n, d, n_clusters, seed, spread=5.0):
    rng = np.random.default_rng(seed)
    centres = rng.normal(0, spread, size=(n_clusters, d))
    lab = rng.integers(0, n_clusters, size=n)
    return (centres[lab] + rng.normal(size=(n, d))).astype(np.float32), lab


def assert_close(got, want, tol, ctx=""):
    got = np.asarray(got, np.float64).ravel()
    want = np.asarray(want, np.float64).ravel()
    assert got.shape == want.shape, f"{ctx}: shape {got.shape} vs {want.shape}"
    d = np.abs(got - want)
    i = int(np.argmax(d)) if d.size else 0
    assert d.size == 0 or d[i] <= tol, f"{ctx}: elem {i}: got {got[i]} want {want[i]} (diff {d[i]})"


def rel_err(got, want, floor):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    return float(np.max(np.abs(got - want) / np.maximum(np.abs(want), floor)))


def assert_scaled_close(got, want, rtol, ctx=""):
    """max |got - want| <= rtol * max |want|.

    Sums of many clipped terms cancel, so an element-wise relative bar is
    meaningless near zero; the error of such a sum is bounded by the size of
    its terms, for which the largest magnitude in the array stands in.
    """
    got = np.asarray(got, np.float64).ravel()
    want = np.asarray(want, np.float64).ravel()
    assert got.shape == want.shape, f"{ctx}: shape {got.shape} vs {want.shape}"
    if want.size == 0:
        return
    d = np.abs(got - want)
    scale = max(float(np.max(np.abs(want))), 1e-30)
    i = int(np.argmax(d))
    assert np.all(np.isfinite(got)), f"{ctx}: non-finite output"
    assert d[i] <= rtol * scale, (
        f"{ctx}: max diff {d[i]:.3e} at {i} (got {got[i]!r}, want {want[i]!r}); scale {scale:.3e}, "
        f"scaled {d[i] / scale:.3e} > {rtol:.1e}; median diff {float(np.median(d)):.3e}")


def assert_grad_close(got, want, absum, rtol=1e-4, ctx=""):
    """Per coordinate |got - want| <= rtol * sum|terms| of that coordinate.

    A node's gradient is a sum of many (clipped) terms that can cancel; its
    rounding error, and the error any reordering of the sum can make, is
    bounded by the sum of the terms' magnitudes, which the oracle returns with
    the gradient.  A coordinate no term touches (sum 0) must be exactly 0."""
    got = np.asarray(got, np.float64).ravel()
    want = np.asarray(want, np.float64).ravel()
    absum = np.asarray(absum, np.float64).ravel()
    assert got.shape == want.shape == absum.shape, f"{ctx}: shapes"
    assert np.all(np.isfinite(got)), f"{ctx}: non-finite output"
    d = np.abs(got - want)
    bar = rtol * absum
    bad = np.flatnonzero(d > bar)
    worst = float(np.max(d / np.maximum(absum, 1e-30))) if d.size else 0.0
    assert bad.size == 0, (
        f"{ctx}: {bad.size} coords over {rtol:g} * sum|terms|, first {bad[0]}: got {got[bad[0]]!r} "
        f"want {want[bad[0]]!r} diff {d[bad[0]]:.3e} sum|terms| {absum[bad[0]]:.3e}")
    return worst


def diff_stats(got, want):
    d = np.abs(np.asarray(got, np.float64).ravel() - np.asarray(want, np.float64).ravel())
    return (f"max {d.max():.3e} p99 {np.percentile(d, 99):.3e} median {np.median(d):.3e} "
            f"frac>1e-4 {np.mean(d > 1e-4):.2e}")
