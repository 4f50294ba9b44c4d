"""Embedding-quality metrics used by the parity tests (textbook definitions).

trustworthiness: Venna & Kaski (2001) as in sklearn.manifold.trustworthiness;
spearman: rank correlation with average ranks for ties;
recall_at_k: |approx top-k  intersect  exact top-k| / k, averaged over rows.
"""
from __future__ import annotations

import numpy as np


def _ranks_avg(x: np.ndarray) -> np.ndarray:
    order = np.argsort(x, kind="mergesort")
    ranks = np.empty(len(x), np.float64)
    xs = x[order]
    i = 0
    while i < len(x):
        j = i
        while j + 1 < len(x) and xs[j + 1] == xs[i]:
            j += 1
        ranks[order[i:j + 1]] = (i + j) / 2.0 + 1.0
        i = j + 1
    return ranks


def spearman(a, b) -> float:
    ra, rb = _ranks_avg(np.asarray(a, np.float64)), _ranks_avg(np.asarray(b, np.float64))
    ra -= ra.mean()
    rb -= rb.mean()
    return float((ra * rb).sum() / np.sqrt((ra * ra).sum() * (rb * rb).sum()))


def trustworthiness(x, y, k: int = 15, block: int = 2048) -> float:
    """1 - 2/(n k (2n - 3k - 1)) * sum_i sum_{j in N_k^Y(i) \\ N_k^X(i)} (r_X(i,j) - k)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    n = x.shape[0]
    xn = (x * x).sum(1)
    yn = (y * y).sum(1)
    penalty = 0.0
    for s in range(0, n, block):
        e = min(n, s + block)
        dx = xn[s:e, None] + xn[None, :] - 2.0 * x[s:e] @ x.T
        dy = yn[s:e, None] + yn[None, :] - 2.0 * y[s:e] @ y.T
        rows = np.arange(e - s)
        dx[rows, np.arange(s, e)] = np.inf
        dy[rows, np.arange(s, e)] = np.inf
        # rank of every j in the original space (1 = nearest)
        order_x = np.argsort(dx, axis=1, kind="mergesort")
        rank_x = np.empty_like(order_x)
        rank_x[rows[:, None], order_x] = np.arange(1, n + 1)[None, :]
        nn_y = np.argsort(dy, axis=1, kind="mergesort")[:, :k]
        r = rank_x[rows[:, None], nn_y]
        penalty += np.clip(r - k, 0, None).sum()
    return float(1.0 - 2.0 / (n * k * (2.0 * n - 3.0 * k - 1.0)) * penalty)


def recall_at_k(approx_idx, exact_idx) -> float:
    a = np.asarray(approx_idx)
    b = np.asarray(exact_idx)
    k = b.shape[1]
    hits = sum(len(np.intersect1d(a[i], b[i])) for i in range(a.shape[0]))
    return hits / (a.shape[0] * k)
