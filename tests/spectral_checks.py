"""Checks of a spectral layout against dense eigensolvers (numpy), shared by the
oracle's CPU tests and the device's GPU tests.

spectral_layout (src/data/init.rs:986-1004) lays each connected component out
by the eigenvectors 1..n_dim of M = -D^-1/2 W D^-1/2 (smallest first), each
then placed, centred and scaled.  On components small enough that Lanczos
spans the whole space the Ritz vectors are exact, so each layout column
restricted to a component lies in the span of that component's eigenvectors
1..n_dim and the constant vector.
"""
import numpy as np


def rectangle(n, aspect, d, seed):
    """Uniform points in [0, aspect] x [0, 1] rotated into d dimensions."""
    rng = np.random.default_rng(seed)
    pts = rng.uniform(size=(n, 2)) * np.array([aspect, 1.0])
    q, _ = np.linalg.qr(rng.normal(size=(d, d)))
    return (pts @ q[:2]).astype(np.float32)


def dense_m(n, rows, cols, vals):
    """M = -D^-1/2 W D^-1/2 with a zero diagonal (graph_to_negative_normalised_adjacency,
    src/data/init.rs:210-256)."""
    W = np.zeros((n, n))
    W[rows, cols] = vals
    deg = W.sum(1)
    dinv = np.where(deg > 1e-8, 1.0 / np.sqrt(deg), 0.0)
    M = -(dinv[:, None] * W * dinv[None, :])
    np.fill_diagonal(M, 0.0)
    return M


def fit_span(basis, y):
    """(R^2 of y in span(basis), least-squares coefficients)."""
    coef, *_ = np.linalg.lstsq(basis, y, rcond=None)
    return 1.0 - (y - basis @ coef).var() / y.var(), coef


def check_components(emb, rows, cols, vals, parts, n_dim):
    """Each component (slices `parts`) spans its own eigenvectors 1..n_dim;
    returns the per-component intercepts (the placed centres) and the largest
    deviation of a component's points from its centre."""
    centre = np.zeros((len(parts), n_dim))
    reach = np.zeros(len(parts))
    for k, sl in enumerate(parts):
        keep = (rows >= sl.start) & (rows < sl.stop)
        assert np.all((cols[keep] >= sl.start) & (cols[keep] < sl.stop)), "components leak"
        m = sl.stop - sl.start
        w, U = np.linalg.eigh(dense_m(m, rows[keep] - sl.start, cols[keep] - sl.start, vals[keep]))
        assert w[n_dim + 1] - w[n_dim] > 1e-3, "test graph: eigenvalues not separated"
        basis = np.column_stack([U[:, 1:n_dim + 1], np.ones(m)])
        for dd in range(n_dim):
            r2, coef = fit_span(basis, emb[sl, dd])
            assert r2 > 0.9999, f"component {k} column {dd}: R^2 {r2:.6f}"
            centre[k, dd] = coef[-1]
        reach[k] = np.max(np.abs(emb[sl] - centre[k]))
    return centre, reach


def check_meta_spectral(centre, parts_data, n_dim):
    """The centres lie in the span of the spectral layout of the dense affinity
    exp(-|c_a - c_b|^2 / max |c - c'|^2) between the components' data centroids
    (component_spectral_meta, src/data/init.rs:648-720)."""
    cen = np.stack([p.mean(0) for p in parts_data]).astype(np.float64)
    d2 = ((cen[:, None] - cen[None]) ** 2).sum(-1)
    A = np.exp(-d2 / d2.max())
    np.fill_diagonal(A, 0.0)
    dinv = 1.0 / np.sqrt(A.sum(1))
    wm, V = np.linalg.eigh(-(dinv[:, None] * A * dinv[None]))
    assert wm[n_dim + 1] - wm[n_dim] > 1e-3, "test meta graph: eigenvalues not separated"
    basis = np.column_stack([V[:, 1:n_dim + 1], np.ones(len(parts_data))])
    for dd in range(n_dim):
        r2, _ = fit_span(basis, centre[:, dd])
        assert r2 > 0.999, f"meta column {dd}: R^2 {r2:.5f}"
