"""Pin the CPU oracle against the reference's own golden vectors and against
independent mathematics (finite differences, dense restatements, brute force).

CPU only.  Each test names the reference test or the property it follows.
"""
import math

import numpy as np
import pytest
from scipy.spatial.distance import cdist

import pyoracle as orc
from helpers import assert_close, blobs, coo_to_adjacency, triangle_setup

import mrs_amd


def tri_csr():
    # triangle_graph, src/training/umap_optimiser_gpu.rs:1646-1652
    g = [[(1, 1.0), (2, 0.5)], [(0, 1.0), (2, 1.0)], [(0, 0.5), (1, 1.0)]]
    st, c = orc.csr_from_graph(*mrs_amd.adjacency_to_csr(g))
    assert st == 0
    return c


# --------------------------------------------------------------- hash


def test_hash_finaliser_is_lowbias32():
    # the finaliser of gpu_hash_neg (umap_optimiser_gpu.rs:619-623) is C. Wellons'
    # "lowbias32"; its published inverse must undo it.  With node=epoch=...=0 the
    # pre-mix is the seed itself, so the oracle computes lowbias32(seed) % n with
    # n = 2^32 - 1 > h for h < n; check the inverse on those values.
    def inv(x):
        x &= 0xFFFFFFFF
        x ^= x >> 16
        x = (x * 0x43021123) & 0xFFFFFFFF
        x ^= (x >> 15) ^ (x >> 30)
        x = (x * 0x1D69E2A5) & 0xFFFFFFFF
        x ^= x >> 16
        return x

    rng = np.random.default_rng(0)
    n = 0xFFFFFFFF
    for s in rng.integers(0, 2**32 - 1, size=2000, dtype=np.uint64):
        h = orc.hash_neg(int(s), 0, 0, 0, 0, n)
        if h != n:  # h == n would have been reduced to 0
            assert inv(h) == int(s)


def test_hash_premix_uses_xxh32_primes():
    # The per-field multipliers of the pre-mix are the xxHash32 primes
    # PRIME32_1..4 (0x9E3779B1, 0x85EBCA77, 0xC2B2AE3D, 0x27D4EB2F).  Each
    # field enters as field*prime xored into the seed, so flipping a field's
    # value x -> x xor 1 at field value 0 equals seed -> seed xor prime.
    primes = [0x9E3779B1, 0x85EBCA77, 0xC2B2AE3D, 0x27D4EB2F]
    n = 0xFFFFFFFF
    for pos, prime in enumerate(primes):
        args = [0, 0, 0, 0]
        args[pos] = 1
        a = orc.hash_neg(0, *args, n)
        b = orc.hash_neg(prime, 0, 0, 0, 0, n)
        assert a == b


def test_hash_range_and_modulo():
    rng = np.random.default_rng(1)
    for n in [1, 2, 3, 7, 1000, 99_991, 1_000_000, 2**31 + 11]:
        for _ in range(50):
            args = [int(v) for v in rng.integers(0, 2**32 - 1, size=5, dtype=np.uint64)]
            h = orc.hash_neg(*args, n)
            assert 0 <= h < n
            raw = orc.hash_neg(*args, 0xFFFFFFFF)  # == the unreduced hash unless it is 2^32-1
            assert h == raw % n


# ---------------------------------------------------------------- CSR


def test_csr_triangle_layout():
    # test_csr_triangle_layout, umap_optimiser_gpu.rs:1716-1728
    c = tri_csr()
    assert c["n_edges"] == 3
    assert list(c["off"]) == [0, 2, 4, 6]
    assert list(c["edge_idx"]) == [0, 1, 0, 2, 1, 2]
    assert list(c["other"]) == [1, 2, 0, 2, 0, 1]


def test_csr_triangle_sampling_periods():
    # :1730-1737
    assert list(tri_csr()["per"]) == [1.0, 2.0, 1.0]


def test_csr_offsets_sum_to_2x_edges():
    c = tri_csr()
    assert int(c["off"][-1]) == 2 * c["n_edges"]


def test_csr_errors_and_self_loops():
    # :1751-1785
    st, _ = orc.csr_from_graph(np.zeros(1, np.int64), np.zeros(0, np.int32), np.zeros(0, np.float32))
    assert st == 1  # NoData
    st, _ = orc.csr_from_graph(*mrs_amd.adjacency_to_csr([[], [], []]))
    assert st == 2  # NoGraphEdges
    st, c = orc.csr_from_graph(*mrs_amd.adjacency_to_csr([[(0, 1.0), (1, 1.0)], [(1, 1.0), (0, 1.0)]]))
    assert st == 0 and c["n_edges"] == 1 and list(c["other"]) == [1, 0]
    st, c = orc.csr_from_graph(*mrs_amd.adjacency_to_csr([[(1, 1.0), (2, 0.5)], [], []]))
    assert c["n_edges"] == 2 and list(c["off"]) == [0, 2, 3, 4]


# ----------------------------------------------------------- gradient


def test_grad_empty_epoch_is_zero():
    # test_grad_empty_epoch, :1995-2003
    embd, graph, params = triangle_setup()
    c = tri_csr()
    g, hu = orc.grad_accum(embd, c, c["per"], 0, 42, 5, params)
    assert np.all(g == 0) and np.all(hu == 0)


def test_grad_single_edge_hand_computed():
    # test_grad_single_edge_hand_computed, :2060-2092 -> (1.2, 0, -1.2, 0)
    embd = np.array([[0.0, 0.0], [3.0, 0.0]], np.float32)
    st, c = orc.csr_from_graph(*mrs_amd.adjacency_to_csr([[(1, 1.0)], [(0, 1.0)]]))
    p = mrs_amd.UmapOptimParams(a=1.0, b=1.0, lr=1.0, gamma=1.0, n_epochs=10, neg_sample_rate=0,
                                min_dist=0.1, beta1=0.9, beta2=0.999, eps=1e-7)
    g, _ = orc.grad_accum(embd, c, c["per"], 5, 42, 0, p)
    assert_close(g, [1.2, 0.0, -1.2, 0.0], 1e-5, "hand-computed")


def _fd_grad(f, y, h=1e-6):
    out = np.zeros_like(y)
    for d in range(y.size):
        e = np.zeros_like(y)
        e[d] = h
        out[d] = (f(y + e) - f(y - e)) / (2 * h)
    return out


@pytest.mark.parametrize("a,b", [(1.5, 0.9), (1.0, 1.0), (0.58, 1.9)])
def test_grad_attractive_is_minus_two_grad_log_kernel(a, b):
    # attractive term = -2 * d/dy_i log(1 + a |y_i - y_j|^{2b}) (UMAP cross-entropy)
    rng = np.random.default_rng(3)
    embd = rng.normal(size=(2, 3)).astype(np.float32)
    st, c = orc.csr_from_graph(*mrs_amd.adjacency_to_csr([[(1, 1.0)], [(0, 1.0)]]))
    p = mrs_amd.UmapOptimParams(a=a, b=b, lr=1.0, gamma=1.0, n_epochs=10, neg_sample_rate=0,
                                min_dist=0.1, beta1=0.9, beta2=0.999, eps=1e-7)
    g, _ = orc.grad_accum(embd, c, c["per"], 5, 42, 0, p)
    yj = embd[1].astype(np.float64)
    f = lambda y: math.log1p(a * np.sum((y - yj) ** 2) ** b)
    want = -2.0 * _fd_grad(f, embd[0].astype(np.float64))
    np.testing.assert_allclose(g[0], want, rtol=1e-4, atol=1e-6)


def test_grad_repulsive_is_grad_log_one_minus_phi():
    # repulsive term of one negative = gamma * d/dy_i log(1 - phi(s + eps)),
    # phi(s) = 1 / (1 + a s^b), away from the +-4 clip.  One active edge whose
    # endpoint sits far away (attractive part removed by subtracting the
    # neg_rate=0 run), negatives counted through the hash.
    a, b, gamma = 1.5, 0.9, 1.0
    rng = np.random.default_rng(5)
    n = 40
    embd = (rng.normal(size=(n, 2)) * 3).astype(np.float32)
    graph = [[] for _ in range(n)]
    graph[0].append((1, 1.0))
    graph[1].append((0, 1.0))
    st, c = orc.csr_from_graph(*mrs_amd.adjacency_to_csr(graph))
    p = mrs_amd.UmapOptimParams(a=a, b=b, lr=1.0, gamma=gamma, n_epochs=10, neg_sample_rate=3,
                                min_dist=0.1, beta1=0.9, beta2=0.999, eps=1e-7)
    g3, _ = orc.grad_accum(embd, c, c["per"], 5, 7, 3, p)
    g0, _ = orc.grad_accum(embd, c, c["per"], 5, 7, 0, p)
    rep = (g3 - g0)[0].astype(np.float64)
    want = np.zeros(2)
    y0 = embd[0].astype(np.float64)
    for s in range(3):
        k = orc.hash_neg(7, 0, 5, 0, s, n)
        if k == 0:
            continue
        yk = embd[k].astype(np.float64)

        def f(y):
            ss = np.sum((y - yk) ** 2) + 0.001
            return gamma * math.log(1.0 - 1.0 / (1.0 + a * ss ** b))

        gk = _fd_grad(f, y0)
        coeff = 2 * gamma * b / ((np.sum((y0 - yk) ** 2) + 0.001) * (1 + a * (np.sum((y0 - yk) ** 2) + 0.001) ** b))
        assert coeff < 4.0  # away from the clip
        want += gk
    np.testing.assert_allclose(rep, want, rtol=2e-4, atol=1e-6)


# --------------------------------------------------------------- Adam


def test_adam_single_step_hand_computed():
    # test_adam_single_step_hand_computed, :2211-2242
    p = mrs_amd.UmapOptimParams(a=1.5, b=0.9, lr=1.0, gamma=1.0, n_epochs=10, neg_sample_rate=5,
                                min_dist=0.1, beta1=0.9, beta2=0.999, eps=1e-7)
    e, m, v = orc.adam_update([1.0], [1], [0.0], [0.0], [0.0], p, 0, 10)
    assert_close(m, [0.1], 1e-6, "m")
    assert_close(v, [0.001], 1e-6, "v")
    assert_close(e, [1.0], 1e-4, "embd")


def test_adam_inactive_node_preserved():
    # :2192-2209
    p = mrs_amd.UmapOptimParams.default_2d()
    embd = np.array([1.5, -2.5, 7.0, 8.0], np.float32)
    m = np.array([0.1, 0.2, 0.3, 0.4], np.float32)
    v = np.array([0.01, 0.02, 0.03, 0.04], np.float32)
    e2, m2, v2 = orc.adam_update(np.full(4, 5.0), [0, 0], m, v, embd, p, 3, 10)
    assert np.array_equal(e2, embd) and np.array_equal(m2, m) and np.array_equal(v2, v)


def test_adam_matches_textbook_bias_corrected_adam():
    # fused form == lr_t * mhat / (sqrt(vhat) + eps) of Kingma & Ba (float64)
    rng = np.random.default_rng(11)
    p = mrs_amd.UmapOptimParams(a=1.5, b=0.9, lr=1.0, gamma=1.0, n_epochs=50, neg_sample_rate=5,
                                min_dist=0.1, beta1=0.5, beta2=0.9, eps=1e-7)
    n = 64
    y = rng.normal(size=n).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    y64, m64, v64 = y.astype(np.float64), np.zeros(n), np.zeros(n)
    for t in range(10):
        g = rng.normal(size=n).astype(np.float32)
        y, m, v = orc.adam_update(g, np.ones(n, np.uint32), m, v, y, p, t, 50)
        g64 = g.astype(np.float64)
        m64 = 0.5 * m64 + 0.5 * g64
        v64 = 0.9 * v64 + 0.1 * g64 * g64
        mhat = m64 / (1 - 0.5 ** (t + 1))
        vhat = v64 / (1 - 0.9 ** (t + 1))
        y64 = y64 + (1.0 - t / 50) * mhat / (np.sqrt(vhat) + 1e-7)
    np.testing.assert_allclose(y, y64, rtol=0, atol=2e-5)


# ----------------------------------------------------------- schedule


def test_schedule_goldens():
    # :2342-2379
    per = np.array([1.0, 2.0, 1.0], np.float32)
    assert list(orc.schedule_update(per, per, 5)) == [2.0, 4.0, 2.0]
    assert list(orc.schedule_update(per, per, 0)) == [1.0, 2.0, 1.0]
    assert list(orc.schedule_update(per, per, 1)) == [2.0, 2.0, 2.0]


# ------------------------------------------------------------ densMAP


def test_correlation_stats_known_values():
    # src/utils/density.rs:651-676
    r = np.array([-1.0, 0.0, 1.0])
    mean, sd, cov = orc.correlation_stats(3 * r + 5, r, 0.0)
    assert abs(mean - 5) < 1e-6 and abs(sd - 3) < 1e-6 and abs(cov - 3) < 1e-6
    _, sd_plain, _ = orc.correlation_stats(r, r, 0.0)
    _, sd_shift, _ = orc.correlation_stats(r, r, 0.1)
    assert abs(sd_plain - 1.0) < 1e-7 and abs(sd_shift - math.sqrt(1.1)) < 1e-7


def test_correlation_stats_matches_numpy_many_chunks():
    rng = np.random.default_rng(2)
    re = rng.normal(size=20000).astype(np.float32)
    r = rng.normal(size=20000).astype(np.float32)
    mean, sd, cov = orc.correlation_stats(re, r, 0.1)
    re64, r64 = re.astype(np.float64), r.astype(np.float64)
    assert abs(mean - re64.mean()) < 1e-12
    assert abs(sd - math.sqrt(re64.var(ddof=1) + float(np.float32(0.1)))) < 1e-12
    assert abs(cov - np.sum((re64 - re64.mean()) * r64) / (re64.size - 1)) < 1e-12


def test_dens_state_known_value():
    # test_original_log_radii_known_value, src/utils/density.rs:592-622
    rows, cols = np.array([0, 0, 1, 2]), np.array([1, 2, 0, 0])
    vals = np.array([1.0, 3.0, 1.0, 3.0])
    knn_idx = np.array([[1, 2], [0, 2], [0, 1]])
    knn_dist = np.array([[4.0, 8.0], [4.0, 99.0], [8.0, 99.0]])
    st, r, ms, tot = orc.dens_state(3, rows, cols, vals, knn_idx, knn_dist, True)
    assert st == 0
    assert list(ms) == [4.0, 1.0, 3.0] and tot == 8.0
    raw = np.log(1e-8 + np.array([7.0, 4.0, 8.0]))
    want = (raw - raw.mean()) / raw.std(ddof=1)
    assert_close(r, want, 1e-6, "r")


def test_dens_state_degenerate_and_isolated():
    # :624-649 and the reverse-row fallback :578-590
    st, *_ = orc.dens_state(2, [0, 1], [1, 0], [1.0, 1.0], [[1], [0]], [[4.0], [4.0]])
    assert st == 3
    st, r, ms, _ = orc.dens_state(3, [0, 1], [1, 0], [1.0, 2.0], [[1], [0], [0]], [[4.0], [9.0], [1.0]])
    assert st == 0 and ms[2] == 0 and np.all(np.isfinite(r))
    # edge (0,1) where 1 is not in 0's row: falls back to row 1's distance 4
    st, r, ms, _ = orc.dens_state(3, [0, 1, 2], [1, 0, 0], [1.0, 1.0, 1.0],
                                  [[7], [0], [0]], [[1.0], [4.0], [9.0]])
    raw = np.log(1e-8 + np.array([4.0, 4.0, 9.0]))
    assert_close(r, (raw - raw.mean()) / raw.std(ddof=1), 1e-6, "fallback")


def test_dens_is_active():
    # :691-718
    assert not orc.dens_is_active(2.0, 0.3, 69, 100)
    assert orc.dens_is_active(2.0, 0.3, 70, 100)
    assert orc.dens_is_active(2.0, 0.3, 99, 100)
    assert not orc.dens_is_active(0.0, 1.0, 99, 100)


def test_dens_radii_closed_form():
    # one neighbour: re = log(1e-8 + d^2), phi_sum = 1/(1 + a d^{2b})
    embd = np.array([[0.0, 0.0], [3.0, 4.0]], np.float32)
    st, c = orc.csr_from_graph(*mrs_amd.adjacency_to_csr([[(1, 1.0)], [(0, 1.0)]]))
    re, phi = orc.dens_radii(embd, c, 1.5, 0.9)
    assert_close(re, [math.log(1e-8 + 25.0)] * 2, 1e-5, "re")
    assert_close(phi, [1 / (1 + 1.5 * 25.0 ** 0.9)] * 2, 1e-7, "phi")


# -------------------------------------------------------------- graph


def test_smooth_knn_dist_basic():
    # src/data/graph.rs:1441-1471
    sigma, rho = orc.smooth_knn_dist([[1.0, 2.0], [1.5, 3.0], [0.5, 1.5]], 2)
    assert_close(rho, [1.0, 1.5, 0.5], 1e-4, "rho")
    assert np.all(sigma > 0)
    sigma, rho = orc.smooth_knn_dist([[0.0, 1.0, 2.0], [0.0, 1.0, 2.0]], 3, local_connectivity=0.0)
    assert np.all(rho == 0.0)


def test_smooth_knn_dist_invariant():
    # |sum_j exp(-max(0, d_j - rho)/sigma) - ln k| < bandwidth unless the 64
    # iterations ran out; rho = d[0] (tests/umap_tests.rs:152-164)
    rng = np.random.default_rng(4)
    x, _ = blobs(600, 8, 4, seed=4)
    idx, dist = orc.knn_exact(x, 15)
    sigma, rho = orc.smooth_knn_dist(dist, 15)
    assert np.array_equal(rho, dist[:, 0])
    adj = np.maximum(dist.astype(np.float64) - rho[:, None], 0.0)
    val = np.exp(-adj / sigma[:, None].astype(np.float64)).sum(axis=1)
    ok = np.abs(val - math.log(15)) < 2e-5
    assert ok.mean() > 0.99


def test_knn_to_coo_closed_form_and_self():
    # :1473-1510
    idx = np.array([[1, 2], [0, 2], [0, 1]])
    dist = np.array([[1.0, 2.0], [1.0, 1.5], [2.0, 1.5]])
    r, c, v = orc.knn_to_coo(idx, dist, [1.0, 1.0, 1.0], [0.0, 0.0, 0.0])
    assert len(v) == 6 and np.all((v >= 0) & (v <= 1))
    assert_close(v, np.exp(-dist.ravel()), 1e-7, "weights")
    r, c, v = orc.knn_to_coo([[0, 1], [1, 0]], [[0.0, 1.0], [0.0, 1.0]], [1.0, 1.0], [0.0, 0.0])
    assert len(v) == 2 and np.all(r != c)


def test_symmetrise_goldens():
    # test_symmetrise_graph_full_union / _directed, src/data/graph.rs:1512-1563
    r, c, v = orc.symmetrise(2, [0, 1], [1, 0], [0.8, 0.6], 0.5)
    d = {(int(a), int(b)): float(w) for a, b, w in zip(r, c, v)}
    assert abs(d[(0, 1)] - 0.86) < 1e-6 and abs(d[(1, 0)] - 0.76) < 1e-6
    r, c, v = orc.symmetrise(2, [0, 1], [1, 0], [0.8, 0.6], 1.0)
    assert np.all(np.abs(v - 0.92) < 1e-6) and len(v) == 2


@pytest.mark.parametrize("mix", [1.0, 0.5, 0.0])
def test_symmetrise_matches_dense_restatement(mix):
    rng = np.random.default_rng(8)
    n, k = 30, 5
    rows = np.repeat(np.arange(n), k)
    cols = np.array([rng.choice(np.setdiff1d(np.arange(n), [i]), k, replace=False) for i in range(n)]).ravel()
    vals = rng.uniform(0.01, 1.0, size=n * k).astype(np.float32)
    W = np.zeros((n, n), np.float64)
    W[rows, cols] = vals
    S = mix * (W + W.T - W * W.T) + (1 - mix) * W
    r, c, v = orc.symmetrise(n, rows, cols, vals, mix)
    nz = np.argwhere(S > 0)  # row-major, columns ascending
    assert np.array_equal(np.stack([r, c], 1), nz)
    np.testing.assert_allclose(v, S[r, c], rtol=1e-6)


def test_filter_weak_edges_threshold():
    rng = np.random.default_rng(9)
    v = rng.uniform(0, 1, size=1000).astype(np.float32)
    v[::7] = 1e-4
    r = np.arange(1000, dtype=np.int32)
    c = (r + 1) % 1000
    fr, fc, fv = orc.filter_weak_edges(r, c, v, 500)
    keep = v >= v.max() / np.float32(500)
    assert np.array_equal(fr, r[keep]) and np.array_equal(fv, v[keep])


def test_knn_exact_matches_brute_force():
    x, _ = blobs(300, 10, 3, seed=12)
    x[5] = x[6]  # an exact duplicate: tie at distance 0 broken by index
    idx, dist = orc.knn_exact(x, 7)
    D = cdist(x.astype(np.float64), x.astype(np.float64), "sqeuclidean")
    np.fill_diagonal(D, np.inf)
    for q in range(300):
        order = np.lexsort((np.arange(300), D[q]))[:7]
        assert np.array_equal(idx[q], order)
        np.testing.assert_allclose(dist[q], D[q, order], rtol=1e-5, atol=1e-5)
    idx64, d64 = orc.knn_exact(x, 7, f64=True)
    assert np.array_equal(idx64, idx)


def test_fit_params_invariants():
    # src/training/umap_optimisers.rs:1559-1593
    a, b = orc.fit_params(0.1, 1.0)
    assert 0.001 <= a <= 10 and 0.1 <= b <= 2
    f = lambda x: 1.0 / (1.0 + a * x ** (2 * b))
    assert f(0.1) > 0.9 and f(3.0) < 0.1 and f(0.1) > f(1.5) > f(3.0)


def test_cpu_optimisers_separate_two_cliques():
    # test_gpu_driver_preserves_graph_structure twin, :2386-2448, for the CPU paths
    graph = [[(1, 1.0), (2, 1.0)], [(0, 1.0), (2, 1.0)], [(0, 1.0), (1, 1.0), (3, 0.1)],
             [(2, 0.1), (4, 1.0), (5, 1.0)], [(3, 1.0), (5, 1.0)], [(3, 1.0), (4, 1.0)]]
    embd = np.array([[0, 0], [10, 0], [0, 10], [10, 10], [-5, -5], [15, 15]], np.float32)
    p = mrs_amd.UmapOptimParams.default_2d()
    p.n_epochs = 200
    adj = mrs_amd.adjacency_to_csr(graph)
    for out in (orc.optimise_adam_gpu(embd, *adj, p, 42),
                orc.optimise_adam_parallel(embd, *adj, p, 42)):
        dist = lambda i, j: float(np.linalg.norm(out[i] - out[j]))
        intra = (dist(0, 1) + dist(0, 2) + dist(1, 2) + dist(3, 4) + dist(3, 5) + dist(4, 5)) / 6
        inter = np.mean([dist(i, j) for i in (0, 1) for j in (3, 4, 5)])
        assert inter > 1.5 * intra and np.all(np.isfinite(out))


def test_cpu_densmap_raises_radius_correlation():
    # densMAP's purpose (Narayan et al. 2021): embedding local radii track the
    # original ones better than plain UMAP's (tests/densmap_tests.rs analogue)
    from helpers import umap_graph_from_data
    from metrics import spearman

    rng = np.random.default_rng(3)
    sds = [0.3, 1.0, 3.0]
    x = np.concatenate([rng.normal(scale=s, size=(150, 6)) + 15 * i for i, s in enumerate(sds)])
    x = x.astype(np.float32)
    idx, dist, (r, c, v), adj = umap_graph_from_data(x, 15, n_epochs=200)
    st, R, ms, mu_tot = orc.dens_state(x.shape[0], r, c, v, idx, dist, True)
    assert st == 0
    init = rng.uniform(-10, 10, size=(x.shape[0], 2)).astype(np.float32)
    p = mrs_amd.UmapOptimParams.default_2d()
    p.n_epochs = 200
    plain = orc.optimise_adam_parallel(init, *adj, p, 42)
    dens = orc.optimise_adam_parallel(init, *adj, p, 42, dens_r=R, dens_params=(2.0, 0.3, 0.1),
                                      mu_tot=mu_tot)

    def emb_radii(e):
        out = np.zeros(x.shape[0])
        off = adj[0]
        for i in range(x.shape[0]):
            nb = adj[1][off[i]:off[i + 1]]
            d2 = ((e[nb] - e[i]) ** 2).sum(1)
            phi = 1 / (1 + p.a * d2 ** p.b)
            out[i] = np.log(1e-8 + (phi * d2).sum() / phi.sum())
        return out

    assert np.all(np.isfinite(dens))
    assert spearman(R, emb_radii(dens)) > spearman(R, emb_radii(plain)) + 0.05


# ------------------------------------------------ fit_params, LUT, floors


def _fit_loss_f64(min_dist, spread, a, b):
    """L(a, b) = 1/(2 * 300) * sum_{x > 0} (1 / (1 + a x^(2b)) - y(x))^2 on the 300
    points of fit_params (src/training/umap_optimisers.rs:145-164), in float64."""
    x = np.arange(300) * (3.0 * spread / 299.0)
    y = np.where(x < min_dist, 1.0, np.exp(-(x - min_dist) / spread))
    x, y = x[1:], y[1:]
    return float(((1.0 / (1.0 + a * x ** (2.0 * b)) - y) ** 2).sum() / 600.0)


@pytest.mark.parametrize("md,sp,a,b", [(0.1, 1.0, 1.0, 1.0), (0.1, 1.0, 0.5, 1.5),
                                       (0.5, 2.0, 3.0, 0.7), (0.0, 1.0, 1.6, 0.9)])
def test_fit_params_gradient_is_gradient_of_its_loss(md, sp, a, b):
    # fit_params steps (a, b) -= (grad_a, grad_b): those must be the gradient of
    # the mean squared curve error (with the 1/2), checked by central
    # differences in float64.  A dropped factor 2 or ln x, or a wrong sign,
    # fails here.
    ga, gb = orc.fit_params_grad(md, sp, a, b)
    h = 1e-6
    fa = (_fit_loss_f64(md, sp, a + h, b) - _fit_loss_f64(md, sp, a - h, b)) / (2 * h)
    fb = (_fit_loss_f64(md, sp, a, b + h) - _fit_loss_f64(md, sp, a, b - h)) / (2 * h)
    assert abs(ga - fa) <= 1e-5 + 1e-4 * abs(fa), (ga, fa)
    assert abs(gb - fb) <= 1e-5 + 1e-4 * abs(fb), (gb, fb)


def test_fit_params_is_projected_gradient_descent():
    # n_iter steps of (a, b) <- clamp((a, b) - grad(a, b)) from (1, 1): one step
    # from (1, 1), and the composition of steps, each through the gradient above
    a1, b1 = orc.fit_params(0.1, 1.0, n_iter=1)
    ga, gb = orc.fit_params_grad(0.1, 1.0, 1.0, 1.0)
    assert (a1, b1) == (np.float32(1.0 - ga), np.float32(1.0 - gb))
    a, b = 1.0, 1.0
    for _ in range(5):
        ga, gb = orc.fit_params_grad(0.8, 0.5, a, b)
        a = float(np.clip(np.float32(a - ga), np.float32(0.001), np.float32(10.0)))
        b = float(np.clip(np.float32(b - gb), np.float32(0.1), np.float32(2.0)))
    assert orc.fit_params(0.8, 0.5, n_iter=5) == (np.float32(a), np.float32(b))
    # spread 0.5 / min_dist 0.8 drives b into its upper clamp
    assert orc.fit_params(0.8, 0.5)[1] == 2.0


def test_fit_params_library_equals_oracle():
    # the product's from_min_dist_spread computes the same a, b, bit for bit
    for md, sp in [(0.1, 1.0), (0.5, 1.0), (0.0, 1.0), (0.25, 2.0), (0.8, 0.5), (1.5, 1.0)]:
        p = mrs_amd.UmapOptimParams.from_min_dist_spread(md, sp)
        a, b = orc.fit_params(md, sp)
        assert np.float32(p.a).tobytes() == np.float32(a).tobytes(), (md, sp)
        assert np.float32(p.b).tobytes() == np.float32(b).tobytes(), (md, sp)
        assert (p.lr, p.gamma, p.n_epochs, p.neg_sample_rate) == (1.0, 1.0, 500, 5)
        assert (np.float32(p.beta1), np.float32(p.beta2), np.float32(p.eps)) == (
            np.float32(0.5), np.float32(0.9), np.float32(1e-7))
    q = mrs_amd.UmapOptimParams.from_min_dist_spread(0.1, 1.0, lr=0.5, n_epochs=200,
                                                     neg_sample_rate=3, beta1=0.9)
    assert (q.lr, q.n_epochs, q.neg_sample_rate, np.float32(q.beta1)) == (0.5, 200, 3,
                                                                          np.float32(0.9))
    with pytest.raises(mrs_amd.ManifoldsError):
        mrs_amd.UmapOptimParams.from_min_dist_spread(0.1, 0.0)
    # new_default_2d: min_dist None -> 0.5, spread None -> 1.0
    g = mrs_amd.UmapParamsGpu.new_default_2d()
    a, b = orc.fit_params(0.5, 1.0)
    assert (g.optim_params.a, g.optim_params.b) == (a, b) and g.n_dim == 2 and g.k == 15


@pytest.mark.parametrize("b", [0.9, 0.5, 1.5, 1.8, 0.2])
def test_lut_pow_within_interpolation_bound(b):
    # FastPowLut (src/training/umap_optimisers.rs:325-397): x^b on 65 536 points
    # of [0, 25], linear interpolation between them, exact powf from 25 up.
    # On a cell [x0, x0 + h] the interpolation error of a smooth f is at most
    # h^2/8 max|f''|; f'' = b(b-1) x^(b-2) is largest at the cell's end nearer
    # zero for b < 2.  The first cell [0, h] is bounded by h^b max_t |t^b - t|.
    # An index off by one costs ~h f'(x), far above these bounds.
    rng = np.random.default_rng(int(b * 100))
    h = np.float64(np.float32(25.0) / np.float32(65535))
    x = np.concatenate([rng.uniform(0, 25, 20000), rng.uniform(0, 3 * h, 2000),
                        rng.uniform(25, 40, 500), [0.0, h, 25.0, 24.9999]]).astype(np.float32)
    got = orc.lut_pow(b, x).astype(np.float64)
    xd = x.astype(np.float64)
    want = xd ** b
    x0 = np.floor(xd / h) * h
    f2 = abs(b * (b - 1)) * np.maximum(np.where(x0 > 0, x0, h) ** (b - 2), (x0 + h) ** (b - 2))
    t = np.linspace(0, 1, 10001)
    first = h ** b * np.max(np.abs(t ** b - t))
    bound = np.where(xd < h, first, h * h / 8 * f2) + 8 * np.finfo(np.float32).eps * want + 1e-30
    above = xd >= 25
    bound[above] = 2 * np.finfo(np.float32).eps * want[above]
    bad = np.flatnonzero(np.abs(got - want) > bound)
    assert bad.size == 0, (x[bad[:3]], got[bad[:3]], want[bad[:3]], bound[bad[:3]])


def test_density_edge_coeff_values_and_floors():
    # density_edge_coeff (src/training/umap_optimisers.rs:564-592) at a = b = 1,
    # d^2 = 1: phi = 1/2, dphi_term = a b d^(2b)/d^2 phi = 1/2, common =
    # 1 - b (1 - phi) = 1/2, so with unit phi sums and re = 0 both dr = 1/2 and
    # coeff (w_i dr + w_o dr) / mu = 3 (0.5 + 1.0) / 0.5 = 9.
    c = orc.density_edge_coeff(1.0, 1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 0.0, 0.0, 1.0, 2.0, 3.0, 0.5)
    assert abs(c - 9.0) < 1e-6
    # re enters as exp(-re): re_i = ln 2 halves common / exp(re_i)
    c = orc.density_edge_coeff(1.0, 1.0, 1.0, 1.0, 2.0, 1.0, 1.0, math.log(2.0), 0.0, 1.0, 0.0,
                               1.0, 1.0)
    assert abs(c - 0.5 * (0.25 + 0.5)) < 1e-6
    # the max(eps) floors: a zero phi sum and a zero edge weight are floored at
    # f32 epsilon (2^-23) instead of dividing by zero
    eps = 2.0 ** -23
    c = orc.density_edge_coeff(1.0, 1.0, 1.0, 1.0, 2.0, 0.0, 1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 1.0)
    assert math.isfinite(c) and abs(c - 0.5 / eps) <= 1e-5 * (0.5 / eps)
    c = orc.density_edge_coeff(1.0, 1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0)
    assert math.isfinite(c) and abs(c - 0.5 / eps) <= 1e-5 * (0.5 / eps)


# ------------------------------------- CPU optimisers on two-node graphs
#
# Two nodes joined by one edge of weight 1: epochs_per_sample = 1, so the edge
# fires at epochs 1, 2, 3, ... (never at 0).  Negatives are scheduled at a
# period of 1/rate (0.2 for rate 5) from next_neg = 0.2: the first firing
# draws floor((1 - 0.2) / 0.2) = 4 = rate - 1 negatives, every later one
# floor(1 / 0.2) = 5 (both exact in f32).  With n = 2 a draw is the node
# itself (skipped) or the other node, so the repulsive sum is a count times one
# term.  The draws come from orc_hash_neg (pinned above) with the keys
# (node, epoch, CSR slot 0, s).


def _attr(a, b, yi, yo):
    """2 * coeff * (y_o - y_i), coeff = 2ab d^(2b) / (d^2 (1 + a d^(2b))) (unclipped)."""
    d2 = float(((yi - yo) ** 2).sum())
    db = d2 ** b
    return 2.0 * (2.0 * a * b * db / (d2 * (1.0 + a * db))) * (yo - yi)


def _rep(a, b, gamma, yi, yk):
    """clip(2 gamma b / ((d^2 + 0.001)(1 + a (d^2 + 0.001)^b)), 4) * (y_i - y_k)."""
    s = float(((yi - yk) ** 2).sum()) + 0.001
    return min(2.0 * gamma * b / (s * (1.0 + a * s ** b)), 4.0) * (yi - yk)


def _draws_other(seed, node, epoch, count):
    return sum(orc.hash_neg(seed, node, epoch, 0, s, 2) != node for s in range(count))


def _two_node_params(n_epochs, a=1.0, b=1.0, eps=10.0):
    # eps large: Adam's first steps stay proportional to the gradient
    # (m / (sqrt(v) + eps_c) with eps_c >> sqrt(v)), so the gradient's size,
    # not only its sign, shows in the result
    return mrs_amd.UmapOptimParams(a=a, b=b, lr=1.0, gamma=1.0, n_epochs=n_epochs,
                                   neg_sample_rate=5, min_dist=0.1, beta1=0.5, beta2=0.9,
                                   eps=eps)


@pytest.mark.parametrize("n_epochs", [2, 3, 4])
def test_adam_parallel_two_nodes_hand_computed(n_epochs):
    # optimise_embedding_adam_parallel (umap_optimisers.rs:1170-1492): node 0
    # (the smaller endpoint) takes the edge's negatives, node 1 none; Adam's
    # bias correction counts epochs (t = epoch + 1), not updates; lr decays as
    # lr (1 - epoch / n_epochs)
    p = _two_node_params(n_epochs)
    y0 = np.array([[0.0, 0.0], [3.0, 1.0]])
    adj = mrs_amd.adjacency_to_csr([[(1, 1.0)], [(0, 1.0)]])
    seed = 7
    y = y0.copy()
    m = np.zeros_like(y)
    v = np.zeros_like(y)
    for e in range(1, n_epochs):
        n_neg = 4 if e == 1 else 5
        g0 = _attr(p.a, p.b, y[0], y[1]) + _draws_other(seed, 0, e, n_neg) * _rep(
            p.a, p.b, p.gamma, y[0], y[1])
        g1 = _attr(p.a, p.b, y[1], y[0])
        g = np.stack([g0, g1])
        m += (1 - p.beta1) * (g - m)
        v += (1 - p.beta2) * (g * g - v)
        t = e + 1
        sq = math.sqrt(1 - p.beta2 ** t)
        y = y + p.lr * (1 - e / n_epochs) * sq / (1 - p.beta1 ** t) * m / (np.sqrt(v) + sq * p.eps)
    got = orc.optimise_adam_parallel(y0.astype(np.float32), *adj, p, seed, use_lut=False)
    np.testing.assert_allclose(got, y, rtol=2e-5, atol=2e-6)
    # with the LUT (b != 1 takes it): the same within the LUT's error
    q = _two_node_params(n_epochs, a=1.5, b=0.9)
    yl = orc.optimise_adam_parallel(y0.astype(np.float32), *adj, q, seed, use_lut=True)
    ye = orc.optimise_adam_parallel(y0.astype(np.float32), *adj, q, seed, use_lut=False)
    np.testing.assert_allclose(yl, ye, rtol=0, atol=1e-4)
    # and the draws matter: another seed draws another count, another result
    assert any(_draws_other(seed, 0, e, 4 if e == 1 else 5) !=
               _draws_other(seed + 1, 0, e, 4 if e == 1 else 5) for e in range(1, 4))


def test_adam_parallel_two_nodes_densmap_hand_computed():
    # the density term of the parallel optimiser (umap_optimisers.rs:1303-1317,
    # 1369-1387) on two nodes: each node's only neighbour is the other, so both
    # log radii are log(1e-8 + d^2), their variance is 0, std = sqrt(var_shift),
    # cov = 0, the weights are r itself and coeff = lambda mu_tot / std.
    n_epochs, seed = 3, 11
    p = _two_node_params(n_epochs)
    lam, frac, vs = 2.0, 1.0, 0.1  # frac 1: every epoch is a density epoch
    r = np.array([1.0, 0.5], np.float32)
    mu_tot = 2.0
    y0 = np.array([[0.0, 0.0], [1.0, 0.5]])
    adj = mrs_amd.adjacency_to_csr([[(1, 1.0)], [(0, 1.0)]])
    y = y0.copy()
    m = np.zeros_like(y)
    v = np.zeros_like(y)
    for e in range(1, n_epochs):
        d2 = float(((y[0] - y[1]) ** 2).sum())
        phi = 1.0 / (1.0 + p.a * d2 ** p.b)
        re = math.log(1e-8 + d2)
        coeff = lam * mu_tot / math.sqrt(vs)
        dr = (1.0 - p.b * (1.0 - phi)) / math.exp(re) + p.a * p.b * d2 ** p.b / d2 * phi
        cor = coeff * (r[0] + r[1]) * dr / 1.0
        gs = []
        for i, o in ((0, 1), (1, 0)):
            gi = _attr(p.a, p.b, y[i], y[o]) + 2.0 * np.clip(2.0 * cor * (y[i] - y[o]), -4, 4)
            if i == 0:
                gi = gi + _draws_other(seed, 0, e, 4 if e == 1 else 5) * _rep(
                    p.a, p.b, p.gamma, y[0], y[1])
            gs.append(gi)
        g = np.stack(gs)
        m += (1 - p.beta1) * (g - m)
        v += (1 - p.beta2) * (g * g - v)
        t = e + 1
        sq = math.sqrt(1 - p.beta2 ** t)
        y = y + p.lr * (1 - e / n_epochs) * sq / (1 - p.beta1 ** t) * m / (np.sqrt(v) + sq * p.eps)
    got = orc.optimise_adam_parallel(y0.astype(np.float32), *adj, p, seed, use_lut=False,
                                     dens_r=r, dens_params=(lam, frac, vs), mu_tot=mu_tot)
    np.testing.assert_allclose(got, y, rtol=2e-5, atol=2e-6)
    plain = orc.optimise_adam_parallel(y0.astype(np.float32), *adj, p, seed, use_lut=False)
    assert np.max(np.abs(got - plain)) > 1e-3  # the density term moved the points


@pytest.mark.parametrize("a,yb", [(1.0, (3.0, 1.0)), (25.0, (0.2, 0.0))])
def test_sgd_two_nodes_hand_computed(a, yb):
    # optimise_embedding_sgd (umap_optimisers.rs:653-865): the adjacency list is
    # walked as directed edges (0 -> 1 then 1 -> 0), each moving both endpoints
    # in place by lr * clip(coeff * delta, 4) per coordinate, then drawing its
    # source's negatives from the updated positions.  a = 25, d = 0.2 puts the
    # attractive step above the clip (coeff * delta = 5).
    n_epochs, seed = 3, 5
    p = _two_node_params(n_epochs, a=a)
    y0 = np.array([[0.0, 0.0], list(yb)])
    adj = mrs_amd.adjacency_to_csr([[(1, 1.0)], [(0, 1.0)]])
    y = y0.copy()
    for e in range(1, n_epochs):
        lr = p.lr * (1 - e / n_epochs)
        n_neg = 4 if e == 1 else 5
        for i, j in ((0, 1), (1, 0)):
            d2 = float(((y[i] - y[j]) ** 2).sum())
            db = d2 ** p.b
            coeff = 2.0 * p.a * p.b * db / (d2 * (1.0 + p.a * db))
            gd = np.clip(coeff * (y[j] - y[i]), -4.0, 4.0)
            y[i] += gd * lr
            y[j] -= gd * lr
            for s in range(n_neg):
                if orc.hash_neg(seed, i, e, 0, s, 2) == i:
                    continue
                y[i] += _rep(p.a, p.b, p.gamma, y[i], y[j]) * lr
    got = orc.optimise_sgd(y0.astype(np.float32), *adj, p, seed)
    np.testing.assert_allclose(got, y, rtol=2e-5, atol=2e-6)


def test_adam_parallel_three_nodes_fractional_schedule():
    # Edges (0, 1) weight 1 and (1, 2) weight 0.45.  Edge (1, 2): per =
    # 1 / 0.45 = 2.222, negative period 0.4444, first negative due at 0.4444;
    # it fires at epochs 3, 5 and 7 with floor((3 - 0.444) / 0.444) =
    # floor(5.75) = 5, floor((5 - 2.667) / 0.444) = floor(5.25) = 5 and
    # floor((7 - 4.889) / 0.444) = floor(4.75) = 4 negatives, drawn by node 1
    # (its smaller endpoint) under CSR slot 1.  Edge (0, 1) fires every epoch
    # from 1 (4, then 5 negatives, node 0, slot 0).  The fractional quotients
    # make the floor visible: ceil or round would draw 6 at epoch 3 and 5 at
    # epoch 7 (and with this seed the extra draws hit another node).
    n_epochs, seed, n = 8, 3, 3
    p = _two_node_params(n_epochs)
    y0 = np.array([[0.0, 0.0], [2.0, 1.0], [4.0, -1.0]])
    adj = mrs_amd.adjacency_to_csr([[(1, 1.0)], [(0, 1.0), (2, 0.45)], [(1, 0.45)]])
    fires = {0: {e: (4 if e == 1 else 5) for e in range(1, n_epochs)}, 1: {3: 5, 5: 5, 7: 4}}
    edges = {0: (0, 1, 0), 1: (1, 2, 1)}  # edge -> (smaller, larger, slot under smaller)
    y = y0.copy()
    m = np.zeros_like(y)
    v = np.zeros_like(y)
    for e in range(n_epochs):
        g = np.zeros_like(y)
        hu = np.zeros(n, bool)
        for ed, (i, j, slot) in edges.items():
            if e not in fires[ed]:
                continue
            hu[i] = hu[j] = True
            g[i] += _attr(p.a, p.b, y[i], y[j])
            g[j] += _attr(p.a, p.b, y[j], y[i])
            for s in range(fires[ed][e]):
                k = orc.hash_neg(seed, i, e, slot, s, n)
                if k != i:
                    g[i] += _rep(p.a, p.b, p.gamma, y[i], y[k])
        upd = hu[:, None]
        m = np.where(upd, m + (1 - p.beta1) * (g - m), m)
        v = np.where(upd, v + (1 - p.beta2) * (g * g - v), v)
        t = e + 1
        sq = math.sqrt(1 - p.beta2 ** t)
        step = p.lr * (1 - e / n_epochs) * sq / (1 - p.beta1 ** t) * m / (np.sqrt(v) + sq * p.eps)
        y = np.where(upd, y + step, y)
    got = orc.optimise_adam_parallel(y0.astype(np.float32), *adj, p, seed, use_lut=False)
    np.testing.assert_allclose(got, y, rtol=2e-5, atol=2e-6)


# ------------------------------------------------------ spectral layout


def _graph(x, k=15):
    from helpers import umap_graph_from_data

    _, _, (r, c, v), _ = umap_graph_from_data(x, k)
    return r, c, v


@pytest.mark.parametrize("n_dim", [2, 3])
def test_spectral_layout_single_component_spans_eigenvectors(n_dim):
    # one component of 40 vertices: 4 (n_dim + 1) + 30 >= 40 Lanczos steps span
    # the space, so the layout columns are exact eigenvectors 1..n_dim, centred,
    # scaled to max |x| = 10, plus N(0, 1e-4) noise
    from spectral_checks import check_components, rectangle

    x = rectangle(40, 4.0, 8, seed=3)
    r, c, v = _graph(x)
    emb = orc.spectral_layout(40, r, c, v, n_comp=n_dim, seed=5).astype(np.float64)
    check_components(emb, r, c, v, [slice(0, 40)], n_dim)
    assert abs(np.max(np.abs(emb)) - 10.0) < 1e-3
    assert np.all(np.abs(emb.mean(0)) < 1e-3)


def test_spectral_layout_alignment_only_chooses_signs():
    # align= picks each eigenvector's sign and nothing else: aligned to the
    # negated layout, the single-component layout is the negated layout, up to
    # the N(0, 1e-4) noise added after the sign is fixed (twice the noise)
    from spectral_checks import rectangle
    from scipy.spatial.distance import cdist

    x = rectangle(40, 4.0, 8, seed=3)
    d2 = cdist(x, x, "sqeuclidean")
    np.fill_diagonal(d2, np.inf)
    w = np.exp(-d2 / np.median(d2[np.isfinite(d2)]))
    w[w < 1e-3] = 0.0
    w = np.maximum(w, w.T)
    r, c = np.nonzero(w)
    v = w[r, c].astype(np.float32)
    ref = orc.spectral_layout(40, r, c, v, seed=5).astype(np.float64)
    for flip in ((1, 1), (-1, 1), (1, -1), (-1, -1)):
        target = (ref * np.array(flip)).astype(np.float32)
        got = orc.spectral_layout(40, r, c, v, seed=5, align=target).astype(np.float64)
        assert np.max(np.abs(got - target)) < 1e-3, flip


def test_spectral_layout_simplex_meta_and_expansion():
    # three components (<= 2 n_dim): centres on the simplex rows (+e1, +e2, -e1)
    # up to the final centring and scaling; each component reaches half its
    # centre's largest distance to another centre
    from spectral_checks import check_components, rectangle

    parts = [rectangle(40, 4.0, 8, seed=s) + 100.0 * s for s in (1, 2, 3)]
    x = np.concatenate(parts)
    r, c, v = _graph(x)
    emb = orc.spectral_layout(120, r, c, v, seed=5).astype(np.float64)
    centre, reach = check_components(emb, r, c, v, [slice(40 * k, 40 * k + 40) for k in range(3)], 2)
    simplex = np.array([[1.0, 0.0], [0.0, 1.0], [-1.0, 0.0]])
    # centre = s * (simplex - mean) for one scale s
    s = np.linalg.lstsq((simplex - simplex.mean(0)).ravel()[:, None],
                        (centre - centre.mean(0)).ravel(), rcond=None)[0][0]
    np.testing.assert_allclose(centre - centre.mean(0), s * (simplex - simplex.mean(0)), atol=1e-3)
    half = 0.5 * np.sqrt(((centre[:, None] - centre[None]) ** 2).sum(-1)).max(1)
    np.testing.assert_allclose(reach, half, rtol=1e-3)


def test_spectral_layout_data_centroid_meta():
    # six components (> 2 n_dim) with data: centres from the centroid affinity's
    # spectral layout, each component expanded to half its reach
    from spectral_checks import check_components, check_meta_spectral, rectangle

    rng = np.random.default_rng(29)
    offsets = rng.normal(0.0, 60.0, size=(6, 8))
    parts = [rectangle(40, 4.0, 8, seed=s) + offsets[s] for s in range(6)]
    x = np.concatenate(parts).astype(np.float32)
    r, c, v = _graph(x)
    emb = orc.spectral_layout(240, r, c, v, seed=5, data=x).astype(np.float64)
    centre, reach = check_components(emb, r, c, v, [slice(40 * k, 40 * k + 40) for k in range(6)], 2)
    check_meta_spectral(centre, parts, 2)
    half = 0.5 * np.sqrt(((centre[:, None] - centre[None]) ** 2).sum(-1)).max(1)
    np.testing.assert_allclose(reach, half, rtol=1e-3)
    # without data the same graph falls back to the simplex: only 2 n_dim = 4
    # distinct centres, the other two at the origin of the meta layout
    emb0 = orc.spectral_layout(240, r, c, v, seed=5).astype(np.float64)
    c0, _ = check_components(emb0, r, c, v, [slice(40 * k, 40 * k + 40) for k in range(6)], 2)
    simplex = np.array([[1, 0], [0, 1], [0, 0], [-1, 0], [0, -1], [0, 0]], np.float64)
    s = np.linalg.lstsq((simplex - simplex.mean(0)).ravel()[:, None],
                        (c0 - c0.mean(0)).ravel(), rcond=None)[0][0]
    np.testing.assert_allclose(c0 - c0.mean(0), s * (simplex - simplex.mean(0)), atol=1e-3)


def test_spectral_layout_small_component_is_scattered_near_its_centre():
    # a component of 3 < 2 n_dim vertices is placed uniformly within its data
    # range of its centre, not spectrally
    from spectral_checks import rectangle

    big = rectangle(40, 4.0, 8, seed=1)
    r, c, v = _graph(big)
    # three extra vertices joined to each other only
    extra = [(40, 41), (41, 42), (40, 42)]
    r2 = np.concatenate([r, [a for a, b in extra] + [b for a, b in extra]])
    c2 = np.concatenate([c, [b for a, b in extra] + [a for a, b in extra]])
    v2 = np.concatenate([v, np.full(6, 0.5, np.float32)])
    o = np.lexsort((c2, r2))
    emb = orc.spectral_layout(43, r2[o], c2[o], v2[o], seed=3).astype(np.float64)
    assert np.all(np.isfinite(emb)) and abs(np.max(np.abs(emb)) - 10.0) < 1e-3
    # the two centres (+e1, +e2 on the simplex) are sqrt(2) apart, so the small
    # component stays within sqrt(2)/2 of its centre in meta units
    small = emb[40:43]
    assert np.ptp(small, axis=0).max() > 0  # scattered, not collapsed


# ----------------------------------------------------------------- PCA layout

_PCA_5x3 = np.array([[1.0, 2.0, 1.0], [2.0, 3.0, 2.0], [3.0, 4.0, 1.5], [4.0, 5.0, 2.5],
                     [5.0, 6.0, 2.0]], np.float32)


def test_pca_layout_reference_known_answers():
    # src/data/init.rs:1495-1566: on the 5 x 3 matrix the layout has mean 0
    # (1e-6) and the first component's standard deviation is PCA_RANGE = 1
    # (1e-4), or the custom range; repeated calls are identical; a 3 x 4
    # matrix gives 3 x 2
    e = orc.pca_layout(_PCA_5x3, 2).astype(np.float64)
    assert e.shape == (5, 2)
    assert np.all(np.abs(e.mean(0)) < 1e-6)
    assert abs(e[:, 0].std() - 1.0) < 1e-4
    for rng in (1.0, 0.5, 3.0):
        er = orc.pca_layout(_PCA_5x3, 2, range_=rng).astype(np.float64)
        assert abs(er[:, 0].std() - rng) < 1e-4 * rng
    assert np.array_equal(orc.pca_layout(_PCA_5x3, 2), e.astype(np.float32))
    x34 = np.array([[1, 2, 3, 4], [5, 6, 7, 8], [9, 10, 11, 12]], np.float32)
    assert orc.pca_layout(x34, 2).shape == (3, 2)


def _numpy_pca(x, n_comp, range_=1.0):
    # the reference's construction with numpy's SVD in float64: scores U S of
    # the centred data, each column centred and scaled to std range_ (a
    # column of std <= 1e-8 keeps scale 1); the sign of each component makes
    # its largest-magnitude loading positive
    xc = x.astype(np.float64) - x.astype(np.float64).mean(0)
    u, s, vt = np.linalg.svd(xc, full_matrices=False)
    out = np.empty((x.shape[0], n_comp))
    for c in range(n_comp):
        sg = 1.0 if vt[c][np.argmax(np.abs(vt[c]))] >= 0 else -1.0
        sc = u[:, c] * s[c] * sg
        sd = sc.std()
        out[:, c] = (sc - sc.mean()) * (range_ / sd if sd > 1e-8 else 1.0)
    return out


@pytest.mark.parametrize("n,d,n_comp", [(400, 12, 2), (300, 7, 3), (50, 3, 3), (2000, 40, 5)])
def test_pca_layout_equals_numpy_svd(n, d, n_comp):
    rng = np.random.default_rng(n + d)
    scales = np.linspace(3.0, 0.2, d)  # distinct variances: a unique basis
    x = (rng.normal(size=(n, d)) * scales @ np.linalg.qr(rng.normal(size=(d, d)))[0] + 5.0)
    x = x.astype(np.float32)
    got = orc.pca_layout(x, n_comp).astype(np.float64)
    want = _numpy_pca(x, n_comp)
    assert np.max(np.abs(got - want)) < 1e-5


def test_pca_layout_rank_one_data_recovers_the_line():
    # closed form: points a_r u + m on a line have one component, a_r scaled to
    # unit std (up to the sign fixed by u's largest entry)
    rng = np.random.default_rng(3)
    a = rng.normal(size=300)
    u = np.array([0.2, -0.9, 0.3, 0.25])
    u /= np.linalg.norm(u)
    x = (a[:, None] * u + np.array([1.0, 2.0, -3.0, 0.5])).astype(np.float32)
    got = orc.pca_layout(x, 1).astype(np.float64)[:, 0]
    want = (a - a.mean()) / a.std() * -1.0  # u's largest-magnitude entry is negative
    assert np.max(np.abs(got - want)) < 1e-4


def test_pca_layout_rejects_bad_shapes():
    x = np.ones((5, 3), np.float32)
    for nc in (0, 4):
        with pytest.raises(RuntimeError):
            orc.pca_layout(x, nc)
