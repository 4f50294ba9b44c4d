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
