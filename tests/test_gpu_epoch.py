"""GPU parity of the "adam_gpu" epoch kernels against the CPU oracle.

Ports the reference's kernel tests (src/training/umap_optimiser_gpu.rs:1786-2505)
and adds seeded larger graphs.  Tolerances: the reference's own (bit-exact for
attractive-only, 1e-4 with negatives, 1e-5/1e-6 for Adam).  On larger graphs a
node gradient is a sum of hundreds of clipped terms that can cancel, so each
coordinate is held to 1e-4 relative to the sum of its terms' magnitudes
(|got - want| <= 1e-4 * sum|terms|, the oracle returning the sum), with want
the exact (float64) sum of the oracle's f32 terms: that bounds any reordering
of the sum (the edge kernel's FP32 atomics) and the device's powf/expf ulps,
and a term dropped or doubled fails it.  The exact sum is the yardstick
because a sequential f32 sum is itself off by more than the bar on a hub: on
the 70 000-neighbour star below the oracle's own f32 sum misses the exact one
by 3.7e-4 of sum|terms|.  Whole runs are
compared epoch by epoch from the same state (teacher forcing): Adam divides by
sqrt(v), so a free-running comparison amplifies ulp differences in near-zero
gradients into O(lr) steps within a few epochs.
"""
import numpy as np
import pytest

import mrs_amd as M
import pyoracle as orc
from helpers import (assert_close, assert_grad_close, assert_scaled_close, blobs, coo_to_adjacency,
                     diff_stats, triangle_setup, umap_graph_from_data)

pytestmark = pytest.mark.gpu


def make_state(ctx, embd, graph, params, dens=None):
    adj = M.adjacency_to_csr(graph) if isinstance(graph, list) else graph
    return M.UmapGpuState(ctx, embd, *adj, params, dens), adj


def oracle_csr(adj):
    st, c = orc.csr_from_graph(*adj)
    assert st == 0
    return c


# ------------------------------------------------------- upload / state


def test_state_roundtrips(ctx):
    embd, graph, params = triangle_setup()
    st, adj = make_state(ctx, embd, graph, params)
    c = oracle_csr(adj)
    assert np.array_equal(st.read(M.BUF_EMBD), embd.ravel())
    for b in (M.BUF_M, M.BUF_V, M.BUF_GRAD):
        assert np.all(st.read(b) == 0)
    assert np.array_equal(st.read(M.BUF_OFFSETS), c["off"])
    assert np.array_equal(st.read(M.BUF_EDGE_IDX), c["edge_idx"])
    assert np.array_equal(st.read(M.BUF_OTHER), c["other"])
    assert np.array_equal(st.read(M.BUF_PER), c["per"])
    assert np.array_equal(st.read(M.BUF_NEXT), c["per"])


def test_upload_3d_embedding(ctx):
    _, graph, params = triangle_setup()
    embd = np.arange(9, dtype=np.float32).reshape(3, 3)
    st, _ = make_state(ctx, embd, graph, params)
    assert st.n_dim == 3
    assert np.array_equal(st.read(M.BUF_EMBD), embd.ravel())


def test_state_errors(ctx):
    _, _, params = triangle_setup()
    with pytest.raises(M.ManifoldsError) as e:
        make_state(ctx, np.zeros((3, 2), np.float32), [[], [], []], params)
    assert e.value.status == M.MRS_NO_GRAPH_EDGES


# ------------------------------------------------------------ gradient


def test_grad_empty_epoch(ctx):
    embd, graph, params = triangle_setup()
    st, _ = make_state(ctx, embd, graph, params)
    st.launch_grad_accum(0, 42)
    assert np.all(st.read(M.BUF_GRAD) == 0)


def test_grad_attractive_only_bit_exact(ctx):
    # :2005-2032
    embd, graph, params = triangle_setup()
    params.neg_sample_rate = 0
    st, adj = make_state(ctx, embd, graph, params)
    st.launch_grad_accum(5, 42)
    c = oracle_csr(adj)
    want, _ = orc.grad_accum(embd, c, st.read(M.BUF_NEXT), 5, 42, 0, params)
    assert_close(st.read(M.BUF_GRAD), want, 0.0, "attractive-only")


def test_grad_with_negatives_matches_cpu(ctx):
    # :2034-2058
    embd, graph, params = triangle_setup()
    st, adj = make_state(ctx, embd, graph, params)
    st.launch_grad_accum(5, 42)
    c = oracle_csr(adj)
    want, _ = orc.grad_accum(embd, c, st.read(M.BUF_NEXT), 5, 42, 5, params)
    assert_close(st.read(M.BUF_GRAD), want, 1e-4, "with-negatives")


def test_grad_single_edge_hand_computed(ctx):
    # :2060-2092
    embd = np.array([[0.0, 0.0], [3.0, 0.0]], np.float32)
    p = M.UmapOptimParams(a=1.0, b=1.0, lr=1.0, gamma=1.0, n_epochs=10, neg_sample_rate=0,
                          min_dist=0.1, beta1=0.9, beta2=0.999, eps=1e-7)
    st, _ = make_state(ctx, embd, [[(1, 1.0)], [(0, 1.0)]], p)
    st.launch_grad_accum(5, 42)
    assert_close(st.read(M.BUF_GRAD), [1.2, 0.0, -1.2, 0.0], 1e-5, "hand-computed")
    st.launch_epoch_edges(5, 42)
    assert_close(st.read(M.BUF_GRAD), [1.2, 0.0, -1.2, 0.0], 1e-5, "hand-computed edges")


def test_grad_deterministic(ctx):
    embd, graph, params = triangle_setup()
    st, _ = make_state(ctx, embd, graph, params)
    st.launch_grad_accum(5, 42)
    g1 = st.read(M.BUF_GRAD)
    st.launch_grad_accum(5, 42)
    assert np.array_equal(g1, st.read(M.BUF_GRAD))


def _random_graph(n=3000, d=8, k=15, seed=21):
    x, _ = blobs(n, d, 6, seed=seed)
    _, _, _, adj = umap_graph_from_data(x, k)
    rng = np.random.default_rng(seed)
    embd = rng.uniform(-10, 10, size=(n, 2)).astype(np.float32)
    return embd, adj


@pytest.mark.parametrize("epoch", [1, 3, 17])
def test_grad_random_graph_node_and_edge_kernels(ctx, epoch):
    embd, adj = _random_graph()
    params = M.UmapOptimParams.default_2d()
    st, _ = make_state(ctx, embd, adj, params)
    c = oracle_csr(adj)
    nxt = st.read(M.BUF_NEXT)
    want, hu, absum = orc.grad_accum(embd, c, nxt, epoch, 42, 5, params, with_abs=True,
                                     exact=True)
    st.launch_grad_accum(epoch, 42)
    g_node = st.read(M.BUF_GRAD).reshape(-1, 2)
    assert np.array_equal(st.read(M.BUF_HAS_UPDATE), hu), "node kernel has_update flags"
    w1 = assert_grad_close(g_node, want, absum, 1e-4, f"node kernel epoch {epoch}")
    # edge-parallel: same gradient by atomics, same flags, schedule advanced
    st.launch_epoch_edges(epoch, 42)
    g_edge = st.read(M.BUF_GRAD).reshape(-1, 2)
    assert np.array_equal(st.read(M.BUF_HAS_UPDATE), hu), "edge kernel has_update flags"
    w2 = assert_grad_close(g_edge, want, absum, 1e-4, f"edge kernel epoch {epoch}")
    assert np.array_equal(st.read(M.BUF_NEXT), orc.schedule_update(c["per"], nxt, epoch))
    print(f"epoch {epoch}: max |diff| / sum|terms|: node {w1:.2e}, edge {w2:.2e}")


# ------------------------------------------------------------------ Adam


def _run_adam(ctx, embd, m, v, grad, hu, n, dim, params, epoch):
    graph = [[((i + 1) % max(n, 2), 1.0)] for i in range(max(n, 2))]
    e = np.zeros((max(n, 2), dim), np.float32)
    e[:n] = np.asarray(embd, np.float32).reshape(n, dim)
    st, _ = make_state(ctx, e, graph, params)

    def pad(a, dt=np.float32):
        out = np.zeros(max(n, 2) * (dim if dt == np.float32 else 1), dt)
        a = np.asarray(a, dt).ravel()
        out[:a.size] = a
        return out

    st.write(M.BUF_M, pad(m))
    st.write(M.BUF_V, pad(v))
    st.write(M.BUF_GRAD, pad(grad))
    st.write(M.BUF_HAS_UPDATE, pad(hu, np.uint32))
    st.launch_adam_update(epoch)
    k = n * dim
    return st.read(M.BUF_EMBD)[:k], st.read(M.BUF_M)[:k], st.read(M.BUF_V)[:k]


def test_adam_inactive_node_preserved(ctx):
    p = M.UmapOptimParams.default_2d()
    p.beta1, p.beta2, p.n_epochs = 0.9, 0.999, 10
    embd = np.array([1.5, -2.5, 7.0, 8.0], np.float32)
    m = np.array([0.1, 0.2, 0.3, 0.4], np.float32)
    v = np.array([0.01, 0.02, 0.03, 0.04], np.float32)
    e, mm, vv = _run_adam(ctx, embd, m, v, np.full(4, 5.0), [0, 0], 2, 2, p, 3)
    assert np.array_equal(e, embd) and np.array_equal(mm, m) and np.array_equal(vv, v)


def test_adam_single_step_hand_computed(ctx):
    p = M.UmapOptimParams(a=1.5, b=0.9, lr=1.0, gamma=1.0, n_epochs=10, neg_sample_rate=5,
                          min_dist=0.1, beta1=0.9, beta2=0.999, eps=1e-7)
    e, mm, vv = _run_adam(ctx, [0.0], [0.0], [0.0], [1.0], [1], 1, 1, p, 0)
    assert_close(mm, [0.1], 1e-6, "m")
    assert_close(vv, [0.001], 1e-6, "v")
    assert_close(e, [1.0], 1e-4, "embd")


def test_adam_matches_cpu_reference(ctx):
    # :2244-2295
    p = M.UmapOptimParams(a=1.5, b=0.9, lr=1.0, gamma=1.0, n_epochs=10, neg_sample_rate=5,
                          min_dist=0.1, beta1=0.9, beta2=0.999, eps=1e-7)
    n, dim = 5, 2
    embd = np.arange(n * dim, dtype=np.float32) * 0.5 - 1.0
    m = np.arange(n * dim, dtype=np.float32) * 0.01
    v = np.arange(n * dim, dtype=np.float32) * 0.001 + 1e-6
    grad = (((np.arange(n * dim) * 7 + 3) % 5) * 0.1).astype(np.float32)
    hu = np.array([1, 0, 1, 1, 0], np.uint32)
    e, mm, vv = _run_adam(ctx, embd, m, v, grad, hu, n, dim, p, 4)
    ew, mw, vw = orc.adam_update(grad, hu, m, v, embd, p, 4, 10)
    assert_close(e, ew, 1e-5, "embd")
    assert_close(mm, mw, 1e-6, "m")
    assert_close(vv, vw, 1e-6, "v")


# -------------------------------------------------------------- schedule


@pytest.mark.parametrize("epoch,want", [(5, [2.0, 4.0, 2.0]), (0, [1.0, 2.0, 1.0]),
                                        (1, [2.0, 2.0, 2.0])])
def test_schedule_goldens(ctx, epoch, want):
    # :2342-2379
    embd, graph, params = triangle_setup()
    st, _ = make_state(ctx, embd, graph, params)
    st.launch_edge_schedule_update(epoch)
    assert list(st.read(M.BUF_NEXT)) == want


# --------------------------------------------------------------- densMAP


def _dens_setup(n):
    # dens_setup, :1873-1885
    r = np.arange(n, dtype=np.float32) - (n - 1) / 2.0
    norm = np.sqrt((r * r).sum() / (n - 1))
    r = (r / norm).astype(np.float32)
    mu_sum = np.full(n, 2.0, np.float32)
    ds = M.DensState(params=M.DensParams(lambda_=2.0, frac=1.0, var_shift=0.1),
                     r=r.ctypes.data, mu_sum=mu_sum.ctypes.data, mu_tot=2.0 * n)
    return ds, (r, mu_sum)


def test_dens_radii_matches_cpu(ctx):
    # :1900-1924
    embd, graph, params = triangle_setup()
    ds, keep = _dens_setup(3)
    st, adj = make_state(ctx, embd, graph, params, ds)
    st.launch_dens_update()
    re_w, phi_w = orc.dens_radii(embd, oracle_csr(adj), params.a, params.b)
    assert_close(st.read(M.BUF_DENS_RE), re_w, 1e-5, "dens-re")
    assert_close(st.read(M.BUF_DENS_PHI), phi_w, 1e-5, "dens-phi")


def test_grad_with_density_matches_cpu(ctx):
    # :1926-1993
    embd, graph, params = triangle_setup()
    params.neg_sample_rate = 0
    ds, (r, _) = _dens_setup(3)
    st, adj = make_state(ctx, embd, graph, params, ds)
    coeff = st.launch_dens_update()
    st.launch_grad_accum(5, 42, coeff)
    got = st.read(M.BUF_GRAD)
    c = oracle_csr(adj)
    re, phi = orc.dens_radii(embd, c, params.a, params.b)
    w, coeff_w = orc.dens_weights(r, re, 0.1, 2.0, 6.0)
    assert abs(coeff - coeff_w) <= 1e-6 * abs(coeff_w)
    want, _ = orc.grad_accum(embd, c, st.read(M.BUF_NEXT), 5, 42, 0, params,
                             dict(re=re, phi_sum=phi, weight=w, coeff=coeff_w))
    assert_close(got, want, 1e-4, "with-density")
    st.launch_epoch_edges(5, 42, coeff)
    assert_close(st.read(M.BUF_GRAD), want, 1e-4, "with-density edges")
    plain, _ = orc.grad_accum(embd, c, st.read(M.BUF_NEXT), 5, 42, 0, params)
    assert np.max(np.abs(got - plain.ravel())) > 1e-6


def test_dens_random_graph_stats_and_grad(ctx):
    embd, adj = _random_graph(n=2500, seed=31)
    rng = np.random.default_rng(31)
    r = rng.normal(size=embd.shape[0]).astype(np.float32)
    r = ((r - r.mean()) / r.std(ddof=1)).astype(np.float32)
    mu_sum = np.ones_like(r)
    ds = M.DensState(params=M.DensParams(lambda_=2.0, frac=0.3, var_shift=0.1),
                     r=r.ctypes.data, mu_sum=mu_sum.ctypes.data, mu_tot=1234.5)
    params = M.UmapOptimParams.default_2d()
    st, _ = make_state(ctx, embd, adj, params, ds)
    c = oracle_csr(adj)
    coeff = st.launch_dens_update()
    re, phi = orc.dens_radii(embd, c, params.a, params.b)
    assert_scaled_close(st.read(M.BUF_DENS_RE), re, 1e-5, "dens re")
    assert_scaled_close(st.read(M.BUF_DENS_PHI), phi, 1e-5, "dens phi_sum")
    w, coeff_w = orc.dens_weights(r, re, 0.1, 2.0, 1234.5)
    assert abs(coeff - coeff_w) <= 1e-5 * abs(coeff_w), f"coeff {coeff!r} vs {coeff_w!r}"
    assert_scaled_close(st.read(M.BUF_DENS_W), w, 1e-5, "dens weights")
    want, _, absum = orc.grad_accum(embd, c, st.read(M.BUF_NEXT), 7, 42, 5, params,
                                    dict(re=re, phi_sum=phi, weight=w, coeff=coeff_w),
                                    with_abs=True, exact=True)
    st.launch_grad_accum(7, 42, coeff)
    assert_grad_close(st.read(M.BUF_GRAD).reshape(-1, 2), want, absum, 1e-4, "dens node kernel")
    st.launch_epoch_edges(7, 42, coeff)
    assert_grad_close(st.read(M.BUF_GRAD).reshape(-1, 2), want, absum, 1e-4, "dens edge kernel")


# ------------------------------------------------------------- driver


def test_driver_preserves_graph_structure(ctx):
    # :2386-2448
    graph = [[(1, 1.0), (2, 1.0)], [(0, 1.0), (2, 1.0)], [(0, 1.0), (1, 1.0), (3, 0.1)],
             [(2, 0.1), (4, 1.0), (5, 1.0)], [(3, 1.0), (5, 1.0)], [(3, 1.0), (4, 1.0)]]
    embd = np.array([[0, 0], [10, 0], [0, 10], [10, 10], [-5, -5], [15, 15]], np.float32)
    p = M.UmapOptimParams.default_2d()
    p.n_epochs = 200
    adj = M.adjacency_to_csr(graph)
    for strat in (M.EPOCH_EDGE_ATOMIC, M.EPOCH_NODE_EXACT):
        out = M.optimise_embedding_adam_gpu(ctx, embd, *adj, p, seed=42, strategy=strat)
        dist = lambda i, j: float(np.linalg.norm(out[i] - out[j]))
        intra = (dist(0, 1) + dist(0, 2) + dist(1, 2) + dist(3, 4) + dist(3, 5) + dist(4, 5)) / 6
        inter = np.mean([dist(i, j) for i in (0, 1) for j in (3, 4, 5)])
        assert inter > 1.5 * intra
        assert np.all(np.isfinite(out))


def _teacher_forced(ctx, embd, adj, p, strategy, gtol=1e-4, seed=42, epochs=None):
    """Every epoch runs on the GPU (run_epochs(e, 1)) from the GPU's own state
    and is replayed on the oracle (grad_accum + adam_update + schedule_update)
    from that same state.  Schedule: exact.  m: within (1-beta1) * the gradient
    bar gtol * sum|terms| of each coordinate.  Embedding: that bar carried
    through Adam's step fused_lr * m / (sqrt(v) + eps), i.e. |d embd| <=
    4 lr gtol sum|terms| / (sqrt(v) + eps) plus rounding."""
    st, _ = make_state(ctx, embd, adj, p)
    c = oracle_csr(adj)
    n, dim = embd.shape
    worst = 0.0
    for e in range(p.n_epochs if epochs is None else epochs):
        e0 = st.read(M.BUF_EMBD).reshape(n, dim)
        m0, v0, nx0 = st.read(M.BUF_M), st.read(M.BUF_V), st.read(M.BUF_NEXT)
        st.run_epochs(e, 1, seed, strategy)
        g, hu, absum = orc.grad_accum(e0, c, nx0, e, seed, p.neg_sample_rate, p, with_abs=True,
                                      exact=True)
        ew, mw, vw = orc.adam_update(g, hu, m0, v0, e0, p, e, p.n_epochs)
        assert np.array_equal(st.read(M.BUF_NEXT), orc.schedule_update(c["per"], nx0, e)), \
            f"schedule epoch {e}"
        dg = gtol * absum.ravel()
        got_m = st.read(M.BUF_M)
        dm = np.abs(got_m - mw.ravel())
        mbad = np.flatnonzero(dm > 1e-6 * (1 + np.abs(mw.ravel())) + dg)
        assert mbad.size == 0, (f"m epoch {e}: {mbad.size} coords over the bar, first {mbad[0]}: "
                                f"diff {dm[mbad[0]]:.3e} bar {dg[mbad[0]]:.3e}; "
                                f"{diff_stats(got_m, mw)}")
        got = st.read(M.BUF_EMBD)
        d = np.abs(got - ew.ravel())
        bar = 1e-5 * (1 + np.abs(ew.ravel())) + 4 * p.lr * dg / (np.sqrt(vw.ravel()) + p.eps)
        bad = np.flatnonzero(d > bar)
        assert bad.size == 0, (f"embd epoch {e}: {bad.size} coords over the bar, first {bad[0]}: "
                               f"diff {d[bad[0]]:.3e} bar {bar[bad[0]]:.3e}; {diff_stats(got, ew)}")
        worst = max(worst, float(d.max()))
    return st, worst


def test_driver_node_exact_bit_identical_rerun_and_oracle(ctx):
    # :2474-2505 (bit-identical rerun) and a per-epoch replay on the oracle
    embd, adj = _random_graph(n=1500, seed=41)
    p = M.UmapOptimParams.default_2d()
    p.n_epochs = 30
    a = M.optimise_embedding_adam_gpu(ctx, embd, *adj, p, seed=42, strategy=M.EPOCH_NODE_EXACT)
    b = M.optimise_embedding_adam_gpu(ctx, embd, *adj, p, seed=42, strategy=M.EPOCH_NODE_EXACT)
    assert np.array_equal(a, b)
    _teacher_forced(ctx, embd, adj, p, M.EPOCH_NODE_EXACT, 1e-4)
    # the free-running driver against the oracle's replay of the whole driver:
    # reported, with a loose bar (ulp differences grow through Adam)
    want = orc.optimise_adam_gpu(embd, *adj, p, 42)
    d = np.abs(a - want)
    print("node-exact free run vs oracle:", diff_stats(a, want))
    assert np.all(np.isfinite(a)) and float(np.median(d)) < 0.1


@pytest.mark.parametrize("neg_rate", [5, 3])
def test_driver_edge_atomic_matches_oracle(ctx, neg_rate):
    embd, adj = _random_graph(n=1500, seed=43)
    p = M.UmapOptimParams.default_2d()
    p.n_epochs = 30
    p.neg_sample_rate = neg_rate
    _teacher_forced(ctx, embd, adj, p, M.EPOCH_EDGE_ATOMIC, 1e-4)
    got = M.optimise_embedding_adam_gpu(ctx, embd, *adj, p, seed=42, strategy=M.EPOCH_EDGE_ATOMIC)
    want = orc.optimise_adam_gpu(embd, *adj, p, 42)
    print("edge-atomic free run vs oracle:", diff_stats(got, want))
    assert np.all(np.isfinite(got)) and float(np.median(np.abs(got - want))) < 0.1


def test_driver_densmap_matches_oracle(ctx):
    embd, adj = _random_graph(n=1200, seed=45)
    rng = np.random.default_rng(45)
    r = rng.normal(size=embd.shape[0]).astype(np.float32)
    mu_sum = np.ones_like(r)
    ds = M.DensState(params=M.DensParams(lambda_=2.0, frac=0.3, var_shift=0.1),
                     r=r.ctypes.data, mu_sum=mu_sum.ctypes.data, mu_tot=float(adj[2].sum()))
    p = M.UmapOptimParams.default_2d()
    p.n_epochs = 20
    got = M.optimise_embedding_adam_gpu(ctx, embd, *adj, p, dens=ds, seed=42,
                                        strategy=M.EPOCH_NODE_EXACT)
    want = orc.optimise_adam_gpu(embd, *adj, p, 42, dens_r=r, dens_params=(2.0, 0.3, 0.1),
                                 mu_tot=float(adj[2].sum()))
    assert np.max(np.abs(got - want)) < 1e-3


@pytest.mark.parametrize("strategy", [M.EPOCH_EDGE_ATOMIC, M.EPOCH_NODE_EXACT])
def test_dens_radii_match_oracle(ctx, strategy):
    # umap_dens_radii (umap_optimiser_gpu.rs:487-544) at the start of a density
    # epoch, from the same embedding as the oracle.  The edge-parallel strategy
    # sums each node's row in 16 lanes (another order), the node-exact one in
    # the reference's; device and host powf differ in the last place either
    # way: phi sums within 1e-5 relative, log radii within 1e-5.
    embd, adj = _random_graph(n=1200, seed=47)
    rng = np.random.default_rng(47)
    r = rng.normal(size=embd.shape[0]).astype(np.float32)
    ms = np.ones_like(r)
    ds = M.DensState(params=M.DensParams(lambda_=2.0, frac=1.0, var_shift=0.1),
                     r=r.ctypes.data, mu_sum=ms.ctypes.data, mu_tot=float(adj[2].sum()))
    p = M.UmapOptimParams.default_2d()
    p.n_epochs = 10
    st, _ = make_state(ctx, embd, adj, p, ds)
    st.run_epochs(0, 1, 42, strategy)  # frac 1: epoch 0 is a density epoch
    want_re, want_phi = orc.dens_radii(embd, oracle_csr(adj), p.a, p.b)
    got_re, got_phi = st.read(M.BUF_DENS_RE), st.read(M.BUF_DENS_PHI)
    np.testing.assert_allclose(got_phi, want_phi, rtol=1e-5, atol=0)
    np.testing.assert_allclose(got_re, want_re, rtol=0, atol=1e-5)


@pytest.mark.parametrize("dens_on", [False, True])
def test_graph_replay_matches_eager(ctx, dens_on):
    # run_epochs(count >= 16) replays a captured HIP graph; it must equal the
    # eager epoch-by-epoch launches bit for bit (node-exact strategy), on every
    # replay, with eager work and syncs in between
    embd, adj = _random_graph(n=1200, seed=51)
    p = M.UmapOptimParams.default_2d()
    p.n_epochs = 40
    ds = None
    if dens_on:
        rng = np.random.default_rng(51)
        r = rng.normal(size=embd.shape[0]).astype(np.float32)
        ms = np.ones_like(r)
        ds = M.DensState(params=M.DensParams(lambda_=2.0, frac=0.5, var_shift=0.1),
                         r=r.ctypes.data, mu_sum=ms.ctypes.data, mu_tot=float(adj[2].sum()))
    st, _ = make_state(ctx, embd, adj, p, ds)
    for e in range(40):
        st.run_epochs(e, 1, 42, M.EPOCH_NODE_EXACT)
    eager = st.read(M.BUF_EMBD)
    for rep in range(3):
        st.reset(embd)
        st.run_epochs(0, 40, 42, M.EPOCH_NODE_EXACT)
        got = st.read(M.BUF_EMBD)
        assert np.array_equal(got, eager), f"replay {rep}"
        st.launch_grad_accum(3, 7)  # unrelated eager work between replays
    # the atomic strategy through the graph stays close to its eager run
    st.reset(embd)
    for e in range(40):
        st.run_epochs(e, 1, 42, M.EPOCH_EDGE_ATOMIC)
    eager_a = st.read(M.BUF_EMBD)
    st.reset(embd)
    st.run_epochs(0, 40, 42, M.EPOCH_EDGE_ATOMIC)
    got_a = st.read(M.BUF_EMBD)
    # atomics reorder FP32 sums, so the two runs drift apart through Adam:
    # reported, loose bar
    print("edge-atomic graph vs eager:", diff_stats(got_a, eager_a))
    assert np.all(np.isfinite(got_a)) and float(np.median(np.abs(got_a - eager_a))) < 0.1


# ------------------------------------------- wide spans and hub nodes


def _random_sparse_graph(n, per_node, seed, w_lo=0.05):
    """A random symmetric graph: every node picks per_node random partners;
    weights uniform in (w_lo, 1].  Returns the adjacency (CSR of both
    directions, columns ascending per row) and the undirected edge count."""
    rng = np.random.default_rng(seed)
    i = np.repeat(np.arange(n, dtype=np.int64), per_node)
    j = rng.integers(0, n, size=i.size)
    keep = i != j
    a, b = np.minimum(i[keep], j[keep]), np.maximum(i[keep], j[keep])
    key = np.unique(a * n + b)
    a, b = key // n, key % n
    w = rng.uniform(w_lo, 1.0, size=a.size).astype(np.float32)
    rows = np.concatenate([a, b])
    cols = np.concatenate([b, a])
    vals = np.concatenate([w, w])
    order = np.lexsort((cols, rows))
    return coo_to_adjacency(n, rows[order], cols[order], vals[order]), a.size


def test_edge_kernel_wide_span_teacher_forced(ctx):
    # Graphs of 2^20 or more edges take the 128-edge-span instantiation of the
    # edge kernel (every headline fit does); below that the 64-edge span runs.
    # Eight epochs on a 1.1M-edge graph, each replayed on the oracle from the
    # GPU's state: schedule exact, gradients and Adam within the per-node bar.
    adj, n_edges = _random_sparse_graph(160_000, 7, seed=61)
    assert n_edges >= 1 << 20, n_edges
    n = adj[0].size - 1
    embd = np.random.default_rng(61).uniform(-10, 10, size=(n, 2)).astype(np.float32)
    p = M.UmapOptimParams.default_2d()
    p.n_epochs = 8
    st, worst = _teacher_forced(ctx, embd, adj, p, M.EPOCH_EDGE_ATOMIC, epochs=8)
    fired = int(np.count_nonzero(st.read(M.BUF_NEXT) > st.read(M.BUF_PER)))
    print(f"wide span: {n_edges} edges, {fired} fired at least once, max |embd diff| {worst:.3e}")
    assert fired > n_edges // 4


@pytest.mark.parametrize("dim", [2, 1, 3, 8])
def test_edge_loop_firing_bitmask_windows(ctx, dim):
    # Edge-parallel runs of 8 or more epochs take each epoch's firing edges
    # from bitmask windows of up to 128 epochs, built ahead by an exact replay
    # of the schedule.  With lr = 0 the embedding stays put, so every epoch's
    # gradient is a function of its firing set alone, and the oracle replays
    # the same epochs from the same embedding.  Calls end just after window
    # boundaries (8 | 130 = 128 + 2 | 129 = 128 + 1 | 33); after each call:
    # schedule cursors exact; m (decay 0.5: the last few epochs' gradients
    # dominate) within the gradient bar accumulated over the epochs; v likewise
    # through g^2.  An edge that fires in the wrong epoch moves m of its
    # endpoints by a whole term.  Each dimension has its own lane layout for
    # the gradient atomics (a node's coordinates in one instruction).
    adj, n_edges = _random_sparse_graph(5000, 7, seed=67)
    n = adj[0].size - 1
    embd = np.random.default_rng(67).uniform(-10, 10, size=(n, dim)).astype(np.float32)
    p = M.UmapOptimParams.default_2d()
    p.n_epochs = 300
    p.lr = 0.0
    st, _ = make_state(ctx, embd, adj, p)
    c = oracle_csr(adj)
    gtol = 1e-4
    m = np.zeros(n * dim, np.float32)
    v = np.zeros(n * dim, np.float32)
    nx = c["per"].copy()
    dm = np.zeros(n * dim)  # accumulated bars of m and v
    dv = np.zeros(n * dim)
    e = 0
    for count in (8, 130, 129, 33):
        st.run_epochs(e, count, 42, M.EPOCH_EDGE_ATOMIC)
        for t in range(e, e + count):
            g, hu, absum = orc.grad_accum(embd, c, nx, t, 42, p.neg_sample_rate, p, with_abs=True,
                                          exact=True)
            ew, m2, v2 = orc.adam_update(g, hu, m, v, embd, p, t, p.n_epochs)
            assert np.array_equal(ew, embd)
            bar = gtol * absum.ravel()
            upd = np.repeat(hu.astype(bool), dim)
            dm = np.where(upd, p.beta1 * dm + (1 - p.beta1) * bar, dm)
            dv = np.where(upd, p.beta2 * dv + (1 - p.beta2) * (2 * np.abs(g.ravel()) * bar + bar ** 2),
                          dv)
            m, v = m2.ravel(), v2.ravel()
            nx = orc.schedule_update(c["per"], nx, t)
        e += count
        assert np.array_equal(st.read(M.BUF_NEXT), nx), f"schedule after epoch {e - 1}"
        got_m, got_v = st.read(M.BUF_M), st.read(M.BUF_V)
        mbad = np.flatnonzero(np.abs(got_m - m) > dm + 1e-6 * (1 + np.abs(m)))
        vbad = np.flatnonzero(np.abs(got_v - v) > dv + 1e-6 * (1 + np.abs(v)))
        assert mbad.size == 0, f"m after epoch {e - 1}: {mbad.size} coords over the bar"
        assert vbad.size == 0, f"v after epoch {e - 1}: {vbad.size} coords over the bar"
        assert np.array_equal(st.read(M.BUF_EMBD), embd.ravel())
    print(f"firing bitmask windows, dim {dim}: {n_edges} edges, 300 epochs, schedule exact, "
          f"m and v in bars")


def test_edge_kernel_hub_degree_above_65535(ctx):
    # Node 0 is joined to every other node (degree 69 999 > 65 535), the others
    # also form a ring.  The negatives of an edge are keyed by its CSR slot
    # under each endpoint, so the hub's slots must reach past 16 bits: the edge
    # kernel against the oracle (which walks the CSR itself), and the slots read
    # back from the edge records.
    n = 70_000
    rows, cols = [], []
    for i in range(1, n):
        rows += [0, i]
        cols += [i, 0]
        j = i + 1 if i + 1 < n else 1
        rows += [i, j]
        cols += [j, i]
    rows, cols = np.array(rows), np.array(cols)
    order = np.lexsort((cols, rows))
    vals = np.where((rows == 0) | (cols == 0), 1.0, 0.5).astype(np.float32)
    adj = coo_to_adjacency(n, rows[order], cols[order], vals[order])
    embd = np.random.default_rng(63).uniform(-10, 10, size=(n, 2)).astype(np.float32)
    p = M.UmapOptimParams.default_2d()
    p.n_epochs = 4
    st, _ = make_state(ctx, embd, adj, p)
    c = oracle_csr(adj)
    slots = st.read(M.BUF_EDGE_SLOTS).reshape(-1, 2)
    ij = st.read(M.BUF_EDGE_IJ).reshape(-1, 2)
    hub = ij[:, 0] == 0
    assert slots[hub, 0].max() == n - 2  # the hub's last slot
    # every slot points at its edge in the endpoint's CSR row
    for end in (0, 1):
        pos = c["off"][ij[:, end]] + slots[:, end]
        assert np.array_equal(c["edge_idx"][pos], np.arange(ij.shape[0]))
    _teacher_forced(ctx, embd, adj, p, M.EPOCH_EDGE_ATOMIC, epochs=4)
