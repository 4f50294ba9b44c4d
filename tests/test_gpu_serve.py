"""The fit server (mrs_server_*): continuous batching of uneven fits.

Bars:
  * node-exact requests return exactly what a single mrs_umap_gpu call
    returns (bit for bit), whether served alone or in a mix of other fits;
  * edge-parallel requests differ from the single fit only by FP32 atomic
    order, which Adam amplifies into a different arrangement inside clusters
    after a few epochs: the bar is the same quality -- trustworthiness@15
    within 0.01 of the single fit and the fraction of 10-NN in the embedding
    that share the point's cluster label within 0.02;
  * admission never holds more estimated bytes than the budget, and every
    request completes.
"""
import numpy as np
import pytest

import mrs_amd as M
from metrics import trustworthiness

pytestmark = pytest.mark.gpu


def mixture(n, d, n_clusters, seed):
    rng = np.random.default_rng(seed)
    centres = rng.normal(0, 5, size=(n_clusters, d))
    lab = rng.integers(0, n_clusters, size=n)
    return (centres[lab] + rng.normal(size=(n, d))).astype(np.float32), lab


def params(strategy, n_epochs=60, precision=M.KNN_FP32):
    p = M.UmapParamsGpu.default()
    p.epoch_strategy = strategy
    p.knn_precision = precision
    p.optim_params.n_epochs = n_epochs
    return p


def label_agreement(emb, lab, k=10):
    from scipy.spatial import cKDTree

    nb = cKDTree(emb).query(emb, k + 1)[1][:, 1:]
    return float(np.mean(lab[nb] == lab[:, None]))


SIZES = [(700, 32, 5), (2600, 50, 12), (1500, 100, 8), (4000, 32, 20), (900, 50, 6),
         (3100, 50, 9)]


def test_serve_node_exact_equals_single_fit_alone_and_mixed(ctx):
    data = [mixture(n, d, c, seed=i)[0] for i, (n, d, c) in enumerate(SIZES)]
    p = params(M.EPOCH_NODE_EXACT)
    single = [M.umap_gpu(ctx, x, p, seed=100 + i) for i, x in enumerate(data)]
    with M.Server(prep_workers=3, max_active=8) as sv:
        alone, _ = sv.wait(sv.submit(data[1], p, seed=101))
        tickets = [sv.submit(x, p, seed=100 + i) for i, x in enumerate(data)]
        mixed = [sv.wait(t)[0] for t in tickets]
        st = sv.stats()
    assert np.array_equal(alone, single[1])
    for i, (a, b) in enumerate(zip(mixed, single)):
        assert a.shape == b.shape and np.array_equal(a, b), f"request {i}"
    assert st["completed"] == len(data) + 1 and st["failed"] == 0
    assert st["peak_resident"] >= 2  # the mix really overlapped


def test_serve_edge_parallel_close_to_single_fit(ctx):
    x, lab = mixture(2000, 50, 10, seed=7)
    others = [mixture(n, d, c, seed=20 + i)[0] for i, (n, d, c) in enumerate(SIZES)]
    p = params(M.EPOCH_EDGE_ATOMIC, n_epochs=200)
    single = M.umap_gpu(ctx, x, p, seed=5).T
    with M.Server(prep_workers=4, max_active=16) as sv:
        tickets = [sv.submit(o, p, seed=30 + i) for i, o in enumerate(others)]
        t = sv.submit(x, p, seed=5)
        tickets += [sv.submit(o, p, seed=40 + i) for i, o in enumerate(others)]
        got = sv.wait(t)[0].T
        rest = [sv.wait(q)[0] for q in tickets]
    assert np.all(np.isfinite(got)) and all(np.all(np.isfinite(r)) for r in rest)
    t_single = trustworthiness(x, single, 15)
    t_served = trustworthiness(x, got, 15)
    a_single, a_served = label_agreement(single, lab), label_agreement(got, lab)
    print(f"trustworthiness single {t_single:.4f} served {t_served:.4f}; "
          f"10-NN label agreement single {a_single:.4f} served {a_served:.4f}")
    assert abs(t_served - t_single) <= 0.01
    assert abs(a_served - a_single) <= 0.02


def test_serve_budget_admission(ctx):
    data = [mixture(n, d, c, seed=50 + i)[0] for i, (n, d, c) in enumerate(SIZES)]
    big = max(M.fit_bytes_estimate(x.shape[0], x.shape[1]) for x in data)
    budget = int(1.5 * big)
    p = params(M.EPOCH_EDGE_ATOMIC, n_epochs=40, precision=M.KNN_BF16_RERANK)
    with M.Server(hbm_budget=budget, prep_workers=4, max_active=64) as sv:
        tickets = [sv.submit(x, p, seed=i) for i, x in enumerate(data)]
        outs = [sv.wait(t) for t in tickets]
        st = sv.stats()
    assert st["completed"] == len(data) and st["failed"] == 0
    assert st["budget_bytes"] == budget and st["peak_bytes"] <= budget
    for (o, times), x in zip(outs, data):
        assert o.shape == (2, x.shape[0]) and np.all(np.isfinite(o))
        assert times["latency_ms"] >= times["prepare_ms"] >= 0


def test_serve_errors(ctx):
    x, _ = mixture(500, 16, 4, seed=1)
    with M.Server(hbm_budget=M.fit_bytes_estimate(500, 16) + 1, prep_workers=1,
                  max_active=4) as sv:
        with pytest.raises(M.ManifoldsError) as e:
            sv.submit(mixture(5000, 64, 4, seed=2)[0], params(M.EPOCH_EDGE_ATOMIC))
        assert e.value.status == M.MRS_INVALID_ARGUMENT
        bad = params(M.EPOCH_EDGE_ATOMIC)
        bad.k = 600
        with pytest.raises(M.ManifoldsError) as e:
            sv.submit(x, bad)
        assert e.value.status == M.MRS_INVALID_ARGUMENT
        # a request that fails inside the server reports its status at wait()
        cpu_opt = params(M.EPOCH_EDGE_ATOMIC)
        cpu_opt.optimiser = b"sgd"
        t = sv.submit(x, cpu_opt)
        with pytest.raises(M.ManifoldsError) as e:
            sv.wait(t)
        assert e.value.status == M.MRS_INVALID_ARGUMENT
        # the server keeps serving after a failed request
        ok, _ = sv.wait(sv.submit(x, params(M.EPOCH_EDGE_ATOMIC, n_epochs=20)))
        assert np.all(np.isfinite(ok))
