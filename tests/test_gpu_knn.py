"""GPU parity of the exhaustive kNN (MFMA shortlist + exact FP32 re-rank)
against the CPU oracle's brute force.

Bar: indices bit-exact (ties by index) and distances within 1e-5 relative of
the FP32 oracle (the re-rank uses the oracle's own fmaf order, so they are in
fact equal); the BF16 path is held to the same bar on these inputs.
"""
import numpy as np
import pytest

import mrs_amd as M
import pyoracle as orc
from helpers import blobs

pytestmark = pytest.mark.gpu


def _check(ctx, x, k, precision, shortlist=0):
    idx, dist = M.knn_exhaustive(ctx, x, k, precision, shortlist)
    widx, wdist = orc.knn_exact(x, k)
    bad = np.nonzero(np.any(idx != widx, axis=1))[0]
    assert bad.size == 0, f"{bad.size} rows differ, first {bad[:5]}: {idx[bad[0]]} vs {widx[bad[0]]}"
    np.testing.assert_allclose(dist, wdist, rtol=1e-5, atol=0)
    assert np.all(idx != np.arange(x.shape[0])[:, None])
    assert np.all(np.diff(dist, axis=1) >= 0)
    return idx, dist


@pytest.mark.parametrize("precision", [M.KNN_FP32, M.KNN_BF16_RERANK])
@pytest.mark.parametrize("n,d,k", [(3000, 50, 15), (2500, 64, 15), (1500, 8, 15),
                                   (1200, 100, 15), (1000, 33, 30)])
def test_knn_matches_oracle(ctx, precision, n, d, k):
    x, _ = blobs(n, d, 7, seed=n + d)
    _check(ctx, x, k, precision, shortlist=4 * k)


@pytest.mark.parametrize("precision", [M.KNN_FP32, M.KNN_BF16_RERANK])
def test_knn_duplicates_ties_by_index(ctx, precision):
    x, _ = blobs(1500, 16, 4, seed=2)
    x[10:40] = x[10]  # 30 identical points: ties at distance 0
    x[500:503] = x[700]
    _check(ctx, x, 15, precision)


@pytest.mark.parametrize("precision", [M.KNN_FP32, M.KNN_BF16_RERANK])
def test_knn_tiny_inputs(ctx, precision):
    rng = np.random.default_rng(0)
    for n, k in [(2, 1), (17, 15), (70, 15), (300, 30)]:
        x = rng.normal(size=(n, 5)).astype(np.float32)
        _check(ctx, x, k, precision)


def test_knn_invalid_arguments(ctx):
    x = np.zeros((10, 4), np.float32)
    with pytest.raises(M.ManifoldsError) as e:
        M.knn_exhaustive(ctx, x, 10)
    assert e.value.status == M.MRS_INVALID_ARGUMENT
    with pytest.raises(M.ManifoldsError) as e:
        M.knn_exhaustive(ctx, np.zeros((1, 4), np.float32), 1)
    assert e.value.status == M.MRS_NO_DATA


def test_knn_device_tensors(ctx):
    import torch

    x, _ = blobs(2000, 32, 5, seed=4)
    xt = torch.from_numpy(x).cuda()
    idx, dist = M.knn_exhaustive(ctx, xt, 15, M.KNN_BF16_RERANK, 60)
    torch.cuda.synchronize()
    widx, wdist = orc.knn_exact(x, 15)
    assert np.array_equal(idx.cpu().numpy(), widx)
    np.testing.assert_allclose(dist.cpu().numpy(), wdist, rtol=1e-5)
