"""mrs_knn_exhaustive_noalloc: the exhaustive kNN with no allocation, working in
one per-GPU scratch array.  Called through the C ABI on device tensors; the
results must equal mrs_knn_exhaustive's bit for bit (its own parity with the
oracle is tested elsewhere) on two calls in a row, the second smaller (the
scratch keeps the first call's contents), in FP32 and BF16 + re-rank; sizes
beyond the stated maximum and host buffers are refused with InvalidArgument.
"""
import numpy as np
import pytest

import mrs_amd as M
import pyoracle as orc
from helpers import blobs

pytestmark = pytest.mark.gpu


def _noalloc(ctx, x_np, k, precision, shortlist):
    import torch

    x = torch.from_numpy(x_np).cuda()
    n = x.shape[0]
    idx = torch.full((n, k), -7, dtype=torch.int32, device="cuda")
    dist = torch.full((n, k), -7.0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    M.knn_exhaustive_noalloc(ctx, x, idx, dist, k, precision, shortlist)
    ctx.synchronize()
    return idx.cpu().numpy(), dist.cpu().numpy()


@pytest.mark.parametrize("precision", [M.KNN_FP32, M.KNN_BF16_RERANK])
def test_noalloc_knn_equals_allocating_knn_twice(ctx, precision):
    cases = [(16384, 64, 15, 60, 1), (3000, 50, 30, 0, 2), (17, 5, 15, 0, 3)]
    for n, d, k, sl, seed in cases:  # the largest first: later calls reuse a dirty scratch
        x, _ = blobs(n, d, 9, seed=seed)
        got_i, got_d = _noalloc(ctx, x, k, precision, sl)
        want_i, want_d = M.knn_exhaustive(ctx, x, k, precision, sl)
        assert np.array_equal(got_i, want_i), (n, d, k)
        assert np.array_equal(got_d.view(np.uint32), want_d.view(np.uint32)), (n, d, k)
    widx, _ = orc.knn_exact(x, 15)
    assert np.array_equal(got_i, widx)  # and the oracle on the last (tiny) case


def test_noalloc_knn_refuses_what_it_cannot_do(ctx):
    import torch

    x = torch.zeros((16385, 8), dtype=torch.float32, device="cuda")
    idx = torch.zeros((16385, 15), dtype=torch.int32, device="cuda")
    dist = torch.zeros((16385, 15), dtype=torch.float32, device="cuda")
    with pytest.raises(M.ManifoldsError) as e:
        M.knn_exhaustive_noalloc(ctx, x, idx, dist, 15)
    assert e.value.status == M.MRS_INVALID_ARGUMENT and "16384" in str(e.value)
    with pytest.raises(M.ManifoldsError) as e:
        M.knn_exhaustive_noalloc(ctx, x[:100], idx[:100], dist[:100], 15, M.KNN_FP32, 128)
    assert e.value.status == M.MRS_INVALID_ARGUMENT
    xh = np.zeros((100, 8), np.float32)
    with pytest.raises(M.ManifoldsError) as e:
        M.knn_exhaustive_noalloc(ctx, xh, np.zeros((100, 15), np.int32),
                                 np.zeros((100, 15), np.float32), 15)
    assert e.value.status == M.MRS_INVALID_ARGUMENT and "device" in str(e.value)
