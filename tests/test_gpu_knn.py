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


# ---- pruned sweep (n >= 65536): rows in cell order, stages ruled out by bounds


def _both_sweeps(ctx, x, k, precision, shortlist):
    """(idx, dist, timings) with pruning, and (idx, dist) of the full sweep in
    row order; the full sweep in cell order (every stage, own cell first) is
    checked against the row-order one bit for bit on the way."""
    got = M.knn_exhaustive(ctx, x, k, precision, shortlist, timed=True)
    try:
        ctx.set_knn_pruning(-1)
        full = M.knn_exhaustive(ctx, x, k, precision, shortlist, timed=True)
        ctx.set_knn_pruning(0)
        cfull = M.knn_exhaustive(ctx, x, k, precision, shortlist, timed=True)
    finally:
        ctx.set_knn_pruning(True)
    assert full[2]["swept"] == 1.0 and full[2]["cells"] == 0
    # every stage that holds a row (cells are padded to whole blocks: stages of
    # padding alone are skipped, so the fraction of padded pairs is below 1)
    assert cfull[2]["cells"] > 0 and cfull[2]["swept"] >= got[2]["swept"]
    assert np.array_equal(cfull[0], full[0])
    assert np.array_equal(cfull[1].view(np.uint32), full[1].view(np.uint32))
    return got, full


def _check_pruned(ctx, x, k, precision, shortlist=0):
    (idx, dist, t), (fidx, fdist, _) = _both_sweeps(ctx, x, k, precision, shortlist)
    assert t["cells"] > 0, "the pruned sweep did not run"
    # the same lists as the full sweep, so the same result, bit for bit
    bad = np.nonzero(np.any(idx != fidx, axis=1))[0]
    assert bad.size == 0, f"{bad.size} rows differ from the full sweep, first {bad[:5]}"
    assert np.array_equal(dist.view(np.uint32), fdist.view(np.uint32))
    # and the oracle's brute force on slices of rows (ties by index)
    n = x.shape[0]
    for q0, q1 in ((0, 300), (n // 2, n // 2 + 300), (n - 300, n)):
        widx, wdist = orc.knn_exact(x, k, q0, q1)
        assert np.array_equal(idx[q0:q1], widx), (q0, q1)
        np.testing.assert_allclose(dist[q0:q1], wdist, rtol=1e-5, atol=0)
    return t


@pytest.mark.parametrize("precision", [M.KNN_BF16_RERANK, M.KNN_FP32])
def test_pruned_sweep_clustered(ctx, precision):
    import synthetic

    x, _ = synthetic.subspace(n=70_000, d=64, seed=3)
    t = _check_pruned(ctx, x, 15, precision, shortlist=60)
    print(f"cells {t['cells']}, swept {t['swept']:.3f}")
    assert t["swept"] < 0.9, t  # the bounds rule out stages on clustered data


def test_pruned_sweep_wide_shortlist(ctx):
    import synthetic

    x, _ = synthetic.subspace(n=68_000, d=50, n_clusters=30, seed=9)
    _check_pruned(ctx, x, 30, M.KNN_BF16_RERANK, shortlist=120)  # 128 candidates


def test_pruned_sweep_uniform(ctx):
    # no structure: few stages can be ruled out, the result is still exact
    x = np.random.default_rng(5).uniform(-1, 1, size=(66_000, 8)).astype(np.float32)
    _check_pruned(ctx, x, 15, M.KNN_BF16_RERANK, shortlist=60)


@pytest.mark.parametrize("precision", [M.KNN_BF16_RERANK, M.KNN_FP32])
def test_pruned_sweep_lattice_ties(ctx, precision):
    # integer lattice points: exact duplicates and many equal distances, so
    # the lists are decided by row order at every threshold
    rng = np.random.default_rng(11)
    x = np.zeros((68_000, 16), np.float32)
    x[:, :3] = rng.integers(0, 30, size=(68_000, 3))
    _check_pruned(ctx, x, 15, precision, shortlist=60)


def test_pruned_sweep_far_from_origin(ctx):
    # Clustered data under a common offset of 400 in every coordinate: |x|^2
    # dwarfs the neighbour distances, so the pass-1 floors (scores at a squared
    # distance d2, computed as |x|^2 - 2 score) would lose their meaning to
    # cancellation and are not used for most queries.  BF16 cannot resolve the
    # neighbours at this offset, so only the sweeps are compared, bit for bit:
    # pruned, full in cell order and full in row order.
    import synthetic

    x, _ = synthetic.subspace(n=66_000, d=32, n_clusters=30, seed=21)
    x = (x + 400.0).astype(np.float32)
    (idx, dist, t), (fidx, fdist, _) = _both_sweeps(ctx, x, 15, M.KNN_BF16_RERANK, 60)
    assert t["cells"] > 0
    assert np.array_equal(idx, fidx)
    assert np.array_equal(dist.view(np.uint32), fdist.view(np.uint32))


def test_pruned_sweep_wide_rows(ctx):
    # d = 100: 64-row stages (separate stage and block bounds) and the 7-step
    # BF16 kernel
    import synthetic

    x, _ = synthetic.subspace(n=66_000, d=100, n_clusters=40, seed=13)
    _check_pruned(ctx, x, 15, M.KNN_BF16_RERANK, shortlist=60)


def test_pruned_sweep_two_dimensional(ctx):
    # d = 2, dense 2-D blobs: neighbours lie ~0.05 apart at coordinates up to
    # 50, below BF16's resolution there (8 significant bits: steps of 0.25), so
    # this runs in FP32, whose shortlist is exact; the sweep is the same code
    rng = np.random.default_rng(17)
    centres = rng.uniform(-50, 50, size=(40, 2))
    x = (centres[rng.integers(0, 40, size=67_000)] + rng.normal(size=(67_000, 2))).astype(np.float32)
    t = _check_pruned(ctx, x, 15, M.KNN_FP32)
    assert t["swept"] < 0.5, t


@pytest.mark.parametrize("k,shortlist", [(15, 60), (30, 120)])
def test_pruned_sweep_split_work_items(ctx, k, shortlist):
    # Second-pass stage lists longer than the work-item chunk are cut into
    # parts whose aux lists k_merge_aux folds back into the block's list (with
    # the main list's S-th score as a floor).  By default only large problems
    # split (1M rows: lists of up to 8 309 stages); a 64-stage chunk makes the
    # 70k-row lists (~500 stages) split too.  Shortlists of 64 (k = 15) and
    # 128 (k = 30) candidates.  Bit-identical to the full sweep and the oracle.
    import synthetic

    x, _ = synthetic.subspace(n=70_000, d=64, seed=3)
    ctx.set_knn_work_chunk(64)
    try:
        t = _check_pruned(ctx, x, k, M.KNN_BF16_RERANK, shortlist=shortlist)
    finally:
        ctx.set_knn_work_chunk(0)
    print(f"k {k}: {t['aux_items']} split work items, longest list {t['list_max']}")
    assert t["aux_items"] > 0, t
    _, _, t0 = M.knn_exhaustive(ctx, x, k, M.KNN_BF16_RERANK, shortlist, timed=True)
    assert t0["aux_items"] == 0  # the default chunk does not split at this size


def test_plain_fp32_form_knn_anisotropic_100k(ctx):
    # The plain FP32 GPU form's data: 100 000 x 50, 20 anisotropic Gaussians
    # with 5 % uniform background.  FP32 kNN (k = 15) bit-exact against the
    # oracle's brute force on row slices and on 1 000 seeded rows.
    import synthetic

    x, _ = synthetic.anisotropic()
    idx, dist, t = M.knn_exhaustive(ctx, x, 15, M.KNN_FP32, timed=True)
    n = x.shape[0]
    for q0, q1 in ((0, 256), (n // 2, n // 2 + 256), (n - 256, n)):
        widx, wdist = orc.knn_exact(x, 15, q0, q1)
        assert np.array_equal(idx[q0:q1], widx), (q0, q1)
        assert np.array_equal(dist[q0:q1].view(np.uint32), wdist.view(np.uint32)), (q0, q1)
    rows = np.sort(np.random.default_rng(2).choice(n, 1000, replace=False))
    widx, wdist = orc.knn_exact_rows_f64(x, 15, rows)
    assert np.array_equal(idx[rows], widx), "FP32 kNN vs FP64 brute force"
    assert np.max(np.abs(dist[rows] - wdist) / wdist) < 1e-5
    print(f"swept {t['swept']:.3f}, cells {t['cells']}")
