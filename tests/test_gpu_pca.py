"""GPU parity of the PCA initial layout (pca_layout, src/data/init.rs:1065-1140)
against the oracle's restatement (cyclic Jacobi on the centred covariance,
pinned to numpy's SVD and the reference's own known answers in
test_oracle.py).  The device takes another route to the same numbers (tiled
FP64 covariance, Householder + QL on the host, FP64 scores), so the bar is on
the unit-std coordinates: 2e-5 absolute.
"""
import numpy as np
import pytest

import mrs_amd as M
import pyoracle as orc
from helpers import blobs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,d,n_comp", [(20_000, 50, 2), (20_000, 50, 3), (7_000, 130, 5),
                                        (3_001, 32, 16), (40, 3, 2)])
def test_pca_layout_matches_oracle(ctx, n, d, n_comp):
    rng = np.random.default_rng(n + d)
    x, _ = blobs(n, d, 12, seed=n + d)
    x = (x * rng.uniform(0.5, 2.0, size=d)).astype(np.float32)  # unequal column scales
    got = M.pca_layout(ctx, x, n_comp).astype(np.float64)
    want = orc.pca_layout(x, n_comp).astype(np.float64)
    assert got.shape == (n, n_comp)
    err = np.abs(got - want).max(axis=0)
    assert np.all(err < 2e-5), err
    assert np.all(np.abs(got.mean(0)) < 1e-5)
    assert np.all(np.abs(got.std(0) - 1.0) < 1e-5)


def test_pca_layout_range_device_input_and_constant_column(ctx):
    import torch

    x, _ = blobs(5000, 20, 6, seed=5)
    x[:, 3] = 7.0  # a constant feature changes nothing
    want = orc.pca_layout(x, 2, range_=0.25).astype(np.float64)
    got = M.pca_layout(ctx, x, 2, range_=0.25).astype(np.float64)
    assert np.max(np.abs(got - want)) < 1e-5
    # constant data: every score is zero and keeps scale 1 (std <= 1e-8)
    z = np.full((100, 4), 3.0, np.float32)
    assert np.array_equal(M.pca_layout(ctx, z, 2), np.zeros((100, 2), np.float32))
    assert np.array_equal(orc.pca_layout(z, 2), np.zeros((100, 2), np.float32))
    del torch


def test_pca_layout_rejects_bad_arguments(ctx):
    x = np.ones((10, 3), np.float32)
    for nc in (0, 4, 17):
        with pytest.raises(M.ManifoldsError) as e:
            M.pca_layout(ctx, x, nc)
        assert e.value.status == M.MRS_INVALID_ARGUMENT
    with pytest.raises(M.ManifoldsError) as e:
        M.pca_layout(ctx, np.zeros((0, 3), np.float32), 2)
    assert e.value.status == M.MRS_NO_DATA


def test_fit_with_pca_init_and_unknown_strings_fall_back_to_pca(ctx):
    # parse_initilisation (init.rs:75-85) knows "spectral", "random", "pca";
    # umap_gpu_inner maps any other string to PCA (lib.rs:3526-3540).  With
    # the deterministic node-exact epochs the two fits are bit-identical, and
    # the fits differ from the spectral-initialised one.
    x, lab = blobs(3000, 24, 5, seed=9)
    p = M.UmapParamsGpu.default()
    p.epoch_strategy = M.EPOCH_NODE_EXACT
    p.optim_params.n_epochs = 30
    p.initialisation = b"pca"
    a = M.umap_gpu(ctx, x, p, seed=3)
    p.initialisation = b"no-such-init"
    b = M.umap_gpu(ctx, x, p, seed=3)
    assert np.array_equal(a, b)
    assert np.all(np.isfinite(a))
    p.initialisation = b"spectral"
    c = M.umap_gpu(ctx, x, p, seed=3)
    assert not np.array_equal(a, c)
