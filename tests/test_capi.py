"""CPU checks of the C-ABI library: it loads, exports every declared symbol,
and its host-side logic (string keys, the negative hash) matches the oracle."""
import ctypes as C

import numpy as np
import pytest

import mrs_amd as M
import pyoracle as orc


def test_library_exports_every_declared_symbol():
    L = M.lib()
    names = M._declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_parse_string_keys():
    # parse_umap_optimiser_gpu (umap_optimiser_gpu.rs:44-52): case-insensitive
    L = M.lib()
    assert L.mrs_parse_umap_optimiser_gpu(b"adam_gpu") == 0
    assert L.mrs_parse_umap_optimiser_gpu(b"ADAM_GPU") == 0
    assert L.mrs_parse_umap_optimiser_gpu(b"adam_parallel") == 1
    assert L.mrs_parse_umap_optimiser_gpu(b"adam") == 2
    assert L.mrs_parse_umap_optimiser_gpu(b"sgd") == 3
    assert L.mrs_parse_umap_optimiser_gpu(b"nope") == -1
    # parse_ann_search_gpu (nearest_neighbours_gpu.rs:203-210)
    assert L.mrs_parse_ann_search_gpu(b"exhaustive_gpu") == 2
    assert L.mrs_parse_ann_search_gpu(b"Exhaustive") == 2
    assert L.mrs_parse_ann_search_gpu(b"ivf") == 0
    assert L.mrs_parse_ann_search_gpu(b"nndescent_gpu") == 1
    assert L.mrs_parse_ann_search_gpu(b"bogus") == -1


def test_hash_matches_oracle_including_fast_remainder():
    L = M.lib()
    rng = np.random.default_rng(0)
    ns = [1, 2, 3, 5, 7, 10, 641, 65537, 99_991, 100_000, 1_000_000, 2**31 - 1, 2**31,
          2**32 - 5, 2**32 - 1]
    for n in ns:
        for _ in range(400):
            a = [int(v) for v in rng.integers(0, 2**32, size=5, dtype=np.uint64)]
            assert L.mrs_hash_neg(*a, n) == orc.hash_neg(*a, n), (a, n)


def test_ctx_create_without_device_reports_device_error():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    h = C.c_void_p()
    rc = M.lib().mrs_ctx_create(0, 0, C.byref(h))
    assert rc == M.MRS_DEVICE_ERROR
    assert M.lib().mrs_last_error()


def test_fit_bytes_estimate_grows_with_every_size():
    base = M.fit_bytes_estimate(10_000, 50, 15, 2)
    assert base > 10_000 * 50 * 4  # at least the data itself
    assert M.fit_bytes_estimate(20_000, 50, 15, 2) > base
    assert M.fit_bytes_estimate(10_000, 100, 15, 2) > base
    assert M.fit_bytes_estimate(10_000, 50, 30, 2) > base
    assert M.fit_bytes_estimate(0, 50, 15, 2) == 0


def test_server_create_checks_arguments_and_fails_cleanly_without_a_device():
    # bad worker / slot counts are refused before anything starts (at most 8
    # prep workers); on a host with no HIP device creation reports a device
    # error and returns no handle
    for workers, slots in ((0, 8), (9, 8), (12, 8), (17, 8), (2, 0), (2, 5000)):
        h = C.c_void_p()
        assert M.lib().mrs_server_create(0, 0, workers, slots, C.byref(h)) == M.MRS_INVALID_ARGUMENT
        assert not h.value
    h = C.c_void_p()
    rc = M.lib().mrs_server_create(0, 0, 2, 8, C.byref(h))
    assert rc in (M.MRS_OK, M.MRS_DEVICE_ERROR)
    if rc == M.MRS_OK:  # a GPU is present after all
        assert M.lib().mrs_server_destroy(h) == M.MRS_OK
    else:
        assert not h.value and M.lib().mrs_last_error()


@pytest.mark.parametrize("m,seed", [(1, 0), (2, 1), (7, 2), (42, 3), (42, 4), (120, 5)])
def test_tridiag_eigen_matches_numpy(m, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=m)
    e = rng.normal(size=m - 1)
    if seed == 4:  # repeated eigenvalues and an exact split
        d[:] = 0.5
        e[::3] = 0.0
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    ev, V = M.tridiag_eigen(d, e)
    want = np.linalg.eigvalsh(T)
    assert np.all(np.diff(ev) >= 0)
    np.testing.assert_allclose(ev, want, atol=1e-12 * max(1.0, np.abs(want).max()))
    np.testing.assert_allclose(V.T @ V, np.eye(m), atol=1e-12)
    np.testing.assert_allclose(T @ V, V * ev[None, :], atol=1e-11)


def test_autotune_arguments_and_empty_table(tmp_path):
    # host-side checks of the autotuner's ABI (no device work)
    M.autotune_clear()
    assert M.autotune_records() == []
    for fam, var in ((M.TUNE_EPOCH_SPAN, 96), (M.TUNE_KNN_SPLITS, 3), (7, 64)):
        with pytest.raises(M.ManifoldsError) as e:
            M.autotune_force(fam, var)
        assert e.value.status == M.MRS_INVALID_ARGUMENT
    with pytest.raises(M.ManifoldsError):
        M.autotune_set_mode(2)
    M.autotune_force(M.TUNE_EPOCH_SPAN, 256)
    M.autotune_force(M.TUNE_EPOCH_SPAN, -1)
    bad = tmp_path / "t.txt"
    bad.write_text("not a table\n")
    with pytest.raises(M.ManifoldsError):
        M.autotune_load(bad)
    with pytest.raises(M.ManifoldsError):
        M.autotune_load(tmp_path / "missing.txt")


@pytest.mark.parametrize("d,n_comp", [(1, 1), (2, 2), (5, 3), (64, 4), (200, 16)])
def test_host_symmetric_eigensolver_matches_numpy(d, n_comp):
    # mrs_sym_eigen_leading (Householder tridiagonalisation + implicit QL, the
    # PCA layout's host step) against numpy's eigh: eigenvalues, and
    # eigenvectors under the library's sign rule (largest-magnitude entry
    # positive)
    rng = np.random.default_rng(d)
    q = np.linalg.qr(rng.normal(size=(d, d)))[0]
    lam = np.sort(rng.uniform(0.1, 10.0, size=d))[::-1] * np.linspace(1.0, 2.0, d)
    a = (q * lam) @ q.T
    a = (a + a.T) / 2
    vecs = np.zeros((n_comp, d))
    evals = np.zeros(n_comp)
    P = C.POINTER(C.c_double)
    st = M.lib().mrs_sym_eigen_leading(np.ascontiguousarray(a).ctypes.data_as(P), d, n_comp,
                                       vecs.ctypes.data_as(P), evals.ctypes.data_as(P))
    assert st == 0
    w, v = np.linalg.eigh(a)
    for c in range(n_comp):
        r = d - 1 - c
        assert abs(evals[c] - w[r]) < 1e-10 * w[-1]
        want = v[:, r] * (1.0 if v[np.argmax(np.abs(v[:, r])), r] >= 0 else -1.0)
        assert np.max(np.abs(vecs[c] - want)) < 1e-8
