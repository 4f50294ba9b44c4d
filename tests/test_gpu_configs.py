"""The benchmark configurations at their own sizes on the GPU, against the
oracle and the CPU optimisers:

* plain FP32 form: 100 000 x 50, 20 anisotropic Gaussians + 5 % background,
  FP32 kNN k = 15, 500 Adam epochs: trustworthiness@15 of the GPU fit within
  0.005 (either side) of the CPU restatement of "adam_gpu" (identical hashed
  negatives) and of the CPU parallel Adam, both from the fit's own graph and
  spectral layout (5 000-point sample);
* precision form: 1 000 000 x 64 (50 Gaussians on an 8-D subspace): the BF16
  shortlist + exact FP32 re-rank kNN has recall@15 = 1.0 against the FP64
  brute force on 10 000 seeded rows, and its distances are within 1e-5
  relative of FP64;
* densMAP: 500 000 x 50 (30 Gaussians, sd log-uniform in [0.2, 5]), k = 30: on
  a 20 000-point draw the device DensState, radii, weights and density
  gradient against the oracle; the full fit finite with a local-radius
  Spearman correlation >= 0.95.
"""
import numpy as np
import pytest

import mrs_amd as M
import pyoracle as orc
import synthetic
from helpers import assert_grad_close, coo_to_adjacency
from metrics import recall_at_k, spearman, trustworthiness

pytestmark = pytest.mark.gpu


def _sample_trustworthiness(x, emb, m=5000, k=15, seed=11):
    rows = np.sort(np.random.default_rng(seed).choice(x.shape[0], size=m, replace=False))
    return trustworthiness(x[rows], np.asarray(emb)[rows], k)


def test_plain_fp32_form_fit_matches_cpu_parallel_adam(ctx):
    n, k, n_epochs = 100_000, 15, 500
    x, _ = synthetic.anisotropic(n=n, d=50)
    p = M.UmapParamsGpu.default()
    p.knn_precision = M.KNN_FP32
    p.optim_params.n_epochs = n_epochs
    e_gpu = M.umap_gpu(ctx, x, p, seed=42).T
    assert np.all(np.isfinite(e_gpu))
    idx, dist = M.knn_exhaustive(ctx, x, k, M.KNN_FP32)
    g = M.FuzzyGraph(ctx, idx, dist, n_epochs=n_epochs)
    init = g.spectral_layout(2, seed=42, data=x)
    q = M.UmapOptimParams.default_2d()
    q.n_epochs = n_epochs
    adj = coo_to_adjacency(n, *g.coo())
    e_cpu = orc.optimise_adam_parallel(init, *adj, q, 42)
    e_ref = orc.optimise_adam_gpu(init, *adj, q, 42)
    t_gpu, t_cpu = _sample_trustworthiness(x, e_gpu), _sample_trustworthiness(x, e_cpu)
    t_ref = _sample_trustworthiness(x, e_ref)
    print(f"trustworthiness@15 (5 000-point sample): GPU {t_gpu:.4f}, CPU adam_gpu restated "
          f"{t_ref:.4f}, CPU parallel Adam {t_cpu:.4f}")
    assert abs(t_gpu - t_ref) <= 0.005, (t_gpu, t_ref)
    assert abs(t_gpu - t_cpu) <= 0.005, (t_gpu, t_cpu)


def test_precision_form_recall_on_10k_rows(ctx):
    x, _ = synthetic.subspace()  # 1 000 000 x 64, seed 42
    idx, dist, t = M.knn_exhaustive(ctx, x, 15, M.KNN_BF16_RERANK, 60, timed=True)
    rows = np.sort(np.random.default_rng(42).choice(x.shape[0], 10_000, replace=False))
    widx, wdist = orc.knn_exact_rows_f64(x, 15, rows)
    rec = recall_at_k(idx[rows], widx)
    rel = float(np.max(np.abs(dist[rows].astype(np.float64) - wdist) / wdist))
    exact_rows = int(np.count_nonzero(np.all(idx[rows] == widx, axis=1)))
    print(f"recall@15 {rec}, rows identical {exact_rows}/10000, max rel dist err {rel:.2e}, "
          f"swept {t['swept']:.3f}, split work items {t['aux_items']}")
    assert rec == 1.0
    assert exact_rows == rows.size  # the same lists, in the same order (ties by index)
    assert rel < 1e-5


def test_densmap_form_sample_against_oracle(ctx):
    # a 20 000-point draw of the densMAP configuration's distribution, k = 30
    n, k, n_epochs = 20_000, 30, 500
    x, _ = synthetic.uneven_density(n=n, seed=7)
    idx, dist = M.knn_exhaustive(ctx, x, k, M.KNN_BF16_RERANK)
    widx, wdist = orc.knn_exact(x, k, 0, 500)
    assert np.array_equal(idx[:500], widx)
    g = M.FuzzyGraph(ctx, idx, dist, n_epochs=n_epochs)
    rows, cols, vals = g.coo()
    # DensState::new on the device against the oracle on the same graph and kNN
    r_dev, ms_dev, tot_dev = g.dens_state(idx, dist)
    st, r_orc, ms_orc, tot_orc = orc.dens_state(n, rows, cols, vals, idx, dist)
    assert st == 0
    np.testing.assert_allclose(r_dev, r_orc, rtol=0, atol=1e-5)
    np.testing.assert_allclose(ms_dev, ms_orc, rtol=1e-6)
    assert abs(tot_dev - tot_orc) <= 1e-6 * tot_orc
    # one density epoch from the spectral layout: radii, weights, coefficient,
    # and the gradient with the density term, against the oracle
    adj = coo_to_adjacency(n, rows, cols, vals)
    init = g.spectral_layout(2, seed=42, data=x)
    dp = M.DensParams.densmap_default()
    ds = M.DensState(params=dp, r=r_dev.ctypes.data, mu_sum=ms_dev.ctypes.data, mu_tot=tot_dev)
    p = M.UmapOptimParams.default_2d()
    p.n_epochs = n_epochs
    s = M.UmapGpuState(ctx, init, *adj, p, ds)
    _, c = orc.csr_from_graph(*adj)
    coeff = s.launch_dens_update()
    re, phi = orc.dens_radii(init, c, p.a, p.b)
    np.testing.assert_allclose(s.read(M.BUF_DENS_PHI), phi, rtol=1e-5, atol=0)
    np.testing.assert_allclose(s.read(M.BUF_DENS_RE), re, rtol=0, atol=1e-5)
    w, coeff_w = orc.dens_weights(r_dev, re, dp.var_shift, dp.lambda_, tot_dev)
    assert abs(coeff - coeff_w) <= 1e-5 * abs(coeff_w)
    np.testing.assert_allclose(s.read(M.BUF_DENS_W), w, rtol=0, atol=1e-4)
    epoch = 400  # a density epoch: (401 / 500) > 1 - 0.3
    nxt = s.read(M.BUF_NEXT)
    want, _, absum = orc.grad_accum(init, c, nxt, epoch, 42, 5, p,
                                    dict(re=re, phi_sum=phi, weight=w, coeff=coeff_w),
                                    with_abs=True, exact=True)
    s.launch_grad_accum(epoch, 42, coeff)
    assert_grad_close(s.read(M.BUF_GRAD), want, absum, 1e-4, "density gradient, node kernel")
    s.launch_epoch_edges(epoch, 42, coeff)
    assert_grad_close(s.read(M.BUF_GRAD), want, absum, 1e-4, "density gradient, edge kernel")


def test_densmap_form_full_fit_properties(ctx):
    n, k = 500_000, 30
    x, _ = synthetic.uneven_density(n=n)
    p = M.UmapParamsGpu.default()
    p.k = k
    p.knn_precision = M.KNN_BF16_RERANK
    out = M.umap_gpu(ctx, x, p, seed=42, dens=M.DensParams.densmap_default())
    emb = out.T
    assert emb.shape == (n, 2) and np.all(np.isfinite(emb))
    idx, dist = M.knn_exhaustive(ctx, x, k, M.KNN_BF16_RERANK)
    g = M.FuzzyGraph(ctx, idx, dist, n_epochs=p.optim_params.n_epochs)
    r_o, _, _ = g.dens_state(idx, dist)
    rows, cols, _ = g.coo()
    e = emb.astype(np.float64)
    d2 = ((e[rows] - e[cols]) ** 2).sum(1)
    phi = 1.0 / (1.0 + p.optim_params.a * d2 ** p.optim_params.b)
    num = np.bincount(rows, weights=phi * d2, minlength=n)
    den = np.maximum(np.bincount(rows, weights=phi, minlength=n), np.finfo(np.float32).eps)
    r_e = np.log(1e-8 + num / den)
    rho = spearman(r_o, r_e)
    print(f"local-radius Spearman(R_o, R_e) at n = {n}: {rho:.4f}")
    assert rho >= 0.95
