"""The launch-shape autotuner: every variant it may choose gives the result the
oracle checks, and tuning on first use records a decision that later calls
use.

- Edge-epoch spans (64, 128, 256 edges per wavefront) differ only in the
  order of the FP32 atomic additions: each forced span passes the per-node
  teacher-forced comparison with the oracle (gradient bar 1e-4 of the sum of
  term magnitudes, schedule exact).
- kNN database splits (1, 2, 4, 8) give bit-identical lists, equal to the
  oracle's brute force.
"""
import numpy as np
import pytest

import mrs_amd as M
import pyoracle as orc
from helpers import blobs, umap_graph_from_data
from test_gpu_epoch import _teacher_forced

pytestmark = pytest.mark.gpu


@pytest.fixture
def clean_tuner():
    M.autotune_clear()
    yield
    M.autotune_force(M.TUNE_EPOCH_SPAN, -1)
    M.autotune_force(M.TUNE_KNN_SPLITS, -1)
    M.autotune_set_mode(M.TUNE_ON_FIRST_USE)
    M.autotune_clear()


def _graph(n=3000, seed=21):
    x, _ = blobs(n, 8, 6, seed=seed)
    _, _, _, adj = umap_graph_from_data(x, 15)
    embd = np.random.default_rng(seed).uniform(-10, 10, size=(n, 2)).astype(np.float32)
    return embd, adj


@pytest.mark.parametrize("span", [64, 128, 256])
def test_every_epoch_span_matches_oracle(ctx, clean_tuner, span):
    M.autotune_force(M.TUNE_EPOCH_SPAN, span)
    embd, adj = _graph()
    p = M.UmapOptimParams.default_2d()
    p.n_epochs = 30
    _teacher_forced(ctx, embd, adj, p, M.EPOCH_EDGE_ATOMIC, epochs=5)


def test_epoch_span_tuned_on_first_use(ctx, clean_tuner):
    embd, adj = _graph(seed=22)
    p = M.UmapOptimParams.default_2d()
    p.n_epochs = 40
    st = M.UmapGpuState(ctx, embd, *adj, p)
    assert not [r for r in M.autotune_records() if r["family"] == M.TUNE_EPOCH_SPAN]
    st.run_epochs(0, 20, 42)  # 12 tuning epochs (spans alternating), then 8 with the choice
    recs = [r for r in M.autotune_records() if r["family"] == M.TUNE_EPOCH_SPAN]
    assert len(recs) == 1
    r = recs[0]
    assert sorted(r["ms"]) == [64, 128, 256]
    assert all(t > 0 for t in r["ms"].values())
    assert r["variant"] == min(r["ms"], key=r["ms"].get)
    assert r["shape"] & 0xF == 2  # dimension in the shape class
    # the rest of the run uses the record: no new decision, a sane layout
    st.run_epochs(20, 20, 42)
    assert len(M.autotune_records()) == 1
    out = st.read(M.BUF_EMBD)
    assert np.all(np.isfinite(out))
    # the tuned run is a valid run: its schedule equals the oracle's replay
    c = orc.csr_from_graph(*adj)[1]
    nx = c["per"].copy()
    for e in range(40):
        nx = orc.schedule_update(c["per"], nx, e)
    assert np.array_equal(st.read(M.BUF_NEXT), nx)


@pytest.mark.parametrize("precision", [M.KNN_FP32, M.KNN_BF16_RERANK])
def test_every_knn_split_count_bit_identical(ctx, clean_tuner, precision):
    x, _ = blobs(9000, 24, 8, seed=31)
    got = {}
    for sp in (1, 2, 4, 8):
        M.autotune_force(M.TUNE_KNN_SPLITS, sp)
        got[sp] = M.knn_exhaustive(ctx, x, 15, precision)
    for sp in (2, 4, 8):
        assert np.array_equal(got[sp][0], got[1][0]), sp
        assert np.array_equal(got[sp][1].view(np.uint32), got[1][1].view(np.uint32)), sp
    idx, _ = orc.knn_exact(x, 15, 0, 600)
    assert np.array_equal(got[1][0][:600], idx)


def test_knn_splits_tuned_on_first_use(ctx, clean_tuner):
    x, _ = blobs(12000, 32, 8, seed=32)
    M.autotune_force(M.TUNE_KNN_SPLITS, 1)
    want = M.knn_exhaustive(ctx, x, 15, M.KNN_BF16_RERANK)
    M.autotune_force(M.TUNE_KNN_SPLITS, -1)
    first = M.knn_exhaustive(ctx, x, 15, M.KNN_BF16_RERANK)  # tunes
    recs = [r for r in M.autotune_records() if r["family"] == M.TUNE_KNN_SPLITS]
    assert len(recs) == 1 and len(recs[0]["ms"]) >= 3
    assert recs[0]["variant"] == min(recs[0]["ms"], key=recs[0]["ms"].get)
    again = M.knn_exhaustive(ctx, x, 15, M.KNN_BF16_RERANK)  # uses the record
    assert len(M.autotune_records()) == 1
    for got in (first, again):
        assert np.array_equal(got[0], want[0])
        assert np.array_equal(got[1].view(np.uint32), want[1].view(np.uint32))


def test_tuning_off_records_nothing(ctx, clean_tuner):
    M.autotune_set_mode(M.TUNE_OFF)
    x, _ = blobs(5000, 16, 5, seed=33)
    M.knn_exhaustive(ctx, x, 15)
    embd, adj = _graph(n=1500, seed=23)
    p = M.UmapOptimParams.default_2d()
    p.n_epochs = 20
    M.optimise_embedding_adam_gpu(ctx, embd, *adj, p, seed=42)
    assert M.autotune_records() == []


def test_records_save_and_load(ctx, clean_tuner, tmp_path):
    x, _ = blobs(6000, 16, 5, seed=34)
    M.knn_exhaustive(ctx, x, 15)
    recs = M.autotune_records()
    assert recs
    path = tmp_path / "tune.txt"
    M.autotune_save(path)
    M.autotune_clear()
    assert M.autotune_records() == []
    M.autotune_load(path)
    assert M.autotune_records() == recs
    bad = tmp_path / "other.txt"
    bad.write_text("mrs-autotune 1 gfx000:1\n2 5 4 1:1.0\n")
    with pytest.raises(M.ManifoldsError):
        M.autotune_load(bad)
