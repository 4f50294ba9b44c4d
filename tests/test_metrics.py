"""Pin the test metrics against scipy / scikit-learn (CPU only)."""
import numpy as np
import pytest
from scipy.stats import spearmanr

from metrics import recall_at_k, spearman, trustworthiness


def test_spearman_matches_scipy():
    rng = np.random.default_rng(0)
    a = rng.normal(size=500)
    b = a + rng.normal(size=500)
    b[::10] = b[0]  # ties
    assert abs(spearman(a, b) - spearmanr(a, b).statistic) < 1e-12


def test_trustworthiness_matches_sklearn():
    sk = pytest.importorskip("sklearn.manifold")
    rng = np.random.default_rng(1)
    x = rng.normal(size=(400, 10))
    y = x[:, :2] + 0.3 * rng.normal(size=(400, 2))
    for k in (5, 15):
        assert abs(trustworthiness(x, y, k) - sk.trustworthiness(x, y, n_neighbors=k)) < 1e-10


def test_trustworthiness_identity_is_one():
    x = np.random.default_rng(2).normal(size=(200, 3))
    assert abs(trustworthiness(x, x, 10) - 1.0) < 1e-12


def test_recall():
    assert recall_at_k([[1, 2, 3]], [[3, 2, 1]]) == 1.0
    assert recall_at_k([[1, 2, 4]], [[3, 2, 1]]) == pytest.approx(2 / 3)
