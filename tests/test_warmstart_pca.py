"""CPU checks of the learned warm start's principal-component step
(markov-zip1_amd/warmstart.py, WarmStart._pca):

* with more rows than components it returns the leading right singular
  vectors (checked against numpy's SVD of the centred data: same span, and
  the projections reproduce the data exactly when k covers the rank);
* with fewer rows than components, the directions beyond the data's rank are
  zero columns, so a new row projects to exactly 0 on them.  Without that,
  those directions span rounding noise: the training rows project to ~1e-15,
  their standard deviation floors at 1e-12, and a new row's standardized
  feature on them is ~1e10 (the failure this pins).
"""
import numpy as np

import warmstart


def test_pca_spans_the_leading_directions():
    rng = np.random.default_rng(0)
    X = rng.standard_normal((50, 8)) @ rng.standard_normal((8, 30))  # rank 8
    mean, basis = warmstart.WarmStart._pca(X, 8)
    assert basis.shape == (30, 8)
    np.testing.assert_allclose(mean, X.mean(axis=0))
    np.testing.assert_allclose(basis.T @ basis, np.eye(8), atol=1e-12)
    Xc = X - mean
    np.testing.assert_allclose((Xc @ basis) @ basis.T, Xc, atol=1e-10)
    # the same span as numpy's leading singular vectors
    _, _, vt = np.linalg.svd(Xc, full_matrices=False)
    P = basis @ basis.T
    np.testing.assert_allclose(P @ vt[:8].T, vt[:8].T, atol=1e-10)


def test_pca_beyond_the_rank_maps_new_rows_to_zero():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((12, 40))  # 12 rows: centred rank 11
    mean, basis = warmstart.WarmStart._pca(X, 32)
    assert basis.shape == (40, 32)
    nonzero = np.flatnonzero(np.abs(basis).sum(axis=0) > 0)
    assert nonzero.tolist() == list(range(11))
    new = rng.standard_normal(40)
    F = (new - mean) @ basis
    assert np.all(F[11:] == 0.0)
    # standardized as WarmStart.features does: every feature stays O(1)
    Ftr = (X - mean) @ basis
    z = (F - Ftr.mean(axis=0)) / (Ftr.std(axis=0) + 1e-12)
    assert np.all(np.abs(z) < 1e3), np.abs(z).max()
