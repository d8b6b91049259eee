"""Seeded sweep over random cross-validation runs (leave-one-out, k-fold,
spatial blocks; 2-D and 3-D; local and global neighbourhoods; nested models)
against the oracle's fold restatement.  Bars as in the validation tests:
predictions and variances within 1e-10 relative, NaN patterns equal."""
import numpy as np
import pytest

import geostat_hip as gh
from geostat_hip import validation as gv

pytestmark = pytest.mark.gpu

KINDS = [gh.ModelKind.SPHERICAL, gh.ModelKind.EXPONENTIAL, gh.ModelKind.MATERN15,
         gh.ModelKind.Matern(0.8), gh.ModelKind.Stable(1.5)]


def _close(cv, want, model):
    rc, p, var = want
    assert rc == 0
    assert np.array_equal(np.isnan(cv.predicted), np.isnan(p))
    m = ~np.isnan(p)
    c0 = model.total_sill()
    if m.any():
        assert np.abs(cv.predicted[m] - p[m]).max() <= 1e-10 * max(np.abs(p[m]).max(), 1e-3 * np.sqrt(c0))
        assert np.abs(cv.variance[m] - var[m]).max() <= 1e-10 * max(np.abs(var[m]).max(), 1e-3 * c0)


@pytest.mark.parametrize("seed", range(12))
def test_random_cross_validation_matches_oracle(oracle, seed):
    r = np.random.default_rng(13000 + seed)
    D = 2 if seed % 3 else 3
    n = int(r.integers(80, 600))
    ext = np.array([500.0, 500.0, 80.0][:D])
    c = r.uniform(0.0, 1.0, size=(n, D)) * ext
    v = np.sin(c[:, 0] / 80.0) + np.cos(c[:, 1] / 60.0) + 0.1 * r.standard_normal(n)
    model = gh.VariogramModel(float(r.uniform(0.02, 0.2)),
                              [gh.Structure(KINDS[int(r.integers(len(KINDS)))],
                                            float(r.uniform(0.5, 1.0)), float(r.uniform(60, 250)))
                               for _ in range(int(r.integers(1, 3)))])
    method = [gh.KrigingMethod.Simple(0.1), gh.KrigingMethod.Ordinary(),
              gh.KrigingMethod.Universal(1)][int(r.integers(3))]
    kk = [None, int(r.integers(8, 33))][int(r.integers(2))] if n <= 400 else int(r.integers(8, 33))
    cfg = gh.KrigingConfig(method=method, max_neighbors=kk)
    scheme = ("loo", "kfold", "block")[seed % 3]
    print(f"cv sweep seed {seed}: {scheme} D={D} n={n} method={method.kind} k={kk}", flush=True)
    if scheme == "loo":
        cv = gv.leave_one_out(c, v, model, cfg)
        want = oracle.cv_with_folds(c, v, model, cfg, list(range(n)), [[i] for i in range(n)])
    elif scheme == "kfold":
        folds, sd = int(r.integers(2, 11)), int(r.integers(0, 2 ** 62))
        cv = gv.k_fold(c, v, model, cfg, folds, sd)
        fo, mem = oracle.fold_assignment(n, folds, sd)
        want = oracle.cv_with_folds(c, v, model, cfg, fo, mem)
    else:
        blocks = [int(r.integers(2, 5)) for _ in range(D)]
        cv = gv.block_cv(c, v, model, cfg, blocks)
        fo, mem = oracle.spatial_block_assignment(c, blocks)
        want = oracle.cv_with_folds(c, v, model, cfg, fo, mem)
    _close(cv, want, model)
