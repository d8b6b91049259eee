"""GPU cross-validation (validation.rs:136-400) against the oracle's fold
restatement: identical fold assignments; predictions and variances within
1e-10 relative (floors 1e-3 sqrt(c0) / 1e-3 c0)."""
import math

import numpy as np
import pytest

import geostat_hip as gh
from geostat_hip import validation as gv

pytestmark = pytest.mark.gpu


def smooth_field(oracle, n, seed, scale=20.0):
    r = oracle.Rng(seed)
    c, v = [], []
    for _ in range(n):
        x = r.uniform() * 100.0
        y = r.uniform() * 100.0
        c.append([x, y])
        v.append(math.sin(x / scale) + math.cos(y / scale))
    return np.array(c), np.array(v)


MODEL = gh.VariogramModel(0.01, [gh.Structure(gh.ModelKind.SPHERICAL, 1.0, 50.0)])


def close(cv, want, model=MODEL):
    rc, p, v = want
    assert rc == 0
    c0 = model.total_sill()
    assert np.abs(cv.predicted - p).max() <= 1e-10 * max(np.abs(p).max(), 1e-3 * math.sqrt(c0))
    assert np.abs(cv.variance - v).max() <= 1e-10 * max(np.abs(v).max(), 1e-3 * c0)


def test_leave_one_out_global_and_k_equals_n(oracle):
    # validation.rs:1005-1019 (global neighbourhood, 40 points)
    c, v = smooth_field(oracle, 40, 5)
    cfg = gh.KrigingConfig()
    loo = gv.leave_one_out(c, v, MODEL, cfg)
    close(loo, oracle.cv_with_folds(c, v, MODEL, cfg, list(range(40)), [[i] for i in range(40)]))
    kf = gv.k_fold(c, v, MODEL, cfg, 40, 123)
    assert np.abs(kf.predicted - loo.predicted).max() < 1e-12
    assert loo.rmse() < 0.5 * math.sqrt(np.mean((v - v.mean()) ** 2))


def test_k_fold_matches_oracle_and_is_deterministic(oracle):
    # validation.rs:1021-1039
    c, v = smooth_field(oracle, 53, 8)
    cfg = gh.KrigingConfig()
    a = gv.k_fold(c, v, MODEL, cfg, 5, 42)
    b = gv.k_fold(c, v, MODEL, cfg, 5, 42)
    assert np.array_equal(a.predicted, b.predicted)
    fo, mem = oracle.fold_assignment(53, 5, 42)
    close(a, oracle.cv_with_folds(c, v, MODEL, cfg, fo, mem))
    assert a.vecv() > 50.0


@pytest.mark.parametrize("mn", [12, 24])
def test_block_cv_and_local_loo_match_oracle(oracle, mn):
    c, v = smooth_field(oracle, 300, 7)
    cfg = gh.KrigingConfig(max_neighbors=mn)
    cv = gv.block_cv(c, v, MODEL, cfg, [4, 4])
    fo, mem = oracle.spatial_block_assignment(c, [4, 4])
    close(cv, oracle.cv_with_folds(c, v, MODEL, cfg, fo, mem))
    assert cv.vecv() > 0.0 and np.isfinite(cv.predicted).all()
    loo = gv.leave_one_out(c, v, MODEL, cfg)
    close(loo, oracle.cv_with_folds(c, v, MODEL, cfg, list(range(300)),
                                    [[i] for i in range(300)]))


def test_leave_one_out_with_drift(oracle):
    c, v = smooth_field(oracle, 150, 9)
    d = np.column_stack([np.sin(c[:, 0] / 20.0)])
    cfg = gh.KrigingConfig(method=gh.KrigingMethod.ExternalDrift(1), max_neighbors=16)
    cv = gv.leave_one_out_with_drift(c, v, d, MODEL, cfg)
    close(cv, oracle.cv_with_folds(c, v, MODEL, cfg, list(range(150)),
                                   [[i] for i in range(150)], drift=d))


def test_cv_rejections(oracle):
    # validation.rs:983-990, 1041-1050, 1100-1110
    with pytest.raises(gh.InsufficientData):
        gv.leave_one_out([[0.0, 0.0], [1.0, 1.0]], [1.0, 2.0], MODEL, gh.KrigingConfig())
    c, v = smooth_field(oracle, 10, 1)
    with pytest.raises(gh.InvalidParameter):
        gv.k_fold(c, v, MODEL, gh.KrigingConfig(), 1, 0)
    with pytest.raises(gh.InvalidParameter):
        gv.k_fold(c, v, MODEL, gh.KrigingConfig(), 11, 0)
    c, v = smooth_field(oracle, 20, 2)
    with pytest.raises(gh.InvalidParameter):
        gv.block_cv(c, v, MODEL, gh.KrigingConfig(), [0, 2])
    with pytest.raises(gh.InvalidParameter):
        gv.block_cv(c, v, MODEL, gh.KrigingConfig(), [1, 1])
    gv.block_cv(c, v, MODEL, gh.KrigingConfig(), [3, 3])
