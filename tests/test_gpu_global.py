"""Global neighbourhood above 64 points (kriging.rs:396-406, 609-612): one
LU of the shared system on the GPU, per-target substitutions.  Parity with
the oracle (which solves the same system per target: the same factors) at
1e-10 relative; at larger sizes, exactness at the data points."""
import math

import numpy as np
import pytest

import _cases as cs
import geostat_hip as gh
from geostat_hip import validation as gv
from geostat_hip import workloads

pytestmark = pytest.mark.gpu


def rel_inf(a, b, floor):
    a, b = np.asarray(a), np.asarray(b)
    assert np.array_equal(np.isnan(a), np.isnan(b)), "NaN patterns differ"
    m = ~np.isnan(b)
    if not m.any():
        return 0.0
    return float(np.abs(a[m] - b[m]).max() / max(np.abs(b[m]).max(), floor))


def compare(model, got, want, tol=1e-10):
    v, var, lag, st = got
    rc, ov, ovar, olag, ost, _ = want
    assert rc == 0
    assert np.array_equal(st, ost), (st[:10], ost[:10])
    c0 = model.total_sill()
    assert rel_inf(v, ov, 1e-3 * math.sqrt(c0)) < tol
    assert rel_inf(var, ovar, 1e-3 * c0) < tol
    ok = (st == 0) & ~np.isnan(olag)
    if ok.any():
        assert np.abs(lag[ok] - olag[ok]).max() <= tol * max(np.abs(olag[ok]).max(), 1.0)


@pytest.mark.parametrize("method", [gh.KrigingMethod.Simple(0.2), gh.KrigingMethod.Ordinary(),
                                    gh.KrigingMethod.Universal(1), gh.KrigingMethod.Universal(2)],
                         ids=["sk", "ok", "uk1", "uk2"])
@pytest.mark.parametrize("mname", ["sph", "mat15_aniso", "matern_nu"])
def test_global_matches_oracle(oracle, method, mname):
    c, v = workloads.points(71, 300, [1000.0, 1000.0])
    t = workloads.points(72, 400, [1000.0, 1000.0])[0]
    t[:20] = c[:20]
    m = cs.MODELS_2D[mname]
    cfg = gh.KrigingConfig(method=method)
    k = gh.Kriging(c, v, m, cfg)
    compare(m, k.predict_many(t), oracle.predict_many(c, v, m, cfg, t))


def test_global_3d_measurement_error_drift_and_block(oracle):
    c, v = workloads.points(73, 200, [500.0, 500.0, 50.0], recipe=1)
    t = workloads.points(74, 300, [500.0, 500.0, 50.0])[0]
    m = gh.VariogramModel(0.1, [gh.Structure(gh.ModelKind.EXPONENTIAL, 1.0, 150.0,
                                             gh.Anisotropy(20.0, 0.7, 0.5, 10.0, 0.0))])
    err = np.full(200, 0.05)
    cfg = gh.KrigingConfig(method=gh.KrigingMethod.Universal(1))
    k = gh.Kriging.with_measurement_error(c, v, m, cfg, err)
    compare(m, k.predict_many(t), oracle.predict_many_ex(c, v, m, cfg, t, errors=err))
    d = np.column_stack([np.sin(c[:, 0] / 70.0), c[:, 2] / 50.0])
    td = np.column_stack([np.sin(t[:, 0] / 70.0), t[:, 2] / 50.0])
    cfg = gh.KrigingConfig(method=gh.KrigingMethod.ExternalDrift(2))
    k = gh.Kriging.with_external_drift(c, v, m, cfg, d)
    compare(m, k.predict_many_with_drift(t, td),
            oracle.predict_many_ex(c, v, m, cfg, t, drift=d, tdrift=td))
    offs = np.array([[0, 0, -5.0], [0, 0, 5.0], [10.0, -10.0, 0.0]])
    k = gh.Kriging(c, v, m, gh.KrigingConfig())
    compare(m, k.predict_block_many(t[:100], offs),
            oracle.predict_many_ex(c, v, m, gh.KrigingConfig(), t[:100], offsets=offs,
                                   block=True))


def test_global_singular_system_fails_every_target(oracle):
    # collinear data: the degree-1 drift block is rank deficient
    x = np.linspace(0.0, 100.0, 80)
    c = np.column_stack([x, np.zeros(80)])
    v = np.sin(x / 10.0)
    m = gh.VariogramModel(0.0, [gh.Structure(gh.ModelKind.SPHERICAL, 1.0, 30.0)])
    cfg = gh.KrigingConfig(method=gh.KrigingMethod.Universal(1))
    t = np.array([[5.0, 1.0], [50.0, -2.0]])
    compare(m, gh.Kriging(c, v, m, cfg).predict_many(t), oracle.predict_many(c, v, m, cfg, t))


def test_loo_global_on_reference_field(oracle):
    # validation.rs:906-935 (120 points, default = global configuration)
    r = oracle.Rng(11)
    c, v = [], []
    for _ in range(120):
        x = r.uniform() * 100.0
        y = r.uniform() * 100.0
        c.append([x, y])
        v.append(math.sin(x / 20.0) + math.cos(y / 20.0))
    c, v = np.array(c), np.array(v)
    m = gh.VariogramModel(0.01, [gh.Structure(gh.ModelKind.SPHERICAL, 1.0, 50.0)])
    cv = gv.leave_one_out(c, v, m, gh.KrigingConfig())
    rc, p, var = oracle.cv_with_folds(c, v, m, gh.KrigingConfig(), list(range(120)),
                                      [[i] for i in range(120)])
    assert rc == 0
    assert np.abs(cv.predicted - p).max() <= 1e-10 * np.abs(p).max()
    assert np.abs(cv.variance - var).max() <= 1e-10 * max(np.abs(var).max(), 1e-3)
    std = math.sqrt(np.mean((v - v.mean()) ** 2))
    assert cv.rmse() < 0.5 * std and abs(cv.mean_error()) < 0.1 and np.isfinite(cv.msdr())


def test_global_large_honours_data():
    # 3000 points: the predictor is exact at the data (value = datum,
    # variance 0) and smooth elsewhere
    c, v = workloads.points(75, 3000, [1000.0, 1000.0])
    m = gh.VariogramModel(0.05, [gh.Structure(gh.ModelKind.EXPONENTIAL, 1.0, 200.0)])
    k = gh.Kriging(c, v, m, gh.KrigingConfig())
    val, var, lag, st = k.predict_many(c[::7])
    assert (st == 0).all()
    assert np.abs(val - v[::7]).max() < 1e-8 * np.abs(v).max()
    assert np.abs(var).max() < 1e-8
    t = workloads.points(76, 5000, [1000.0, 1000.0])[0]
    val, var, lag, st = k.predict_many(t)
    assert (st == 0).all() and np.isfinite(val).all() and (var >= 0).all()
    assert (var <= m.total_sill() + 1e-9).all()
