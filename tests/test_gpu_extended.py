"""GPU parity of the extended predictors against the oracle: measurement
error (kriging.rs:181-218), external drift (223-260, 505-522, 656-672) and
block kriging (678-803).

Bars as for point kriging: per-target status equal, FP64 value and variance
within 1e-10 relative (floors 1e-3 sqrt(c0) / 1e-3 c0).
"""
import numpy as np
import pytest

import _cases as cs
import geostat_hip as gh
from geostat_hip import workloads

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def rel_inf(a, b, floor):
    a, b = np.asarray(a), np.asarray(b)
    assert np.array_equal(np.isnan(a), np.isnan(b)), "NaN patterns differ"
    m = ~np.isnan(b)
    if not m.any():
        return 0.0
    return float(np.abs(a[m] - b[m]).max() / max(np.abs(b[m]).max(), floor))


def compare(model, got, want):
    v, var, lag, st = got
    rc, ov, ovar, olag, ost, _ = want
    assert rc == 0
    assert np.array_equal(st, ost), (np.nonzero(st != ost)[0][:10], st[st != ost][:10],
                                     ost[st != ost][:10])
    c0 = model.total_sill()
    assert rel_inf(v, ov, 1e-3 * np.sqrt(c0)) < TOL64
    assert rel_inf(var, ovar, 1e-3 * c0) < TOL64


MODELS = {
    "sph": cs.MODELS_2D["sph"],
    "mat15_aniso": cs.MODELS_2D["mat15_aniso"],
    "matern_nu": cs.MODELS_2D["matern_nu"],
}


# ---- measurement error ----------------------------------------------------
@pytest.mark.parametrize("mname", sorted(MODELS))
@pytest.mark.parametrize("method", [gh.KrigingMethod.Simple(0.2), gh.KrigingMethod.Ordinary(),
                                    gh.KrigingMethod.Universal(1), gh.KrigingMethod.Universal(2)],
                         ids=["sk", "ok", "uk1", "uk2"])
def test_measurement_error_matches_oracle(oracle, mname, method):
    c, v = workloads.points(31, 3000, [1000.0, 1000.0])
    err = np.abs(np.sin(np.arange(3000) * 0.37)) * 0.2
    t = workloads.points(32, 2000, [1000.0, 1000.0])[0]
    t[:50] = c[:50]    # targets on data points: smoothed, not honoured
    m = MODELS[mname]
    cfg = gh.KrigingConfig(method=method, max_neighbors=24)
    k = gh.Kriging.with_measurement_error(c, v, m, cfg, err)
    compare(m, k.predict_many(t), oracle.predict_many_ex(c, v, m, cfg, t, errors=err))


def test_measurement_error_3d_octant_and_zero_errors(oracle):
    c, v = workloads.points(33, 4000, [500.0, 500.0, 50.0], recipe=1)
    t = workloads.points(34, 1500, [500.0, 500.0, 50.0])[0]
    m = gh.VariogramModel(0.1, [
        gh.Structure(gh.ModelKind.EXPONENTIAL, 0.6, 120.0, gh.Anisotropy(30.0, 0.6, 0.4, 10.0, 5.0)),
        gh.Structure(gh.ModelKind.SPHERICAL, 0.4, 300.0)])
    err = np.full(4000, 0.05)
    cfg = gh.KrigingConfig(max_per_octant=4, min_neighbors=4)
    k = gh.Kriging.with_measurement_error(c, v, m, cfg, err)
    compare(m, k.predict_many(t), oracle.predict_many_ex(c, v, m, cfg, t, errors=err))
    # zero error variances reproduce the plain predictor (kriging.rs:885-897)
    z = gh.Kriging.with_measurement_error(c, v, m, cfg, np.zeros(4000)).predict_many(t)
    p = gh.Kriging(c, v, m, cfg).predict_many(t)
    assert np.array_equal(z[3], p[3])
    assert rel_inf(z[0], p[0], 1e-3) < 1e-12 and rel_inf(z[1], p[1], 1e-3) < 1e-12


def test_measurement_error_rejections():
    c, v = cs.sample_data()
    m = cs.sample_model()
    with pytest.raises(gh.InvalidParameter):
        gh.Kriging.with_measurement_error(c, v, m, gh.KrigingConfig(), np.full(5, -0.1))
    with pytest.raises(gh.DimensionMismatch):
        gh.Kriging.with_measurement_error(c, v, m, gh.KrigingConfig(), np.full(2, 0.1))
    with pytest.raises(gh.InvalidParameter):
        gh.Kriging.with_measurement_error(
            c, v, m, gh.KrigingConfig(method=gh.KrigingMethod.ExternalDrift(1)), np.zeros(5))


# ---- external drift -------------------------------------------------------
@pytest.mark.parametrize("nvars", [1, 2, 3])
@pytest.mark.parametrize("dim", [2, 3])
def test_external_drift_matches_oracle(oracle, nvars, dim):
    ext = [1000.0, 1000.0] if dim == 2 else [1000.0, 1000.0, 100.0]
    c, v = workloads.points(40 + nvars, 3000, ext)
    t = workloads.points(50 + nvars, 1500, ext)[0]
    dd = np.column_stack([np.sin(c[:, 0] / (90.0 + 7 * k)) + 0.01 * c[:, 1] / (k + 1)
                          for k in range(nvars)])
    td = np.column_stack([np.sin(t[:, 0] / (90.0 + 7 * k)) + 0.01 * t[:, 1] / (k + 1)
                          for k in range(nvars)])
    v = v + dd.sum(1)
    m = gh.VariogramModel(0.05, [gh.Structure(gh.ModelKind.MATERN15, 1.0, 200.0)])
    cfg = gh.KrigingConfig(method=gh.KrigingMethod.ExternalDrift(nvars), max_neighbors=20)
    k = gh.Kriging.with_external_drift(c, v, m, cfg, dd)
    compare(m, k.predict_many_with_drift(t, td),
            oracle.predict_many_ex(c, v, m, cfg, t, drift=dd, tdrift=td))


def test_external_drift_linear_relation_and_usage_errors(oracle):
    # kriging.rs:1415-1450
    coords, values, drift = [], [], []
    for i in range(5):
        for j in range(5):
            x, y = i * 10.0, j * 10.0
            d = np.sin(x * 0.13 + y * 0.07)
            coords.append([x, y])
            drift.append([d])
            values.append(2.0 + 3.0 * d)
    c, v, dd = np.array(coords), np.array(values), np.array(drift)
    m = cs.sample_model()
    cfg = gh.KrigingConfig(method=gh.KrigingMethod.ExternalDrift(1))
    k = gh.Kriging.with_external_drift(c, v, m, cfg, dd)
    val, _, _, st = k.predict_many_with_drift([[12.0, 33.0], [45.0, 5.0]], [[0.4], [-0.7]])
    assert (st == 0).all()
    assert abs(val[0] - 3.2) < 1e-7 and abs(val[1] - (2.0 - 2.1)) < 1e-7
    # predict on an external-drift predictor, predict_with_drift on an
    # ordinary one: every target fails (NaN, InvalidParameter)
    v2, _, _, st2 = k.predict_many([[0.0, 0.0]])
    assert st2[0] == 4 and np.isnan(v2[0])
    ok = gh.Kriging(c, v, m, gh.KrigingConfig())
    v3, _, _, st3 = ok.predict_many_with_drift([[0.0, 0.0]], [[1.0]])
    assert st3[0] == 4 and np.isnan(v3[0])
    with pytest.raises(gh.DimensionMismatch):
        k.predict_many_with_drift([[0.0, 0.0]], [[1.0, 2.0]])
    with pytest.raises(gh.InvalidParameter):
        gh.Kriging.with_external_drift(c, v, m, gh.KrigingConfig(), dd)
    with pytest.raises(gh.InvalidParameter):
        gh.Kriging(c, v, m, cfg)


# ---- block kriging --------------------------------------------------------
@pytest.mark.parametrize("mname", sorted(MODELS))
@pytest.mark.parametrize("method", [gh.KrigingMethod.Simple(0.1), gh.KrigingMethod.Ordinary()],
                         ids=["sk", "ok"])
def test_block_kriging_matches_oracle(oracle, mname, method):
    c, v = workloads.points(61, 3000, [1000.0, 1000.0])
    grid = gh.Grid2D(0.0, 0.0, 25.0, 25.0, 40, 40)
    offs = gh.block_offsets([25.0, 25.0], [3, 4])
    m = MODELS[mname]
    cfg = gh.KrigingConfig(method=method, max_neighbors=20)
    k = gh.Kriging(c, v, m, cfg)
    compare(m, k.predict_block_many(grid.centers(), offs),
            oracle.predict_many_ex(c, v, m, cfg, grid.centers(), offsets=offs, block=True))
    # with measurement error as well
    err = np.full(3000, 0.02)
    km = gh.Kriging.with_measurement_error(c, v, m, cfg, err)
    compare(m, km.predict_block_many(grid.centers()[:300], offs),
            oracle.predict_many_ex(c, v, m, cfg, grid.centers()[:300], errors=err, offsets=offs,
                                   block=True))


def test_block_kriging_3d_and_reference_properties(oracle):
    # kriging.rs:1524-1566 and three_d.rs:219-250
    c, v = cs.sample_data()
    m = cs.sample_model()
    k = gh.Kriging(c, v, m, gh.KrigingConfig())
    center = np.array([0.6, 0.4])
    offs = gh.block_offsets([0.4, 0.4], [4, 4])
    b = k.predict_block(center, offs)
    pv = k.predict_many(center + offs)[0]
    assert abs(b.value - pv.mean()) < 1e-10
    assert b.variance < k.predict(center).variance
    uk = gh.Kriging(c, v, m, gh.KrigingConfig(method=gh.KrigingMethod.Universal(1)))
    with pytest.raises(gh.InvalidParameter):
        uk.predict_block(center, offs)
    with pytest.raises(gh.InvalidParameter):
        k.predict_block(center, np.zeros((0, 2)))
    r = oracle.Rng(13)
    c3 = np.array([[r.uniform() * 100.0, r.uniform() * 100.0, r.uniform() * 50.0]
                   for _ in range(100)])
    v3 = np.sin(c3[:, 0] / 20.0) + 0.01 * c3[:, 2]
    m3 = gh.VariogramModel(0.05, [gh.Structure(gh.ModelKind.SPHERICAL, 1.0, 30.0)])
    k3 = gh.Kriging(c3, v3, m3, gh.KrigingConfig(max_neighbors=40))
    ctr = np.array([42.0, 37.0, 18.0])
    one = k3.predict_block(ctr, np.zeros((1, 3)))
    pt = k3.predict(ctr)
    assert abs(one.value - pt.value) < 1e-10
    assert abs(one.variance - (pt.variance - m3.nugget)) < 1e-10
    offs3 = np.array([[0, 0, -5.0], [0, 0, 5.0], [3.0, -2.0, 1.0]])
    compare(m3, k3.predict_block_many(c3[:40] + 0.5, offs3),
            oracle.predict_many_ex(c3, v3, m3, gh.KrigingConfig(max_neighbors=40), c3[:40] + 0.5,
                                   offsets=offs3, block=True))
