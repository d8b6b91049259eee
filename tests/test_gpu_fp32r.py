"""GPU parity of the FP32-refined path (GS_PREC_FP32_REFINED).

Bars: neighbour lists are the FP64 path's (the search does not depend on
the precision, so they are bit-exact against the oracle); value and variance
within 1e-6 normwise relative of the oracle (floors 1e-3 sqrt(c0) and
1e-3 c0, as for the FP64 path), the Lagrange multiplier within 1e-5; the
per-target status pattern equal to the oracle's (every singular decision is
taken by the FP64 kernel).  Configurations outside the refined kernels
(universal kriging, extended systems, general model kinds, global
neighbourhoods) are computed in FP64: bit-equal to the FP64 call.  Vecchia:
each log-likelihood within 1e-6 relative of the oracle's on 4,000 points; at
10^6 points the batch within 1e-6 normwise of the FP64 path's, and every
model within 1e-7 per point.
"""
import numpy as np
import pytest

import _cases as cs
import geostat_hip as gh
from geostat_hip import workloads

pytestmark = pytest.mark.gpu

TOL = 1e-6


def rel(a, b, floor):
    a, b = np.asarray(a), np.asarray(b)
    assert np.array_equal(np.isnan(a), np.isnan(b)), "NaN patterns differ"
    m = ~np.isnan(b)
    if not m.any():
        return 0.0
    return float(np.abs(a[m] - b[m]).max() / max(np.abs(b[m]).max(), floor))


def check32(oracle, coords, values, model, config, targets, lag_tol=1e-5):
    k = gh.Kriging(coords, values, model, config)
    v, var, lag, st = k.predict_many(targets, precision="fp32r")
    rc, ov, ovar, olag, ost, _ = oracle.predict_many(coords, values, model, config, targets)
    assert rc == 0
    assert np.array_equal(st, ost), (np.nonzero(st != ost)[0][:10], st[st != ost][:10])
    c0 = model.total_sill()
    ev = rel(v, ov, 1e-3 * np.sqrt(c0))
    evar = rel(var, ovar, 1e-3 * c0)
    assert ev < TOL and evar < TOL, (ev, evar)
    if config.method.kind != 0:
        assert rel(lag, olag, 1e-3) < lag_tol
    return k, v, var, lag, st


def test_ok2d_large_sample(oracle):
    w = workloads.ok2d_large()
    t = w.targets()[::199]
    print(f"fp32r ok2d_large sample: T={len(t)}", flush=True)
    k, v, var, lag, st = check32(oracle, w.coords, w.values, w.model, w.config, t)
    # the same search as the FP64 path
    idx, cnt = k.neighbors(t)
    _, oidx, ocnt = oracle.neighbors(w.coords, w.values, w.model, w.config, t, idx.shape[1])
    assert np.array_equal(idx, oidx) and np.array_equal(cnt, ocnt)


def test_ok2d_small_full(oracle):
    w = workloads.ok2d_small()
    print("fp32r ok2d_small: T=10000", flush=True)
    check32(oracle, w.coords, w.values, w.model, w.config, w.targets())


@pytest.mark.parametrize("mname", ["sph", "exp", "gau", "mat15_aniso", "mat25", "nested"])
@pytest.mark.parametrize("method", [gh.KrigingMethod.Simple(0.4), gh.KrigingMethod.Ordinary()])
@pytest.mark.parametrize("k", [12, 32, 60])
def test_lean_kinds_sizes_methods(oracle, mname, method, k):
    c, v = workloads.points(61, 6000, [1000.0, 1000.0])
    t = workloads.points(62, 1500, [1000.0, 1000.0])[0]
    print(f"fp32r kind={mname} method={method.kind} k={k}", flush=True)
    check32(oracle, c, v, cs.MODELS_2D[mname], gh.KrigingConfig(method=method, max_neighbors=k), t)


@pytest.mark.parametrize("method", [gh.KrigingMethod.Simple(0.0), gh.KrigingMethod.Ordinary()])
def test_3d_rotated_octant_and_radius(oracle, method):
    c, v = workloads.points(71, 20000, [1000.0, 1000.0, 200.0], recipe=1)
    t = workloads.points(72, 1500, [1000.0, 1000.0, 200.0], recipe=1)[0]
    an = gh.Anisotropy(30.0, 0.5, 0.3, 20.0, 40.0)
    m = gh.VariogramModel(0.1, [gh.Structure(gh.ModelKind.EXPONENTIAL, 0.6, 200.0, an),
                                gh.Structure(gh.ModelKind.SPHERICAL, 0.4, 500.0, an)])
    for cfg in [gh.KrigingConfig(method=method, max_neighbors=30),
                gh.KrigingConfig(method=method, max_neighbors=64, max_per_octant=6,
                                 search_radius=150.0, anisotropic_search=an, min_neighbors=10),
                gh.KrigingConfig(method=method, max_neighbors=20, search_radius=40.0)]:
        print(f"fp32r 3-D method={method.kind} cfg={cfg}", flush=True)
        check32(oracle, c, v, m, cfg, t)


def test_ill_conditioned_systems_fall_back_to_fp64(oracle):
    # Gaussian model without nugget: the FP32 pivots collapse, every such
    # system is solved by the FP64 kernel (same bar as the FP64 path)
    c, v = workloads.points(81, 3000, [1000.0, 1000.0])
    t = workloads.points(82, 800, [1000.0, 1000.0])[0]
    m = gh.VariogramModel(0.0, [gh.Structure(gh.ModelKind.GAUSSIAN, 1.0, 60.0)])
    cfg = gh.KrigingConfig(max_neighbors=12)
    k = gh.Kriging(c, v, m, cfg)
    a = k.predict_many(t, precision="fp64")
    b = k.predict_many(t, precision="fp32r")
    assert np.array_equal(a[3], b[3])
    assert rel(b[0], a[0], 1e-3) < TOL and rel(b[1], a[1], 1e-3) < TOL
    # near-duplicate points: singular in both paths, same pattern as the oracle
    c2 = np.array([[0.0, 0.0], [1e-7, 0.0], [5.0, 5.0], [0.0, 1e-7]])
    m2 = gh.VariogramModel(0.0, [gh.Structure(gh.ModelKind.GAUSSIAN, 1.0, 100.0)])
    k2 = gh.Kriging(c2, [1.0, 1.1, 2.0, 0.9], m2, gh.KrigingConfig(max_neighbors=4))
    _, _, _, st = k2.predict_many([[0.5, 0.5], [3.0, 3.0]], precision="fp32r")
    assert (st == 3).all()


@pytest.mark.parametrize("sill", [1e-4, 1e6, 1e7])
def test_scale_relative_singular_test_in_refined_path(oracle, sill):
    c, v = workloads.points(5, 300, [1000.0, 1000.0])
    t = workloads.uniforms(9, 400).reshape(200, 2) * 1000.0
    m = gh.VariogramModel(0.1 * sill, [gh.Structure(gh.ModelKind.SPHERICAL, sill, 300.0)])
    check32(oracle, c, v * np.sqrt(sill), m, gh.KrigingConfig(max_neighbors=32), t)


def test_other_configurations_are_fp64():
    c, v = workloads.points(91, 4000, [1000.0, 1000.0])
    t = workloads.points(92, 500, [1000.0, 1000.0])[0]
    for m, cfg in [(cs.MODELS_2D["exp"], gh.KrigingConfig(method=gh.KrigingMethod.Universal(1),
                                                         max_neighbors=24)),
                   (cs.MODELS_2D["matern_nu"], gh.KrigingConfig(max_neighbors=24)),
                   (cs.MODELS_2D["exp"], gh.KrigingConfig())]:       # global neighbourhood
        k = gh.Kriging(c[:2000], v[:2000], m, cfg)
        a = k.predict_many(t, precision="fp64")
        b = k.predict_many(t, precision="fp32r")
        for x, y in zip(a, b):
            assert np.array_equal(x, y, equal_nan=True)
    ke = gh.Kriging.with_measurement_error(c, v, cs.MODELS_2D["sph"], gh.KrigingConfig(max_neighbors=16),
                                           np.full(len(c), 0.01))
    a = ke.predict_many(t, precision="fp64")
    b = ke.predict_many(t, precision="fp32r")
    for x, y in zip(a, b):
        assert np.array_equal(x, y, equal_nan=True)


def test_full_size_against_fp64_path():
    # every one of the 4*10^6 targets: within 1e-6 of the FP64 GPU path
    # (itself within 1e-10 of the oracle), same statuses, deterministic
    w = workloads.ok2d_large()
    k = gh.Kriging(w.coords, w.values, w.model, w.config)
    t = w.targets()
    print(f"fp32r full size: T={len(t)}", flush=True)
    a = k.predict_many(t, precision="fp64")
    b = k.predict_many(t, precision="fp32r")
    c = k.predict_many(t, precision="fp32r")
    assert np.array_equal(a[3], b[3]) and (b[3] == 0).all()
    assert rel(b[0], a[0], 1e-3) < TOL and rel(b[1], a[1], 1e-3 * 1.05) < TOL
    assert rel(b[2], a[2], 1e-3) < 1e-5
    for x, y in zip(b, c):
        assert np.array_equal(x, y)


def test_vecchia_against_oracle(oracle):
    c, _ = workloads.points(7, 4000, [1.0, 1.0], recipe=2)
    for m in (10, 30, 60):
        p = gh.vecchia_plan(c, m)
        o, nb, cnt = p.arrays()
        vals = workloads.vecchia_field(c, o, nb, cnt, m)
        models = workloads.vecchia_models(16)
        ll, st = p.loglik_batch(vals, models, precision="fp32r")
        print(f"fp32r vecchia m={m}", flush=True)
        for i, mdl in enumerate(models):
            rc, oll = oracle.vecchia_loglik_plan(c, vals, mdl, o, nb, cnt)
            assert rc == 0 and st[i] == 0
            assert abs(ll[i] - oll) <= TOL * abs(oll), (i, ll[i], oll)


def test_vecchia_full_size_against_fp64_path():
    c = workloads.vecchia_c3()
    p = gh.vecchia_plan(c, 30)
    o, nb, cnt = p.arrays()
    vals = workloads.vecchia_field(c, o, nb, cnt, 30)
    models = workloads.vecchia_models(64)
    print("fp32r vecchia full size: n=10^6 m=30, 64 models", flush=True)
    a, sa = p.loglik_batch(vals, models)
    b, sb = p.loglik_batch(vals, models, precision="fp32r")
    assert np.array_equal(sa, sb) and (sa == 0).all()
    # normwise over the batch (the metric's relative error), and per model
    # the mean per-point error: a total that nearly cancels (one of the 64
    # sums to about -7e3 over 10^6 terms of order 1) is no scale for it
    d = np.abs(a - b)
    assert d.max() <= TOL * np.abs(a).max(), d.max() / np.abs(a).max()
    assert (d / len(c) <= 1e-7).all(), (d / len(c)).max()
