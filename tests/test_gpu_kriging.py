"""GPU parity of the kriging predictor against the CPU oracle.

Bars: neighbour index lists bit-exact; FP64 estimates within 1e-10
relative (||x - x_ref||_inf / ||x_ref||_inf per output array); the
FP32-refined path within 1e-6.  NaN patterns (per-target failures) must
match exactly.
"""
import numpy as np
import pytest

import _cases as cs
import geostat_hip as gh
from geostat_hip import workloads

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32R = 1e-6


def rel_inf(a, b, floor=1e-300):
    """||a - b||_inf / max(||b||_inf, floor) over finite entries; the floor
    keeps an all-zero reference (variance at data points) meaningful."""
    a, b = np.asarray(a), np.asarray(b)
    assert np.array_equal(np.isnan(a), np.isnan(b)), "NaN patterns differ"
    m = ~np.isnan(b)
    if not m.any():
        return 0.0
    scale = max(np.abs(b[m]).max(), floor)
    return float(np.abs(a[m] - b[m]).max() / scale)


def closer_to_exact(oracle, coords, values, model, config, targets, got, want, which):
    """Where GPU and oracle differ by more than the bar, both are inside the
    rounding noise of an ill-conditioned system: then the GPU must be at
    least as close as the oracle (the reference's LU) to the exact solution
    of the same system, computed in 40-digit arithmetic."""
    c0 = model.total_sill()
    floor = 1e-3 * (np.sqrt(c0) if which == 0 else c0)     # value; variance or multiplier
    scale = max(np.abs(want[~np.isnan(want)]).max(), floor)
    worst = np.argsort(-np.nan_to_num(np.abs(got - want)))[:8]
    _, idx, cnt = oracle.neighbors(coords, values, model, config, targets[worst],
                                   max(config.max_neighbors or 64, 1))
    for j, t in enumerate(worst):
        ex = cs.exact_estimate(oracle, coords, values, model, config.method, targets[t],
                               idx[j, :cnt[j]])[which]
        assert abs(got[t] - ex) <= max(2.0 * abs(want[t] - ex), TOL64 * scale), \
            (t, got[t], want[t], ex)


def check_predict(oracle, coords, values, model, config, targets, prec="fp64", tol=TOL64):
    k = gh.Kriging(coords, values, model, config)
    v, var, lag, st = k.predict_many(targets, precision=prec)
    rc, ov, ovar, olag, ost, _ = oracle.predict_many(coords, values, model, config, targets)
    assert rc == 0
    assert np.array_equal(st, ost), (np.nonzero(st != ost)[0][:10], st[st != ost][:10], ost[st != ost][:10])
    c0 = model.total_sill()
    for which, got, want in ((0, v, ov), (1, var, ovar)):
        floor = 1e-3 * (np.sqrt(c0) if which == 0 else c0)
        if rel_inf(got, want, floor) >= tol:
            closer_to_exact(oracle, coords, values, model, config, targets, got, want, which)
    if config.method.kind != 0 and rel_inf(lag, olag) >= tol * 10:
        # the multiplier of an ill-conditioned drift block: at least as close
        # as the oracle's to the exact solution's
        closer_to_exact(oracle, coords, values, model, config, targets, lag, olag, 2)
    return k, v, var


def check_neighbors(oracle, coords, values, model, config, targets):
    k = gh.Kriging(coords, values, model, config)
    idx, cnt = k.neighbors(targets)
    rc, oidx, ocnt = oracle.neighbors(coords, values, model, config, targets, idx.shape[1])
    assert rc == 0
    assert np.array_equal(cnt, ocnt)
    assert np.array_equal(idx, oidx)


# ---------------------------------------------------------------- configs --
def test_ok2d_small_full_size(oracle):
    w = workloads.ok2d_small()
    t = w.targets()
    print("ok2d_small: n=2000 k=20 T=10000", flush=True)
    check_neighbors(oracle, w.coords, w.values, w.model, w.config, t)
    check_predict(oracle, w.coords, w.values, w.model, w.config, t)


def test_ok2d_large_subset(oracle):
    w = workloads.ok2d_large(nx=2000, ny=2000)
    t = w.targets()[::199]          # ~20k targets spread over the 4M grid
    print(f"ok2d_large subset: n=200000 k=32 T={len(t)}", flush=True)
    check_neighbors(oracle, w.coords, w.values, w.model, w.config, t)
    check_predict(oracle, w.coords, w.values, w.model, w.config, t)


def test_uk3d_octant_subset(oracle):
    w = workloads.uk3d_octant(n=250_000, nx=200, ny=200, nz=100)
    t = w.targets()[::1000]         # 4000 targets
    print(f"uk3d_octant subset: n=250000 T={len(t)}", flush=True)
    check_neighbors(oracle, w.coords, w.values, w.model, w.config, t)
    k, v, var = check_predict(oracle, w.coords, w.values, w.model, w.config, t)
    _, cnt = k.neighbors(t)
    assert cnt.min() < 32 < cnt.max(), "octant sizes should span several buckets"


# ----------------------------------------------------------- search rules --
@pytest.mark.parametrize("k", [1, 4, 5, 8, 13, 24])
def test_lattice_ties_follow_reference_traversal(oracle, k):
    c, v = cs.lattice(37, 29)
    t = np.array([[x + 0.5, y + 0.5] for x in range(0, 36, 5) for y in range(0, 28, 3)]
                 + [[x, y] for x in range(0, 37, 4) for y in range(0, 29, 4)]
                 + [[x + 0.25, y] for x in range(0, 36, 7) for y in range(0, 28, 7)], dtype=float)
    m = cs.sample_model()
    check_neighbors(oracle, c, v, m, gh.KrigingConfig(max_neighbors=k), t)


def test_radius_limited_and_min_neighbors(oracle):
    c, v = workloads.points(11, 3000, [1000.0, 1000.0])
    t = workloads.points(12, 2000, [1200.0, 1200.0])[0] - 100.0
    m = cs.MODELS_2D["sph"]
    for cfg in [gh.KrigingConfig(max_neighbors=16, search_radius=25.0),
                gh.KrigingConfig(max_neighbors=32, search_radius=40.0, min_neighbors=5),
                gh.KrigingConfig(search_radius=18.0)]:
        check_neighbors(oracle, c, v, m, cfg, t)
        check_predict(oracle, c, v, m, cfg, t)


@pytest.mark.parametrize("mname", sorted(cs.MODELS_2D))
def test_model_kinds(oracle, mname):
    c, v = workloads.points(21, 4000, [1000.0, 1000.0])
    t = workloads.points(22, 3000, [1000.0, 1000.0])[0]
    check_predict(oracle, c, v, cs.MODELS_2D[mname], gh.KrigingConfig(max_neighbors=24), t)


@pytest.mark.parametrize("method", [gh.KrigingMethod.Simple(0.3), gh.KrigingMethod.Ordinary(),
                                    gh.KrigingMethod.Universal(1), gh.KrigingMethod.Universal(2)])
@pytest.mark.parametrize("k", [10, 32, 60])
def test_methods_2d(oracle, method, k):
    c, v = workloads.points(31, 5000, [1000.0, 1000.0])
    t = workloads.points(32, 2000, [1000.0, 1000.0])[0]
    check_predict(oracle, c, v, cs.MODELS_2D["exp"], gh.KrigingConfig(method=method, max_neighbors=k), t)


@pytest.mark.parametrize("method", [gh.KrigingMethod.Ordinary(), gh.KrigingMethod.Universal(1),
                                    gh.KrigingMethod.Universal(2), gh.KrigingMethod.Simple(0.0)])
def test_methods_3d_rotated(oracle, method):
    c, v = workloads.points(41, 20000, [1000.0, 1000.0, 200.0], recipe=1)
    t = workloads.points(42, 2000, [1000.0, 1000.0, 200.0], recipe=1)[0]
    an = gh.Anisotropy(30.0, 0.5, 0.3, 20.0, 40.0)
    m = gh.VariogramModel(0.1, [gh.Structure(gh.ModelKind.EXPONENTIAL, 0.6, 200.0, an),
                                gh.Structure(gh.ModelKind.SPHERICAL, 0.4, 500.0, an)])
    for cfg in [gh.KrigingConfig(method=method, max_neighbors=30),
                gh.KrigingConfig(method=method, max_neighbors=40, anisotropic_search=an),
                gh.KrigingConfig(method=method, max_neighbors=64, max_per_octant=6,
                                 search_radius=150.0, anisotropic_search=an, min_neighbors=10)]:
        check_neighbors(oracle, c, v, m, cfg, t)
        check_predict(oracle, c, v, m, cfg, t)


def test_octant_2d_and_aniso_search(oracle):
    c, v = workloads.points(51, 6000, [1000.0, 1000.0])
    t = workloads.points(52, 2000, [1000.0, 1000.0])[0]
    an = gh.Anisotropy(60.0, 0.3)
    m = cs.MODELS_2D["mat15_aniso"]
    for cfg in [gh.KrigingConfig(max_neighbors=16, max_per_octant=3),
                gh.KrigingConfig(max_per_octant=10, search_radius=60.0, anisotropic_search=an),
                gh.KrigingConfig(max_neighbors=20, anisotropic_search=an),
                gh.KrigingConfig(max_neighbors=12, max_per_octant=4, min_neighbors=12)]:
        check_neighbors(oracle, c, v, m, cfg, t)
        check_predict(oracle, c, v, m, cfg, t)


# ------------------------------------------------- reference behaviour tests --
def test_reference_small_cases(oracle):
    # kriging.rs:1204-1219 exact at data; 1360-1382 far field; 1570-1591 k = n
    c, v = cs.sample_data()
    m = cs.sample_model()
    k = gh.Kriging(c, v, m)
    val, var, _, st = k.predict_many(c)
    assert np.abs(val - v).max() < 1e-8 and var.max() < 1e-8 and (st == 0).all()
    e = k.predict([1e6, 1e6])
    assert e.variance >= 0.99 * m.total_sill() and np.isfinite(e.value)
    sk = gh.Kriging(c, v, m, gh.KrigingConfig(method=gh.KrigingMethod.Simple(1.74)))
    e = sk.predict([1e6, 1e6])
    assert abs(e.value - 1.74) < 1e-9 and abs(e.variance - m.total_sill()) < 1e-9 and e.lagrange is None
    t = np.array([[0.3, 0.3], [0.9, 0.1], [2.0, 2.0]])
    a = gh.Kriging(c, v, m).predict_many(t)
    b = gh.Kriging(c, v, m, gh.KrigingConfig(max_neighbors=5)).predict_many(t)
    assert np.abs(a[0] - b[0]).max() < 1e-12 and np.abs(a[1] - b[1]).max() < 1e-12
    for cfg in [gh.KrigingConfig(), gh.KrigingConfig(max_neighbors=3),
                gh.KrigingConfig(search_radius=0.75, min_neighbors=3)]:
        check_predict(oracle, c, v, m, cfg, np.vstack([c, t, [[1.5, 1.5]]]))


def test_reference_search_behaviour():
    # kriging.rs:960-1188 octant balancing, anisotropic search (2-D / 3-D)
    m = cs.sample_model()
    c, v = cs.octant_cluster()
    near = gh.Kriging(c, v, m, gh.KrigingConfig(max_neighbors=4)).predict([0.0, 0.0]).value
    octv = gh.Kriging(c, v, m, gh.KrigingConfig(max_neighbors=4, max_per_octant=1)).predict([0.0, 0.0]).value
    assert near > 9.0 and octv < near - 2.0
    ab = np.array([[0.0, 3.0], [1.0, 0.0]])
    iso = gh.Kriging(ab, [100.0, 0.0], m, gh.KrigingConfig(max_neighbors=1)).predict([0, 0]).value
    ani = gh.Kriging(ab, [100.0, 0.0], m, gh.KrigingConfig(
        max_neighbors=1, anisotropic_search=gh.Anisotropy(0.0, 0.2))).predict([0, 0]).value
    assert abs(iso) < 1e-8 and abs(ani - 100.0) < 1e-8
    a3 = np.array([[0.0, 0.0, 4.0], [4.0, 0.0, 0.0]])
    est = gh.Kriging(a3, [100.0, 0.0], m, gh.KrigingConfig(
        max_neighbors=1, anisotropic_search=gh.Anisotropy(0.0, 0.2, 1.0, 90.0, 0.0))).predict([0, 0, 0])
    assert abs(est.value - 100.0) < 1e-6


def test_min_neighbors_error_and_build_errors():
    c, v = cs.sample_data()
    m = cs.sample_model()
    k = gh.Kriging(c, v, m, gh.KrigingConfig(search_radius=0.75, min_neighbors=3))
    k.predict([0.5, 0.5])
    with pytest.raises(gh.NoNeighbors):
        k.predict([1.5, 1.5])
    with pytest.raises(gh.DuplicatePoints) as ei:
        gh.Kriging(np.array([[0.0, 0.0], [1.0, 1.0], [1.0, 1.0]]), [1.0, 2.0, 3.0], m)
    assert (ei.value.i, ei.value.j) == (1, 2)
    with pytest.raises(gh.InvalidParameter):
        gh.Kriging(c, v, m, gh.KrigingConfig(min_neighbors=0))
    with pytest.raises(gh.InvalidParameter):
        gh.Kriging(c, v, m, gh.KrigingConfig(max_neighbors=2, min_neighbors=3))
    with pytest.raises(gh.InvalidParameter):
        gh.Kriging(c, v, m, gh.KrigingConfig(search_radius=0.0))
    with pytest.raises(gh.InvalidParameter):
        gh.Kriging(c, v, m, gh.KrigingConfig(anisotropic_search=gh.Anisotropy(0.0, 0.0)))
    with pytest.raises(gh.InsufficientData):
        gh.Kriging(np.zeros((0, 2)), np.zeros(0), m)
    with pytest.raises(gh.InvalidParameter):   # Power: semivariogram form not on GPU
        gh.Kriging(c, v, gh.VariogramModel(0.0, [gh.Structure(gh.ModelKind.Power(1.2), 2.0, 1.0)]))
    # a neighbourhood above 64 points goes through the large path: the
    # reference's estimate, not a truncated system
    c1, v1 = workloads.points(1, 200, [10.0, 10.0])
    big = gh.Kriging(c1, v1, m, gh.KrigingConfig(max_neighbors=100))
    val, _, _, st = big.predict_many([[5.0, 5.0], [1.0, 2.0]])
    rc, ov, _, _, ost, _ = oracle_mod().predict_many(c1, v1, m, gh.KrigingConfig(max_neighbors=100),
                                                     [[5.0, 5.0], [1.0, 2.0]])
    assert (st == 0).all() and (ost == 0).all() and rel_inf(val, ov) < TOL64


def oracle_mod():
    from oracle import refcpu
    return refcpu


def test_radius_neighbourhood_above_64(oracle):
    # radius-only search keeps every in-radius point (kriging.rs:445-451):
    # targets with more than 64 take the large path, every target matches
    # the oracle
    c, v = workloads.points(13, 20000, [1000.0, 1000.0])
    t = workloads.points(14, 3000, [1000.0, 1000.0])[0]
    m = cs.MODELS_2D["exp"]
    cfg = gh.KrigingConfig(search_radius=30.0)
    _, _, ocnt = oracle.neighbors(c, v, m, cfg, t, 400)
    assert (ocnt > 64).any() and (ocnt <= 64).any()
    check_neighbors(oracle, c, v, m, cfg, t)
    check_predict(oracle, c, v, m, cfg, t)


def test_singular_targets_are_nan_with_status(oracle):
    # Gaussian model, no nugget, near-duplicate points: singular systems
    c = np.array([[0.0, 0.0], [1e-7, 0.0], [5.0, 5.0], [0.0, 1e-7]])
    v = np.array([1.0, 1.1, 2.0, 0.9])
    m = gh.VariogramModel(0.0, [gh.Structure(gh.ModelKind.GAUSSIAN, 1.0, 100.0)])
    k = gh.Kriging(c, v, m, gh.KrigingConfig(max_neighbors=4))
    val, var, lag, st = k.predict_many([[0.5, 0.5], [3.0, 3.0]])
    rc, ov, ovar, olag, ost, _ = oracle.predict_many(c, v, m, gh.KrigingConfig(max_neighbors=4),
                                                     [[0.5, 0.5], [3.0, 3.0]])
    assert (ost == 3).all() and (st == 3).all() and np.isnan(val).all()


@pytest.mark.parametrize("sill", [1e-6, 1e-4, 1e6, 3e6, 1e7, 1e8])
@pytest.mark.parametrize("method", [gh.KrigingMethod.Simple(0.0), gh.KrigingMethod.Ordinary(),
                                    gh.KrigingMethod.Universal(1), gh.KrigingMethod.Universal(2)])
def test_singular_test_follows_matrix_scale(oracle, sill, method):
    # The reference's LU is singular when a pivot falls below 1e-12 of the
    # bordered matrix's largest entry (linalg.rs:36-57): with sills far from
    # 1, the drift block's Schur pivots decide it (every target singular
    # for ordinary kriging from a sill of about 1e7, for universal from
    # about 1e6), and the NaN pattern must be the reference's.
    c, v = workloads.points(5, 300, [1000.0, 1000.0])
    t = workloads.uniforms(9, 400).reshape(200, 2) * 1000.0
    for kind in (gh.ModelKind.SPHERICAL, gh.ModelKind.EXPONENTIAL):
        m = gh.VariogramModel(0.1 * sill, [gh.Structure(kind, sill, 300.0)])
        check_predict(oracle, c, v * np.sqrt(sill), m,
                      gh.KrigingConfig(method=method, max_neighbors=32), t)


# ------------------------------------------- size-independent properties --
def test_full_size_properties():
    # At the full 200k x 4M configuration: determinism across calls, batch
    # independence (a target's result does not depend on what else is in
    # the launch), exactness at data locations.
    w = workloads.ok2d_large()
    k = gh.Kriging(w.coords, w.values, w.model, w.config)
    t = w.targets()
    print(f"full size: T={len(t)}", flush=True)
    v1, var1, _, st1 = k.predict_many(t)
    assert (st1 == 0).all() and np.isfinite(v1).all() and (var1 >= 0).all()
    v2, var2, _, _ = k.predict_many(t)
    assert np.array_equal(v1, v2) and np.array_equal(var1, var2)
    sel = np.random.default_rng(0).choice(len(t), 5000, replace=False)
    v3, var3, _, _ = k.predict_many(t[sel])
    assert np.array_equal(v3, v1[sel]) and np.array_equal(var3, var1[sel])
    at = w.coords[:2000]
    v4, var4, _, _ = k.predict_many(at)
    # nugget 0.05 is part of the covariance at h = 0 only: the data point is
    # reproduced exactly and the variance vanishes
    assert np.abs(v4 - w.values[:2000]).max() < 1e-8 and var4.max() < 1e-8
