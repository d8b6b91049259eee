"""Moving neighbourhoods above 64 points (the large path, gs_big.hip):
max_neighbors above 64, radius-only searches whose disks hold more than 64
points, octant quotas above 64 slots, with the extended systems and both
precisions.  Bars as for the FP64 path: neighbour lists bit-exact against
the oracle, estimates within 1e-10 (or at least as close as the oracle to a
40-digit solution of an ill-conditioned system), NaN / status patterns
equal."""
import numpy as np
import pytest

import _cases as cs
import geostat_hip as gh
from geostat_hip import workloads
from test_gpu_kriging import TOL64, check_neighbors, check_predict, rel_inf

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [65, 100, 200])
@pytest.mark.parametrize("method", [gh.KrigingMethod.Simple(0.2), gh.KrigingMethod.Ordinary(),
                                    gh.KrigingMethod.Universal(1)])
def test_max_neighbors_above_64_2d(oracle, k, method):
    c, v = workloads.points(111, 8000, [1000.0, 1000.0])
    t = workloads.points(112, 600, [1000.0, 1000.0])[0]
    cfg = gh.KrigingConfig(method=method, max_neighbors=k)
    print(f"big 2-D: k={k} method={method.kind}", flush=True)
    m = cs.MODELS_2D["mat15_aniso"]
    check_neighbors(oracle, c, v, m, cfg, t)
    check_predict(oracle, c, v, m, cfg, t)


def test_3d_rotated_general_model_and_radius(oracle):
    c, v = workloads.points(121, 30000, [1000.0, 1000.0, 200.0], recipe=1)
    t = workloads.points(122, 400, [1000.0, 1000.0, 200.0], recipe=1)[0]
    an = gh.Anisotropy(30.0, 0.5, 0.3, 20.0, 40.0)
    m = gh.VariogramModel(0.1, [gh.Structure(gh.ModelKind.Matern(0.8), 0.6, 200.0, an),
                                gh.Structure(gh.ModelKind.SPHERICAL, 0.4, 500.0, an)])
    for cfg in [gh.KrigingConfig(max_neighbors=90, anisotropic_search=an),
                gh.KrigingConfig(method=gh.KrigingMethod.Universal(2), max_neighbors=120),
                gh.KrigingConfig(search_radius=45.0, min_neighbors=3)]:
        print(f"big 3-D: {cfg}", flush=True)
        check_neighbors(oracle, c, v, m, cfg, t)
        check_predict(oracle, c, v, m, cfg, t)


def test_octant_quotas_above_64_slots(oracle):
    c, v = workloads.points(131, 20000, [1000.0, 1000.0, 200.0], recipe=1)
    t = workloads.points(132, 300, [1000.0, 1000.0, 200.0], recipe=1)[0]
    m = cs.MODELS_2D["exp"]
    for cfg in [gh.KrigingConfig(max_per_octant=12, search_radius=120.0),
                gh.KrigingConfig(max_per_octant=20, max_neighbors=100, min_neighbors=4)]:
        print(f"big octant 3-D: {cfg}", flush=True)
        check_neighbors(oracle, c, v, m, cfg, t)
        check_predict(oracle, c, v, m, cfg, t)
    c2, v2 = workloads.points(133, 8000, [1000.0, 1000.0])
    t2 = workloads.points(134, 400, [1000.0, 1000.0])[0]
    cfg = gh.KrigingConfig(max_per_octant=20)
    check_neighbors(oracle, c2, v2, m, cfg, t2)
    check_predict(oracle, c2, v2, m, cfg, t2)


def test_lattice_ties_above_64(oracle):
    c, v = cs.lattice(40, 40)
    t = np.array([[x + 0.5, y + 0.5] for x in range(2, 38, 7) for y in range(2, 38, 5)]
                 + [[x, y] for x in range(0, 40, 9) for y in range(0, 40, 6)], dtype=float)
    for k in (69, 81, 100):
        check_neighbors(oracle, c, v, cs.sample_model(), gh.KrigingConfig(max_neighbors=k), t)


def test_extended_systems_above_64(oracle):
    c, v = workloads.points(141, 5000, [1000.0, 1000.0])
    t = workloads.points(142, 300, [1000.0, 1000.0])[0]
    m = cs.MODELS_2D["sph"]
    err = workloads.uniforms(143, len(c)) * 0.05
    cfg = gh.KrigingConfig(max_neighbors=80)
    k = gh.Kriging.with_measurement_error(c, v, m, cfg, err)
    val, var, lag, st = k.predict_many(t)
    rc, ov, ovar, olag, ost, _ = oracle.predict_many_ex(c, v, m, cfg, t, errors=err)
    assert rc == 0 and np.array_equal(st, ost)
    assert rel_inf(val, ov) < TOL64 and rel_inf(var, ovar, 1e-3) < TOL64
    drift = np.stack([np.sin(c[:, 0] / 90.0), c[:, 1] / 1000.0], axis=1)
    tdrift = np.stack([np.sin(t[:, 0] / 90.0), t[:, 1] / 1000.0], axis=1)
    cfg = gh.KrigingConfig(method=gh.KrigingMethod.ExternalDrift(2), max_neighbors=75)
    k = gh.Kriging.with_external_drift(c, v, m, cfg, drift)
    val, var, lag, st = k.predict_many_with_drift(t, tdrift)
    rc, ov, ovar, olag, ost, _ = oracle.predict_many_ex(c, v, m, cfg, t, drift=drift, tdrift=tdrift)
    assert rc == 0 and np.array_equal(st, ost)
    assert rel_inf(val, ov) < TOL64 and rel_inf(var, ovar, 1e-3) < TOL64
    offs = gh.block_offsets([20.0, 20.0], [3, 3])
    cfg = gh.KrigingConfig(max_neighbors=70)
    k = gh.Kriging(c, v, m, cfg)
    val, var, lag, st = k.predict_block_many(t, offs)
    rc, ov, ovar, olag, ost, _ = oracle.predict_many_ex(c, v, m, cfg, t, offsets=offs, block=True)
    assert rc == 0 and np.array_equal(st, ost)
    assert rel_inf(val, ov) < TOL64 and rel_inf(var, ovar, 1e-3) < TOL64


def test_mixed_sizes_precisions_and_limits(oracle):
    # a radius search where most targets stay at or below 64 points: both
    # paths in one call, FP32-refined for the small systems
    c, v = workloads.points(151, 30000, [1000.0, 1000.0])
    t = workloads.points(152, 2000, [1000.0, 1000.0])[0]
    m = cs.MODELS_2D["exp"]
    cfg = gh.KrigingConfig(search_radius=22.0)
    k = gh.Kriging(c, v, m, cfg)
    _, cnt = k.neighbors(t)
    assert (cnt > 64).any() and (cnt <= 64).sum() > len(t) // 2
    a = k.predict_many(t, precision="fp64")
    b = k.predict_many(t, precision="fp32r")
    rc, ov, ovar, _, ost, _ = oracle.predict_many(c, v, m, cfg, t)
    assert np.array_equal(a[3], ost) and np.array_equal(b[3], ost)
    big = cnt > 64
    assert np.array_equal(a[0][big], b[0][big])          # the large path is FP64
    assert rel_inf(b[0], ov, 1e-3) < 1e-6 and rel_inf(b[1], ovar, 1e-3) < 1e-6
    assert k.max_neighbors == 4096
    # the request loop takes the large path too
    srv = gh.Server(n_slots=2)
    got = srv.wait(srv.submit_krige(k, t))
    srv.close()
    for g, w in zip(got, a):
        assert np.array_equal(g, w, equal_nan=True)
