"""GPU parity of the 33-64-neighbour matrix-core kriging solve
(krige_mfma64_kernel).

Universal kriging (linear drift in 3-D) with the nested model of the 3-D
octant workload: nugget + exponential + spherical sharing one rotated
anisotropy; and every lean kind as a single structure, isotropic and
anisotropic, with simple, ordinary, linear and quadratic drift in 2-D and
the first three in 3-D.  Every case checks that the call ran the matrix-core kernel
(gs_last_solve_kernel), that it matches the oracle (statuses and NaN
patterns exact, estimates within 1e-10 relative, or at least as close as
the oracle to the exact solution), and that the vector-unit kernel
(GS_KRIGE_MFMA=0) gives the same statuses and agrees within the same bar.
The cases cover octant neighbourhoods of 8 to 64 points (sizes in every
bucket), full 64-point neighbourhoods, radius-limited ones with empty slots
and too few neighbours (status 5), targets on data points, and both ends of
the singular test's scale.
"""
import os

import numpy as np
import pytest

import geostat_hip as gh
from geostat_hip import workloads
from test_gpu_kriging import check_predict, rel_inf

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
ANIS = gh.Anisotropy(azimuth_deg=45.0, ratio=0.6, ratio_z=0.2, dip_deg=10.0, rake_deg=5.0)


class vector_kernels:
    def __enter__(self):
        os.environ["GS_KRIGE_MFMA"] = "0"

    def __exit__(self, *exc):
        os.environ.pop("GS_KRIGE_MFMA", None)


def nested(sill=1.0, r1=60.0, r2=150.0):
    return gh.VariogramModel(0.1 * sill, [gh.Structure(gh.ModelKind.EXPONENTIAL, 0.6 * sill, r1, ANIS),
                                          gh.Structure(gh.ModelKind.SPHERICAL, 0.4 * sill, r2, ANIS)])


def uk_case(oracle, cfg, model=None, n=20000, nt=1500, seed=3, scale=1.0, on_data=40, lifted=0):
    ext = [500.0, 500.0, 50.0]
    c, v = workloads.points(seed, n, ext, recipe=1)
    v = v * scale
    t, _ = workloads.points(seed + 1, nt, ext)
    # targets on data points, and targets lifted above the data (few or no neighbours)
    t = np.concatenate([t, c[:on_data], t[:lifted] + np.array([0.0, 0.0, 70.0])])
    model = model or nested()
    print(f"mfma64 case: n={n} T={len(t)} k={cfg.max_neighbors} oct={cfg.max_per_octant} "
          f"r={cfg.search_radius} seed={seed}", flush=True)
    kr, val, var = check_predict(oracle, c, v, model, cfg, t)
    assert gh._abi.last_solve_kernel() == "krige_mfma64_kernel"
    _, cnt = kr.neighbors(t)
    _, _, lag, st = kr.predict_many(t)
    with vector_kernels():
        v2, var2, lag2, st2 = kr.predict_many(t)
        assert gh._abi.last_solve_kernel() == "krige_fp64_kernel"
    assert np.array_equal(st, st2)
    c0 = model.total_sill()
    assert rel_inf(val, v2, 1e-3 * np.sqrt(c0)) < TOL64
    assert rel_inf(var, var2, 1e-3 * c0) < TOL64
    return cnt, st


def test_octant_neighbourhoods(oracle):
    cfg = gh.KrigingConfig(method=gh.KrigingMethod.Universal(1), max_neighbors=64, min_neighbors=8,
                           max_per_octant=8, search_radius=60.0, anisotropic_search=ANIS)
    cnt, st = uk_case(oracle, cfg)
    assert cnt.min() < 32 < cnt.max() and cnt.max() == 64


def test_full_neighbourhoods(oracle):
    cfg = gh.KrigingConfig(method=gh.KrigingMethod.Universal(1), max_neighbors=64)
    cnt, st = uk_case(oracle, cfg, n=8000, seed=5)
    assert (cnt == 64).all() and (st == 0).all()


def test_radius_empty_slots(oracle):
    # between none and 64 neighbours: empty slots, and status 5 below the minimum
    cfg = gh.KrigingConfig(method=gh.KrigingMethod.Universal(1), max_neighbors=64, min_neighbors=10,
                           search_radius=24.0)
    cnt, st = uk_case(oracle, cfg, n=20000, seed=9, lifted=60)
    assert (cnt < 64).any() and (cnt == 64).any() and (cnt > 32).any()
    assert (st == 5).any()


@pytest.mark.parametrize("sill", [1e-4, 1e8])
def test_scale_of_singular_test(oracle, sill):
    cfg = gh.KrigingConfig(method=gh.KrigingMethod.Universal(1), max_neighbors=48)
    uk_case(oracle, cfg, model=nested(sill), n=6000, nt=600, seed=13, scale=np.sqrt(sill))


LEAN = [("sph", gh.ModelKind.SPHERICAL), ("exp", gh.ModelKind.EXPONENTIAL),
        ("gau", gh.ModelKind.GAUSSIAN), ("mat15", gh.ModelKind.MATERN15),
        ("mat25", gh.ModelKind.MATERN25)]
METHODS = [gh.KrigingMethod.Simple(0.3), gh.KrigingMethod.Ordinary(),
           gh.KrigingMethod.Universal(1), gh.KrigingMethod.Universal(2)]


def single_case(oracle, D, kind, aniso, method, k, seed, radius=None, nugget=0.05):
    ext = [100.0] * D
    c, v = workloads.points(seed, 3000, ext)
    t, _ = workloads.points(seed + 1, 300, ext)
    t = np.concatenate([t, c[:20]])
    an = (gh.Anisotropy(30.0, 0.5) if D == 2 else ANIS) if aniso else None
    m = gh.VariogramModel(nugget, [gh.Structure(kind, 1.0, 25.0, an)])
    cfg = gh.KrigingConfig(method=method, max_neighbors=k, search_radius=radius)
    print(f"mfma64 single: D={D} kind={kind} aniso={aniso} method={method.kind} k={k} seed={seed}",
          flush=True)
    kr, val, var = check_predict(oracle, c, v, m, cfg, t)
    assert gh._abi.last_solve_kernel() == "krige_mfma64_kernel"
    _, _, _, st = kr.predict_many(t)
    with vector_kernels():
        v2, var2, _, st2 = kr.predict_many(t)
        assert gh._abi.last_solve_kernel() == "krige_fp64_kernel"
    assert np.array_equal(st, st2)
    assert rel_inf(val, v2, 1e-3) < TOL64
    assert rel_inf(var, var2, 1e-3) < TOL64


@pytest.mark.parametrize("name,kind", LEAN)
@pytest.mark.parametrize("aniso", [False, True])
def test_2d_kinds_and_methods(oracle, name, kind, aniso):
    # the smooth kinds get a larger nugget, for the conditioning of their
    # 64-point systems (smaller nuggets were not run against the 1e-10 bar)
    nug = 0.05 if kind in (gh.ModelKind.SPHERICAL, gh.ModelKind.EXPONENTIAL) else 0.2
    for i, meth in enumerate(METHODS):
        single_case(oracle, 2, kind, aniso, meth, (64, 48, 40, 64)[i], seed=31 + 2 * i, nugget=nug)


@pytest.mark.parametrize("name,kind", [LEAN[0], LEAN[1], LEAN[4]])
def test_3d_single_structure(oracle, name, kind):
    nug = 0.05 if kind != gh.ModelKind.MATERN25 else 0.2
    for i, meth in enumerate(METHODS[:3]):
        single_case(oracle, 3, kind, i % 2 == 0, meth, (64, 56, 64)[i], seed=51 + 2 * i, nugget=nug)


def test_radius_single_structure(oracle):
    single_case(oracle, 2, gh.ModelKind.MATERN15, True, gh.KrigingMethod.Ordinary(), 64, seed=61,
                radius=9.0)


def test_unsupported_models_stay_on_vector_kernel(oracle):
    # structures with different anisotropies, and 3-D quadratic drift (ten
    # drift rows): no 64-slot matrix-core kernel; the vector kernel solves them
    c, v = workloads.points(21, 4000, [100.0, 100.0, 100.0])
    t, _ = workloads.points(22, 300, [100.0, 100.0, 100.0])
    other = gh.Anisotropy(azimuth_deg=10.0, ratio=0.5, ratio_z=0.5)
    m = gh.VariogramModel(0.1, [gh.Structure(gh.ModelKind.EXPONENTIAL, 0.6, 30.0, ANIS),
                                gh.Structure(gh.ModelKind.SPHERICAL, 0.4, 60.0, other)])
    check_predict(oracle, c, v, m, gh.KrigingConfig(method=gh.KrigingMethod.Universal(1),
                                                      max_neighbors=60), t)
    assert gh._abi.last_solve_kernel() == "krige_fp64_kernel"
    m1 = gh.VariogramModel(0.1, [gh.Structure(gh.ModelKind.SPHERICAL, 1.0, 40.0)])
    check_predict(oracle, c, v, m1, gh.KrigingConfig(method=gh.KrigingMethod.Universal(2),
                                                       max_neighbors=60), t)
    assert gh._abi.last_solve_kernel() == "krige_fp64_kernel"
