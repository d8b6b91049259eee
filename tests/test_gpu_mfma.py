"""GPU parity of the matrix-core solves (krige_mfma_kernel, vecchia_mfma_kernel,
vecchia_grad_mfma_kernel).

Simple / ordinary kriging with 17-32 neighbours and Vecchia conditioning
sets of at most 32 points, one lean structure: every case checks that the
call ran the matrix-core kernel (gs_last_solve_kernel), that it matches the
oracle (statuses and NaN patterns exact, estimates within 1e-10 relative),
and that the vector-unit kernel (GS_KRIGE_MFMA=0 / GS_VECCHIA_MFMA=0) gives
the same statuses and agrees within the same bar (the likelihood gradient:
log-likelihood 1e-10, gradient 1e-8 normwise).  The cases cover every lean
kind, 2-D and 3-D, isotropic and anisotropic structures, neighbourhoods
that leave slots empty (search radius), targets on data points, and systems
the singular test rejects.
"""
import os

import numpy as np
import pytest

import geostat_hip as gh
from geostat_hip import workloads
from test_gpu_kriging import check_predict, rel_inf

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
LEAN = [("sph", gh.ModelKind.SPHERICAL), ("exp", gh.ModelKind.EXPONENTIAL),
        ("gau", gh.ModelKind.GAUSSIAN), ("mat15", gh.ModelKind.MATERN15),
        ("mat25", gh.ModelKind.MATERN25)]
ANIS2 = gh.Anisotropy(30.0, 0.5)
ANIS3 = gh.Anisotropy(azimuth_deg=45.0, ratio=0.6, ratio_z=0.3, dip_deg=10.0, rake_deg=5.0)


class vector_kernels:
    """Runs the block with the matrix-core kernels switched off."""

    def __enter__(self):
        os.environ["GS_KRIGE_MFMA"] = "0"
        os.environ["GS_VECCHIA_MFMA"] = "0"

    def __exit__(self, *exc):
        os.environ.pop("GS_KRIGE_MFMA", None)
        os.environ.pop("GS_VECCHIA_MFMA", None)


def kriging_case(oracle, D, kind, aniso, method, k, radius=None, sill=1.0, nugget=0.05, seed=7):
    ext = [100.0] * D
    c, v = workloads.points(seed, 3000, ext)
    an = (ANIS2 if D == 2 else ANIS3) if aniso else None
    model = gh.VariogramModel(nugget * sill, [gh.Structure(kind, sill, 25.0, an)])
    cfg = gh.KrigingConfig(method=method, max_neighbors=k, search_radius=radius,
                           min_neighbors=None)
    t, _ = workloads.points(seed + 1, 400, ext)
    t = np.concatenate([t, c[:25]])            # 25 targets on data points
    kr, val, var = check_predict(oracle, c, v, model, cfg, t)
    assert gh._abi.last_solve_kernel() == "krige_mfma_kernel"
    _, _, _, st = kr.predict_many(t)
    with vector_kernels():
        v2, var2, _, st2 = kr.predict_many(t)
        assert gh._abi.last_solve_kernel() == "krige_fp64_kernel"
    assert np.array_equal(st, st2)
    c0 = model.total_sill()
    assert rel_inf(val, v2, 1e-3 * np.sqrt(c0)) < TOL64
    assert rel_inf(var, var2, 1e-3 * c0) < TOL64
    return st


@pytest.mark.parametrize("name,kind", LEAN)
@pytest.mark.parametrize("aniso", [False, True])
def test_kriging_2d_kinds(oracle, name, kind, aniso):
    kriging_case(oracle, 2, kind, aniso, gh.KrigingMethod.Ordinary(), 32)
    kriging_case(oracle, 2, kind, aniso, gh.KrigingMethod.Simple(0.3), 20, seed=11)


@pytest.mark.parametrize("name,kind", [LEAN[0], LEAN[3], LEAN[4]])
def test_kriging_3d(oracle, name, kind):
    kriging_case(oracle, 3, kind, True, gh.KrigingMethod.Ordinary(), 24, seed=13)
    kriging_case(oracle, 3, kind, False, gh.KrigingMethod.Simple(-0.2), 32, seed=17)


def test_kriging_empty_slots(oracle):
    # a search radius leaves between the minimum and 32 neighbours per target
    # (the rest of the 32 slots empty); targets beyond the data get status 5
    kriging_case(oracle, 2, gh.ModelKind.MATERN15, True, gh.KrigingMethod.Ordinary(), 32,
                 radius=6.0)
    kriging_case(oracle, 2, gh.ModelKind.SPHERICAL, False, gh.KrigingMethod.Simple(1.0), 32,
                 radius=5.0, seed=23)


@pytest.mark.parametrize("sill", [1e-4, 1e8])
def test_kriging_scale_of_singular_test(oracle, sill):
    # the singular test's scale is max(c0, 1) with a drift: at sill 1e8 the
    # reference's LU rejects every one of these ordinary systems (the drift
    # block's pivot), at 1e-4 it solves them
    c, v = workloads.points(5, 300, [1000.0, 1000.0])
    t = workloads.uniforms(9, 400).reshape(200, 2) * 1000.0
    m = gh.VariogramModel(0.1 * sill, [gh.Structure(gh.ModelKind.SPHERICAL, sill, 300.0)])
    kr, _, _ = check_predict(oracle, c, v * np.sqrt(sill), m,
                             gh.KrigingConfig(method=gh.KrigingMethod.Ordinary(), max_neighbors=32), t)
    assert gh._abi.last_solve_kernel() == "krige_mfma_kernel"
    _, _, _, st = kr.predict_many(t)
    assert (st == 3).all() if sill > 1 else (st == 0).all()
    with vector_kernels():
        _, _, _, st2 = kr.predict_many(t)
    assert np.array_equal(st, st2)


def vecchia_case(oracle, D, kind, aniso, m, seed=41):
    ext = [1.0] * D
    c, _ = workloads.points(seed, 2500, ext, recipe=2)
    p = gh.vecchia_plan(c, m)
    o, nb, cnt = p.arrays()
    vals = workloads.vecchia_field(c, o, nb, cnt, m, seed=seed + 1)
    an = (ANIS2 if D == 2 else ANIS3) if aniso else None
    params = ((0.05, 1.0, 0.1), (0.0, 2.0, 0.05), (0.2, 0.5, 0.3), (0.01, 1.5, 0.02))
    if kind in (gh.ModelKind.GAUSSIAN, gh.ModelKind.MATERN15, gh.ModelKind.MATERN25):
        # without a nugget the smooth kinds' conditionals are ill-conditioned
        # enough that any two FP64 evaluation orders differ by more than the
        # bar (at (0, 2, 0.05): Matern 3/2 1.1e-10 from the oracle here,
        # 1.2e-10 by the vector kernel; Matern 5/2 1.5e-9 and 2.4e-9)
        params = tuple((max(nug, 0.02), s, r) for nug, s, r in params)
    models = [gh.VariogramModel(nug, [gh.Structure(kind, s, r, an)]) for nug, s, r in params]
    ll, st = p.loglik_batch(vals, models)
    assert gh._abi.last_solve_kernel() == "vecchia_mfma_kernel"
    for i, md in enumerate(models):
        rc, oll = oracle.vecchia_loglik_plan(c, vals, md, o, nb, cnt)
        assert st[i] == rc, (i, st[i], rc)
        if rc == 0:
            assert abs(ll[i] - oll) <= TOL64 * abs(oll), (i, ll[i], oll)
    assert (st == 0).any()
    with vector_kernels():
        ll2, st2 = p.loglik_batch(vals, models)
        assert gh._abi.last_solve_kernel() == "vecchia_kernel"
    assert np.array_equal(st, st2)
    ok = st == 0
    assert np.all(np.abs(ll[ok] - ll2[ok]) <= TOL64 * np.abs(ll2[ok]))


@pytest.mark.parametrize("name,kind", LEAN)
@pytest.mark.parametrize("aniso", [False, True])
def test_vecchia_2d_kinds(oracle, name, kind, aniso):
    vecchia_case(oracle, 2, kind, aniso, 30)


@pytest.mark.parametrize("m", [17, 24, 32])
def test_vecchia_set_sizes_and_3d(oracle, m):
    vecchia_case(oracle, 2, gh.ModelKind.MATERN25, False, m, seed=43 + m)
    vecchia_case(oracle, 3, gh.ModelKind.EXPONENTIAL, True, m, seed=47 + m)


GRAD_TOL = 1e-8      # normwise relative, floor 1 (as the vector kernel's gradient tests)


def gradient_case(oracle, D, kind, m, seed):
    c, _ = workloads.points(seed, 2000, [1.0] * D, recipe=2)
    p = gh.vecchia_plan(c, m)
    o, nb, cnt = p.arrays()
    vals = workloads.vecchia_field(c, o, nb, cnt, m, seed=seed + 1)
    params = [(0.01, 1.0, 0.05), (0.2, 0.7, 0.12), (0.0, 1.3, 0.03), (0.5, 1e-3, 0.08), (0.05, 2.0, 0.2)]
    if kind == gh.ModelKind.GAUSSIAN:      # a nugget-free Gaussian model is numerically singular
        params = [(max(nug, 0.02), s, r) for nug, s, r in params]
    models = [gh.VariogramModel(nug, [gh.Structure(kind, s, r)]) for nug, s, r in params]
    ll, g, st = p.loglik_grad_batch(vals, models)
    assert gh._abi.last_solve_kernel() == "vecchia_grad_mfma_kernel"
    assert (st == 0).all(), st
    for i, md in enumerate(models):
        s0 = md.structures[0]
        rc, oll, og = oracle.vecchia_loglik_grad_plan(c, vals, s0.kind[0], md.nugget, s0.sill,
                                                      s0.range, o, nb, cnt)
        assert rc == 0
        assert abs(ll[i] - oll) <= TOL64 * abs(oll), (i, ll[i], oll)
        assert np.max(np.abs(g[i] - og)) <= GRAD_TOL * max(np.max(np.abs(og)), 1.0), (i, g[i], og)
    with vector_kernels():
        llv, gv, stv = p.loglik_grad_batch(vals, models)
        assert gh._abi.last_solve_kernel() == "vecchia_grad_kernel"
    assert np.array_equal(st, stv)
    assert np.all(np.abs(ll - llv) <= TOL64 * np.abs(llv))
    for i in range(len(models)):
        assert np.max(np.abs(g[i] - gv[i])) <= GRAD_TOL * max(np.max(np.abs(gv[i])), 1.0)


@pytest.mark.parametrize("name,kind", LEAN)
def test_vecchia_gradient_kinds(oracle, name, kind):
    gradient_case(oracle, 2, kind, 30, 61)


@pytest.mark.parametrize("m", [17, 32])
def test_vecchia_gradient_set_sizes_and_3d(oracle, m):
    gradient_case(oracle, 2, gh.ModelKind.MATERN25, m, 71 + m)
    gradient_case(oracle, 3, gh.ModelKind.SPHERICAL, m, 73 + m)
