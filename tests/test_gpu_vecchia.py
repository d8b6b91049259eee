"""GPU parity of the Vecchia plan and log-likelihood against the oracle.

Bars: plan (ordering and conditioning sets) bit-exact; FP64 likelihood
within 1e-10 relative of the oracle (the reference's serial sum and the
GPU's fixed-tree sum differ only by rounding).
"""
import numpy as np
import pytest

import _cases as cs
import geostat_hip as gh
from geostat_hip import workloads

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def plan_equal(oracle, coords, m, order=None):
    p = gh.vecchia_plan(coords, m, order)
    o, nb, cnt = p.arrays()
    rc, oo, onb, ocnt = oracle.vecchia_plan(coords, m, order)
    assert rc == 0
    assert np.array_equal(o, oo)
    assert np.array_equal(cnt, ocnt)
    assert np.array_equal(nb, onb)
    return p, o, nb, cnt


def test_plan_and_loglik_random(oracle):
    c, v = workloads.points(61, 3000, [1.0, 1.0], recipe=2)
    p, o, nb, cnt = plan_equal(oracle, c, 30)
    vals = workloads.vecchia_field(c, o, nb, cnt, 30, seed=62)
    models = workloads.vecchia_models(8)
    ll, st = p.loglik_batch(vals, models)
    assert (st == 0).all()
    for i, m in enumerate(models):
        rc, oll = oracle.vecchia_loglik_plan(c, vals, m, o, nb, cnt)
        assert rc == 0 and abs(ll[i] - oll) <= TOL64 * abs(oll), (i, ll[i], oll)


def test_plan_with_ties_and_given_order(oracle):
    c, _ = cs.lattice(23, 19)
    plan_equal(oracle, c, 12)
    order = np.random.default_rng(3).permutation(len(c))
    plan_equal(oracle, c, 7, order)


def test_plan_3d(oracle):
    c, v = workloads.points(63, 2500, [100.0, 100.0, 30.0], recipe=1)
    p, o, nb, cnt = plan_equal(oracle, c, 20)
    m = gh.VariogramModel(0.05, [gh.Structure(gh.ModelKind.EXPONENTIAL, 1.0, 40.0,
                                              gh.Anisotropy(30.0, 0.5, 0.3, 20.0, 40.0))])
    ll = p.loglik(v, m)
    rc, oll = oracle.vecchia_loglik_plan(c, v, m, o, nb, cnt)
    assert rc == 0 and abs(ll - oll) <= TOL64 * abs(oll)


def test_full_conditioning_equals_dense_mvn(oracle):
    # vecchia.rs:1997-2009
    c, v = cs.vecchia_field(oracle, 14, 3)
    m = gh.VariogramModel(0.05, [gh.Structure(gh.ModelKind.EXPONENTIAL, 1.0, 40.0)])
    ll = gh.vecchia_loglik(c, v, m, 13)
    rc, ex = oracle.dense_mvn_loglik(c, v, m)
    assert abs(ll - ex) < 1e-9


@pytest.mark.parametrize("mname", ["sph", "exp", "mat15_aniso", "mat25", "matern_nu", "nested"])
def test_model_kinds_loglik(oracle, mname):
    c, v = workloads.points(64, 1500, [1000.0, 1000.0])
    m = cs.MODELS_2D[mname]
    p, o, nb, cnt = plan_equal(oracle, c, 25)
    ll = p.loglik(v, m)
    rc, oll = oracle.vecchia_loglik_plan(c, v, m, o, nb, cnt)
    assert rc == 0 and abs(ll - oll) <= TOL64 * abs(oll)


def test_batch_equals_singles_and_is_deterministic():
    c, v = workloads.points(65, 4000, [1.0, 1.0], recipe=2)
    p = gh.vecchia_plan(c, 30)
    o, nb, cnt = p.arrays()
    vals = workloads.vecchia_field(c, o, nb, cnt, 30, seed=66)
    models = workloads.vecchia_models(70)          # > one launch of 64
    ll, st = p.loglik_batch(vals, models)
    ll2, _ = p.loglik_batch(vals, models)
    assert np.array_equal(ll, ll2)
    for i in (0, 17, 63, 64, 69):
        assert p.loglik(vals, models[i]) == ll[i]


def test_singular_and_rejections(oracle):
    c = np.array([[0.0, 0.0], [1e-9, 0.0], [0.0, 1e-9], [1.0, 1.0], [2.0, 0.5]])
    v = np.array([0.1, 0.2, 0.3, 0.4, 0.5])
    m = gh.VariogramModel(0.0, [gh.Structure(gh.ModelKind.GAUSSIAN, 1.0, 10.0)])
    p = gh.vecchia_plan(c, 4)
    ll, st = p.loglik_batch(v, [m])
    o, nb, cnt = p.arrays()
    rc, _ = oracle.vecchia_loglik_plan(c, v, m, o, nb, cnt)
    assert rc == 3 and st[0] == 3 and np.isnan(ll[0])
    with pytest.raises(gh.SingularSystem):
        gh.vecchia_loglik(c, v, m, 4)
    with pytest.raises(gh.InvalidParameter):
        gh.vecchia_loglik(c, v, gh.VariogramModel(0.0, [gh.Structure(gh.ModelKind.Power(1.0), 1.0, 1.0)]), 3)
    with pytest.raises(gh.InsufficientData):
        gh.vecchia_plan(c[:1], 3)
    with pytest.raises(gh.InvalidParameter):
        gh.vecchia_plan(c, 0)
    with pytest.raises(gh.DuplicatePoints) as ei:
        gh.vecchia_plan(np.vstack([c, c[:1]]), 3)
    assert (ei.value.i, ei.value.j) == (0, 5)
