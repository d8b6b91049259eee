"""Request loop (gs_server_*): concurrent requests over resident predictors
and plans give exactly the results of the same requests made alone."""
import numpy as np
import pytest

import geostat_hip as gh
from geostat_hip import workloads

pytestmark = pytest.mark.gpu


def _mix():
    reqs = []
    for i, (n, k, kind) in enumerate([(20000, 16, gh.ModelKind.SPHERICAL),
                                      (5000, 32, gh.ModelKind.MATERN15),
                                      (60000, 64, gh.ModelKind.EXPONENTIAL)]):
        c, v = workloads.points(90 + i, n, [10000.0, 10000.0])
        m = gh.VariogramModel(0.05, [gh.Structure(kind, 1.0, 800.0)])
        reqs.append(gh.Kriging(c, v, m, gh.KrigingConfig(max_neighbors=k)))
    c, v = workloads.points(95, 300, [1000.0, 1000.0])
    m = gh.VariogramModel(0.05, [gh.Structure(gh.ModelKind.MATERN25, 1.0, 200.0)])
    reqs.append(gh.Kriging(c, v, m, gh.KrigingConfig()))          # global neighbourhood
    return reqs


def test_mix_equals_alone():
    handles = _mix()
    cv, vv = workloads.points(96, 20000, [1.0, 1.0])
    plan = gh.vecchia_plan(cv, 20)
    models = workloads.vecchia_models(8)
    jobs = []
    for j in range(14):
        h = handles[j % len(handles)]
        ext = 1000.0 if h.n == 300 else 10000.0
        t = workloads.points(200 + j, 3000 + 7000 * (j % 3), [ext, ext])[0]
        jobs.append(("k", h, t))
        if j % 4 == 1:
            jobs.append(("v", plan, models))
    alone = []
    for kind, h, arg in jobs:
        alone.append(h.predict_many(arg) if kind == "k" else h.loglik_batch(vv, arg))
    srv = gh.Server(n_slots=4, mem_budget=8 << 30)
    tickets = [srv.submit_krige(h, arg) if kind == "k" else srv.submit_vecchia(h, vv, arg)
               for kind, h, arg in jobs]
    for tk, want in zip(tickets, alone):
        got = srv.wait(tk)
        assert tk.queue_ms >= 0.0 and tk.service_ms > 0.0
        for g, w in zip(got, want):
            assert np.array_equal(g, w, equal_nan=True)
    srv.close()


def test_tiny_budget_serializes_and_errors_propagate():
    handles = _mix()
    srv = gh.Server(n_slots=3, mem_budget=1)
    t = workloads.points(300, 5000, [10000.0, 10000.0])[0]
    want = [h.predict_many(t) for h in handles[:3]]
    tickets = [srv.submit_krige(h, t) for h in handles[:3]]
    for tk, w in zip(tickets, want):
        got = srv.wait(tk)
        for g, x in zip(got, w):
            assert np.array_equal(g, x, equal_nan=True)
    cv, vv = workloads.points(97, 2000, [1.0, 1.0])
    plan = gh.vecchia_plan(cv, 10)
    bad = [gh.VariogramModel(0.0, [gh.Structure(gh.ModelKind.Power(1.0), 1.0, 1.0)])]
    tk = srv.submit_vecchia(plan, vv, bad)
    with pytest.raises(gh.InvalidParameter):
        srv.wait(tk)
    with pytest.raises(gh.InvalidParameter):
        gh.Server(n_slots=0)
    srv.close()


def test_random_mix_with_extended_and_octant_predictors():
    # 60 requests over six resident predictors (2-D and 3-D, octant search,
    # measurement error, general-class kinds, a global neighbourhood) and two
    # plans, submitted at once to five slots: every result equals the same
    # request made alone, bit for bit.
    r = np.random.default_rng(17)
    preds = []
    c2, v2 = workloads.points(401, 8000, [5000.0, 5000.0])
    c3, v3 = workloads.points(402, 20000, [5000.0, 5000.0, 500.0], recipe=1)
    m_gen = gh.VariogramModel(0.1, [gh.Structure(gh.ModelKind.Stable(1.5), 0.8, 600.0),
                                    gh.Structure(gh.ModelKind.HOLE, 0.3, 900.0)])
    m_lean = gh.VariogramModel(0.05, [gh.Structure(gh.ModelKind.MATERN25, 1.0, 700.0,
                                                   gh.Anisotropy(30.0, 0.6, 0.4, 5.0, 0.0))])
    preds.append((gh.Kriging(c2, v2, m_gen, gh.KrigingConfig(max_neighbors=12)), 2))
    preds.append((gh.Kriging.with_measurement_error(c2, v2, m_gen, gh.KrigingConfig(max_neighbors=6),
                                                    np.full(8000, 0.03)), 2))
    preds.append((gh.Kriging(c3, v3, m_lean, gh.KrigingConfig(method=gh.KrigingMethod.Universal(1),
                                                               max_per_octant=6, min_neighbors=6)), 3))
    preds.append((gh.Kriging.with_measurement_error(c3, v3, m_gen,
                                                    gh.KrigingConfig(max_per_octant=4, min_neighbors=3),
                                                    np.full(20000, 0.02)), 3))
    m_sph = gh.VariogramModel(0.05, [gh.Structure(gh.ModelKind.SPHERICAL, 1.0, 2000.0)])
    preds.append((gh.Kriging(c2[:400], v2[:400], m_sph, gh.KrigingConfig()), 2))
    preds.append((gh.Kriging(c3, v3, m_gen, gh.KrigingConfig(max_neighbors=40)), 3))
    cp, _ = workloads.points(403, 10000, [1.0, 1.0])
    plans = [gh.vecchia_plan(cp, 10), gh.vecchia_plan(cp, 40)]
    vp = np.sin(cp[:, 0] * 9.0) + np.cos(cp[:, 1] * 7.0)
    models = [gh.VariogramModel(0.02, [gh.Structure(kd, 1.0, 0.1)])
              for kd in (gh.ModelKind.MATERN15, gh.ModelKind.Matern(0.8), gh.ModelKind.EXPONENTIAL)]
    jobs = []
    for j in range(60):
        if j % 6 == 5:
            jobs.append(("v", plans[(j // 6) % 2], models))
            continue
        h, D = preds[int(r.integers(len(preds)))]
        ext = np.array([5000.0, 5000.0, 500.0][:D])
        jobs.append(("k", h, r.uniform(-0.05, 1.05, size=(int(r.integers(1, 20000)), D)) * ext))
    print(f"server random mix: {len(jobs)} requests, 5 slots", flush=True)
    alone = [h.predict_many(a) if kind == "k" else h.loglik_batch(vp, a) for kind, h, a in jobs]
    srv = gh.Server(n_slots=5, mem_budget=16 << 30)
    tickets = [srv.submit_krige(h, a) if kind == "k" else srv.submit_vecchia(h, vp, a)
               for kind, h, a in jobs]
    for tk, want in zip(tickets, alone):
        got = srv.wait(tk)
        for g, w in zip(got, want):
            assert np.array_equal(g, w, equal_nan=True)
    srv.close()
