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
