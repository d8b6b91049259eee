"""The serving mix at its stated sizes (workloads.serving_requests: 40
requests, 10^3-2*10^5 data points, 10^4-10^6 targets, k in {16, 32, 64}, four
model kinds; every fifth request a Vecchia batch of 8 models).

* Every request run through the request loop with four slots, all submitted
  at once, equals the same request made alone, bit for bit, in FP64 and in
  the FP32-refined precision.
* Each kriging request, on a 2,000-target strided sample: neighbour lists
  bit-exact against the oracle, estimates within 1e-10 (FP64) and 1e-6
  (FP32-refined); each Vecchia request: one model's log-likelihood within
  1e-10 of the oracle's on the same plan.
"""
import concurrent.futures as cf

import numpy as np
import pytest

import geostat_hip as gh
from geostat_hip import workloads

pytestmark = pytest.mark.gpu


def _rel(a, b, floor):
    assert np.array_equal(np.isnan(a), np.isnan(b))
    m = ~np.isnan(b)
    return float(np.abs(a[m] - b[m]).max() / max(np.abs(b[m]).max(), floor)) if m.any() else 0.0


@pytest.fixture(scope="module")
def mix():
    reqs = workloads.serving_requests(40)
    jobs = []
    for i, r in enumerate(reqs):
        if i % 5 == 4:
            c, v = workloads.points(r.seed, r.n, [1.0, 1.0], recipe=0)
            plan = gh.vecchia_plan(c, 30)
            jobs.append(("v", plan, (c, v, workloads.vecchia_models(8, seed=r.seed))))
        else:
            c, v, mdl, cfg, t = workloads.serving_case(r)
            jobs.append(("k", gh.Kriging(c, v, mdl, cfg), (c, v, mdl, cfg, t)))
    big = max(r.targets for r in reqs)
    print(f"serving mix: {len(jobs)} requests, largest {max(r.n for r in reqs)} points, "
          f"{big} targets", flush=True)
    return reqs, jobs


@pytest.mark.parametrize("prec", ["fp64", "fp32r"])
def test_mix_equals_alone(mix, prec):
    _, jobs = mix
    alone = [h.predict_many(a[4], precision=prec) if kind == "k"
             else h.loglik_batch(a[1], a[2], precision=prec) for kind, h, a in jobs]
    srv = gh.Server(n_slots=4, mem_budget=32 << 30)
    tickets = [srv.submit_krige(h, a[4], precision=prec) if kind == "k"
               else srv.submit_vecchia(h, a[1], a[2], precision=prec) for kind, h, a in jobs]
    for tk, want in zip(tickets, alone):
        got = srv.wait(tk)
        for g, w in zip(got, want):
            assert np.array_equal(g, w, equal_nan=True)
    srv.close()


def test_requests_against_oracle(mix, oracle):
    reqs, jobs = mix
    vjobs = []
    for (kind, h, a), r in zip(jobs, reqs):
        if kind == "v":
            vjobs.append((h, a))
            continue
        c, v, mdl, cfg, t = a
        sel = t[:: max(1, len(t) // 2000)]
        idx, cnt = h.neighbors(sel)
        _, oidx, ocnt = oracle.neighbors(c, v, mdl, cfg, sel, idx.shape[1])
        assert np.array_equal(idx, oidx) and np.array_equal(cnt, ocnt), r
        rc, ov, ovar, _, ost, _ = oracle.predict_many(c, v, mdl, cfg, sel)
        for prec, tol in (("fp64", 1e-10), ("fp32r", 1e-6)):
            gv, gvar, _, gst = h.predict_many(sel, precision=prec)
            assert rc == 0 and np.array_equal(gst, ost)
            assert _rel(gv, ov, 1e-3 * np.sqrt(1.05)) < tol, (r, prec)
            assert _rel(gvar, ovar, 1e-3 * 1.05) < tol, (r, prec)

    def one(job):
        plan, (c, v, models) = job
        o, nb, cnt = plan.arrays()
        return plan.loglik_batch(v, models[:1])[0][0], oracle.vecchia_loglik_plan(c, v, models[0], o, nb, cnt)

    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        for got, (rc, want) in ex.map(one, vjobs):
            assert rc == 0 and abs(got - want) <= 1e-10 * abs(want), (got, want)
