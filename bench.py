"""Throughput benchmark: moving-neighbourhood kriging on one MI355X.

A step is one predict_many pass over the whole synthetic target grid, with
data, search index and targets already resident in HBM (neighbour search +
covariance build + solves).  Default workload: ok2d_large (n = 200,000 data
points, Matern 3/2 with geometric anisotropy, k = 32, 2000x2000 = 4*10^6
targets).  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                  [--precision fp64|fp32r] [--no-cpu-baseline]

Other workloads (one JSON line each, not the headline):
  --workload ok2d_small | uk3d_octant   the other kriging configurations
  --workload vecchia_c3                 Vecchia log-likelihood, n = 10^6 points,
                                        m = 30, 64 Matern-5/2 parameter vectors per
                                        batch (plan built once, timed separately)
  --workload serving                    request mix (build + predict per request),
                                        p50 / p99 latency under Poisson arrivals
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "geostat-rs_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

METRIC = ("kriged targets/s at k=32 (FP32-refined vs FP64, 1e-6 rel err); "
          "p50/p99 req latency")
FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 (vector and MFMA)
HBM_PEAK_GBS = 8000.0


def kriging_flops_per_target(n: int, m: int, c: int) -> float:
    """Algorithmic flops per target: E*c + (2/3) N^3 + 2 N^2 + 2n + 2N + 1,
    E = n(n+1)/2, N = n + m (LINPACK solve count for the bordered system)."""
    N = n + m
    E = n * (n + 1) / 2
    return E * c + (2.0 / 3.0) * N ** 3 + 2 * N * N + 2 * n + 2 * N + 1


# flops per covariance evaluation (+ - * / sqrt exp each 1), model matrix
# precomputed: spherical iso 2-D 16, Matern 3/2 + 2-D anisotropy 23,
# nugget + exponential + spherical 3-D rotated 60.
COV_FLOPS = {"ok2d_small": 16, "ok2d_large": 23, "uk3d_octant": 60}


def cpu_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(w, targets, budget_s: float = 10.0):
    """The oracle (C restatement of the reference's rayon path) on host
    cores over a strided sample of the same targets."""
    from oracle import refcpu
    import numpy as np
    th = cpu_threads()
    refcpu.lib()
    t0 = time.perf_counter()
    refcpu.predict_many(w.coords, w.values, w.model, w.config, targets[:1], threads=th)
    build_s = time.perf_counter() - t0
    probe = targets[:: max(1, len(targets) // 20000)]
    t0 = time.perf_counter()
    refcpu.predict_many(w.coords, w.values, w.model, w.config, probe, threads=th)
    dt = max(time.perf_counter() - t0 - build_s, 1e-6)
    rate = len(probe) / dt
    n_s = int(min(len(targets), max(len(probe), rate * budget_s)))
    sample = targets[:: max(1, len(targets) // n_s)]
    t0 = time.perf_counter()
    refcpu.predict_many(w.coords, w.values, w.model, w.config, sample, threads=th)
    dt = max(time.perf_counter() - t0 - build_s, 1e-6)
    return {"value": len(sample) / dt, "unit": "targets/s", "cores": th, "kind": "port",
            "sample": f"{len(sample)} of {len(targets)} grid targets (strided), "
                      f"oracle C restatement, {th} threads, kd-tree build excluded "
                      f"({build_s:.2f} s)"}


def vecchia_flops_per_point(s: int, c: int) -> float:
    """E*c + s^3/3 + 2 s^2 + 4 s + 6 with E = s(s+1)/2 (s conditioning points)."""
    return s * (s + 1) / 2 * c + s ** 3 / 3.0 + 2 * s * s + 4 * s + 6


def bench_vecchia(args) -> dict:
    import numpy as np
    import geostat_hip as gh
    from geostat_hip import workloads
    m = 30
    coords = workloads.vecchia_c3(args.points or 1_000_000)
    n = coords.shape[0]
    t0 = time.perf_counter()
    plan = gh.vecchia_plan(coords, m)
    plan_s = time.perf_counter() - t0
    order, nb, cnt = plan.arrays()
    vals = workloads.vecchia_field(coords, order, nb, cnt, m, threads=cpu_threads())
    models = workloads.vecchia_models(64)
    for _ in range(args.warmup):
        plan.loglik_batch(vals, models)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ll, st = plan.loglik_batch(vals, models)
    dt = (time.perf_counter() - t0) / args.steps
    grad_ms = None
    if args.grad:
        gmodels = workloads.vecchia_models(64)
        plan.loglik_grad_batch(vals, gmodels)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            _, _, gst = plan.loglik_grad_batch(vals, gmodels)
        grad_ms = (time.perf_counter() - t0) / args.steps * 1e3
    mle = None
    if args.mle:
        t0 = time.perf_counter()
        fit = plan.mle(vals, gh.ModelKind.MATERN25)
        s = fit.model.structures[0]
        mle = {"seconds": time.perf_counter() - t0, "kind": "matern25", "loglik": fit.loglik,
               "nugget": fit.model.nugget, "psill": s.sill, "range": s.range}
    hist = np.bincount(cnt, minlength=m + 1)
    flops = 64 * float(sum(hist[s] * vecchia_flops_per_point(s, 19) for s in range(m + 1) if hist[s]))
    return {
        "metric": "Vecchia conditionals/s (points x parameter vectors)",
        "value": n * 64 / dt, "unit": "conditionals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded, sequential Vecchia draw over the plan)",
        "config": {"workload": "vecchia_c3", "points": int(n), "m": m, "models": 64,
                   "plan_build_s": plan_s, "ok_models": int((st == 0).sum()),
                   "grad_ms_per_batch": grad_ms, "mle_fit": mle},
        "roofline": {"bound": "mfma", "achieved": flops / dt / 1e12, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": flops / dt / 1e12 / FP64_PEAK_TFLOPS,
                     "traffic": None, "kernel": "vecchia_kernel", "flops_per_launch": flops},
        "cpu_baseline": None,
    }


def bench_serving(args) -> dict:
    """Each request builds its handle (host kd-tree + cell grid, upload)
    and predicts its targets (host in, host out).  Requests are served FIFO
    by one worker; the measured service times are replayed against Poisson
    arrivals at a load of 0.5 (arrival rate = 0.5 / mean service time)."""
    import numpy as np
    import geostat_hip as gh
    from geostat_hip import workloads
    reqs = workloads.serving_requests(args.requests)
    ctx = gh.Context(0)
    # warm the kernels of every k and kind once
    for k in (16, 32, 64):
        for kind in workloads.SERVING_KINDS:
            r = workloads.ServingRequest(0.0, 2000, 10_000, k, kind, 7)
            c, v, mdl, cfg, t = workloads.serving_case(r)
            gh.Kriging(c, v, mdl, cfg, ctx=ctx).predict_many(t)
    service, sizes = [], []
    for r in reqs:
        c, v, mdl, cfg, t = workloads.serving_case(r)
        t0 = time.perf_counter()
        kr = gh.Kriging(c, v, mdl, cfg, ctx=ctx)
        est = kr.predict_many(t)
        service.append(time.perf_counter() - t0)
        sizes.append((r.n, r.targets, r.k))
        del kr, est
    service = np.array(service)
    mean_gap = service.mean() / 0.5
    arrival = np.cumsum([r.arrival_gap * mean_gap for r in reqs])
    finish = np.empty(len(reqs))
    free = 0.0
    for i in range(len(reqs)):
        start = max(arrival[i], free)
        finish[i] = start + service[i]
        free = finish[i]
    lat = (finish - arrival) * 1e3
    total_targets = sum(s[1] for s in sizes)
    return {
        "metric": "serving request latency p50/p99 (build + predict per request)",
        "value": float(np.percentile(lat, 50)), "unit": "ms", "n_gpus": 1,
        "steps": len(reqs), "warmup": 12, "ms_per_step": float(service.mean() * 1e3),
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic request mix (seeded)",
        "config": {"workload": "serving", "requests": len(reqs), "load": 0.5,
                   "p50_ms": float(np.percentile(lat, 50)),
                   "p99_ms": float(np.percentile(lat, 99)),
                   "service_p50_ms": float(np.percentile(service, 50) * 1e3),
                   "service_p99_ms": float(np.percentile(service, 99) * 1e3),
                   "targets_per_s_in_service": float(total_targets / service.sum())},
        "roofline": None, "cpu_baseline": None,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="ok2d_large",
                    choices=["ok2d_large", "ok2d_small", "uk3d_octant", "vecchia_c3", "serving"])
    ap.add_argument("--points", type=int, default=0, help="vecchia_c3: number of points")
    ap.add_argument("--requests", type=int, default=40, help="serving: number of requests")
    ap.add_argument("--grad", action="store_true",
                    help="vecchia_c3: also time the 64-model log-likelihood + gradient batch")
    ap.add_argument("--mle", action="store_true",
                    help="vecchia_c3: also time one maximum-likelihood fit (Matern 5/2)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32r"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--host-io", action="store_true",
                    help="also time predict_many on host arrays (PCIe-inclusive rate, reported "
                         "as config.host_io_targets_per_s; never the headline value)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB three times "
                         "(scales rocprof FETCH_SIZE / WRITE_SIZE to bytes)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "launches": 3}), flush=True)
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(targets shard trivially; see DESIGN.md)", file=sys.stderr)
        return 2

    if args.workload in ("vecchia_c3", "serving"):
        fn = bench_vecchia if args.workload == "vecchia_c3" else bench_serving
        print(json.dumps(fn(args)), flush=True)
        return 0

    import numpy as np
    import torch

    import geostat_hip as gh
    from geostat_hip import workloads

    dev = torch.device("cuda:0")
    w = getattr(workloads, args.workload)()
    targets = w.targets()
    T = targets.shape[0]
    k = gh.Kriging(w.coords, w.values, w.model, w.config)
    D = targets.shape[1]
    d_t = torch.from_numpy(targets).to(dev)
    d_v = torch.empty(T, dtype=torch.float64, device=dev)
    d_var = torch.empty(T, dtype=torch.float64, device=dev)
    d_lag = torch.empty(T, dtype=torch.float64, device=dev)
    d_st = torch.empty(T, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream

    def step():
        k.predict_many_device(d_t, d_v, d_var, d_lag, d_st, precision=args.precision, stream=sh)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    k.timing_summary()                 # drop warmup marks (timing off so far)
    k.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    k.set_timing(False)
    calls, ms_search, ms_solve = k.timing_summary()
    st = d_st.cpu().numpy()
    ok = int((st == 0).sum())

    nb = k.max_neighbors
    m = w.config.method
    mdrift = 0 if m.kind == 0 else (1 if m.kind == 1 else (1 + D if m.degree == 1 else 1 + D + D * (D + 1) // 2))
    if args.workload == "uk3d_octant":
        # uneven systems: sum the per-target count over the size histogram
        _, cnt = k.neighbors(targets)
        hist = np.bincount(cnt[st == 0], minlength=65)
        flops_launch = float(sum(hist[c] * kriging_flops_per_target(c, mdrift, COV_FLOPS[args.workload])
                                 for c in range(len(hist)) if hist[c]))
    else:
        flops_launch = ok * kriging_flops_per_target(nb, mdrift, COV_FLOPS[args.workload])
    solve_s = ms_solve / max(calls, 1) / 1e3
    achieved = flops_launch / solve_s / 1e12 if solve_s > 0 else 0.0
    traffic = None
    tf = ROOT / "profiles" / f"traffic_{args.workload}_{args.precision}.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get("bytes_per_launch")
    out = {
        "metric": METRIC,
        "value": T * args.steps / dt,
        "unit": "targets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded, reference bench recipe)",
        "config": {"workload": args.workload, "n_data": int(w.coords.shape[0]),
                   "max_neighbors": int(nb), "targets": int(T), "precision": args.precision,
                   "ok_targets": ok},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": "krige_fp64_kernel", "kernel_ms": solve_s * 1e3,
                     "search_ms": ms_search / max(calls, 1),
                     "flops_per_launch": flops_launch},
        "cpu_baseline": None,
    }
    if args.host_io:
        k.predict_many(targets)                      # warm the host path
        t0 = time.perf_counter()
        for _ in range(args.steps):
            k.predict_many(targets)
        out["config"]["host_io_targets_per_s"] = T * args.steps / (time.perf_counter() - t0)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, targets)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
