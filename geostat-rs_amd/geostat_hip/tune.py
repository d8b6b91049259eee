"""Autotuner of the solve launch schedules (run on an MI355X).

For every compiled shape of the lean solve kernels (kriging: dim, padded
system size NB, drift terms M; Vecchia: dim, NB) it times the two group-size
variants on a synthetic workload of that shape, and for the uneven octant
workload it times the persistent grid of the size-bucketed launches at
several wavefronts per CU; the shapes of the benchmark workloads are then
timed again at their full sizes, which decide those rows.  The winners are
written as the C table
csrc/gs_tune_table.inc, which the library compiles in: the schedule is then a
fixed function of (device architecture, shape), identical in every process.

  python -m geostat_hip.tune [--out PATH] [--quick]
"""
from __future__ import annotations

import argparse
import os
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
DEFAULT_OUT = HERE.parent / "csrc" / "gs_tune_table.inc"


def _time_solve(k, d_t, outs, reps: int = 3) -> float:
    """Median solve-phase milliseconds of `reps` predict passes."""
    import torch
    k.predict_many_device(d_t, *outs)          # warm code and workspaces
    torch.cuda.synchronize()
    k.timing_summary()
    ms = []
    for _ in range(reps):
        k.set_timing(True)
        k.predict_many_device(d_t, *outs)
        torch.cuda.synchronize()
        k.set_timing(False)
        calls, _, solve = k.timing_summary()
        ms.append(solve / max(calls, 1))
    return float(np.median(ms))


def _with_env(name: str, value, fn):
    old = os.environ.get(name)
    os.environ[name] = str(value)
    try:
        return fn()
    finally:
        if old is None:
            del os.environ[name]
        else:
            os.environ[name] = old


def _outs(T):
    import torch
    dev = torch.device("cuda:0")
    return (torch.empty(T, dtype=torch.float64, device=dev),
            torch.empty(T, dtype=torch.float64, device=dev),
            torch.empty(T, dtype=torch.float64, device=dev),
            torch.empty(T, dtype=torch.int32, device=dev))


def tune_krige(quick: bool):
    import torch

    import geostat_hip as gh
    from geostat_hip import workloads
    rows = []
    n, T = (50_000, 100_000) if quick else (200_000, 400_000)
    for dim in (2, 3):
        ext = [10_000.0, 10_000.0] + ([1000.0] if dim == 3 else [])
        c, v = workloads.points(7 + dim, n, ext, recipe=1 if dim == 3 else 0)
        t = workloads.uniforms(17 + dim, T * dim).reshape(T, dim) * np.array(ext)
        d_t = torch.from_numpy(t).to("cuda:0")
        outs = _outs(T)
        an = gh.Anisotropy(30.0, 0.5, 0.5, 10.0, 0.0)
        model = gh.VariogramModel(0.05, [gh.Structure(gh.ModelKind.MATERN15, 1.0, 800.0, an)])
        methods = {0: gh.KrigingMethod.Simple(0.0), 1: gh.KrigingMethod.Ordinary(),
                   1 + dim: gh.KrigingMethod.Universal(1),
                   1 + dim + dim * (dim + 1) // 2: gh.KrigingMethod.Universal(2)}
        for nb in (16, 32, 64):
            for m, meth in methods.items():
                k = gh.Kriging(c, v, model, gh.KrigingConfig(method=meth, max_neighbors=nb))
                ms = [_with_env("GS_SOLVE_VARIANT", var, lambda: _time_solve(k, d_t, outs))
                      for var in (0, 1)]
                best = int(np.argmin(ms))
                print(f"krige D={dim} NB={nb} M={m}: variant ms {ms[0]:.3f} / {ms[1]:.3f} "
                      f"-> {best}", flush=True)
                rows.append(["krige", dim, nb, m, 1, 0, best, 16])
                del k
    return rows


def tune_octant_grid(rows, quick: bool):
    """Wavefronts per CU of the persistent bucket launches (uneven octant
    systems, universal kriging in 3-D)."""
    import torch

    import geostat_hip as gh
    from geostat_hip import workloads
    w = workloads.uk3d_octant(n=250_000, nx=100, ny=100, nz=20 if quick else 50)
    k = gh.Kriging(w.coords, w.values, w.model, w.config)
    t = w.targets()
    d_t = torch.from_numpy(t).to("cuda:0")
    outs = _outs(len(t))
    best, best_ms = 16, None
    for wpc in (4, 8, 16, 32):
        ms = _with_env("GS_WAVES_PER_CU", wpc, lambda: _time_solve(k, d_t, outs))
        print(f"octant uk3d: {wpc} waves/CU -> {ms:.3f} ms", flush=True)
        if best_ms is None or ms < best_ms:
            best, best_ms = wpc, ms
    for r in rows:
        if r[0] == "krige" and r[1] == 3 and r[3] == 4:
            r[7] = best


def tune_vecchia(quick: bool):
    import geostat_hip as gh
    from geostat_hip import workloads
    rows = []
    n = 50_000 if quick else 200_000
    models = workloads.vecchia_models(16)
    for dim in (2, 3):
        c, _ = workloads.points(3 + dim, n, [1.0] * dim, recipe=2)
        for nb, m in ((16, 16), (32, 30), (64, 60)):
            p = gh.vecchia_plan(c, m)
            o, nbl, cnt = p.arrays()
            vals = workloads.points(5 + dim, n, [1.0] * dim)[1]
            res = []
            for var in (0, 1):
                def run():
                    p.loglik_batch(vals, models)
                    t0 = time.perf_counter()
                    for _ in range(3):
                        p.loglik_batch(vals, models)
                    return (time.perf_counter() - t0) / 3 * 1e3
                res.append(_with_env("GS_SOLVE_VARIANT", var, run))
            best = int(np.argmin(res))
            print(f"vecchia D={dim} NB={nb}: variant ms {res[0]:.3f} / {res[1]:.3f} -> {best}",
                  flush=True)
            rows.append(["vecchia", dim, nb, 0, 1, 0, best, 16])
    return rows


def tune_bench_shapes(rows):
    """The shapes of the benchmark workloads, timed at their full sizes (the
    small generic workloads above can rank close variants the other way):
    2-D ordinary kriging at k = 32 (4*10^6 targets), the uneven 3-D
    universal-kriging octant search (4*10^6 targets, all three size buckets,
    group-size variant and persistent grid together), and the 64-model
    Vecchia batch at 10^6 points."""
    import torch

    import geostat_hip as gh
    from geostat_hip import workloads

    def set_row(family, D, NB, M, variant=None, wpc=None):
        for r in rows:
            if r[0] == family and r[1] == D and r[2] == NB and r[3] == M:
                if variant is not None:
                    r[6] = variant
                if wpc is not None:
                    r[7] = wpc

    for name, shapes in (("ok2d_large", [(2, 32, 1)]),
                         ("uk3d_octant", [(3, 16, 4), (3, 32, 4), (3, 64, 4)])):
        w = getattr(workloads, name)()
        k = gh.Kriging(w.coords, w.values, w.model, w.config)
        t = w.targets()
        d_t = torch.from_numpy(t).to("cuda:0")
        outs = _outs(len(t))
        grid = (16,) if name == "ok2d_large" else (8, 16, 32, 128)
        best, best_ms = None, None
        for var in (0, 1):
            for wpc in grid:
                ms = _with_env("GS_SOLVE_VARIANT", var, lambda: _with_env(
                    "GS_WAVES_PER_CU", wpc, lambda: _time_solve(k, d_t, outs)))
                print(f"{name}: variant {var}, {wpc} waves/CU -> {ms:.3f} ms", flush=True)
                if best_ms is None or ms < best_ms:
                    best, best_ms = (var, wpc), ms
        for D, NB, M in shapes:
            set_row("krige", D, NB, M, *best)
        del k
    c = workloads.vecchia_c3()
    p = gh.vecchia_plan(c, 30)
    o, nb, cnt = p.arrays()
    vals = workloads.vecchia_field(c, o, nb, cnt, 30)
    models = workloads.vecchia_models(64)
    res = []
    for var in (0, 1):
        def run():
            p.loglik_batch(vals, models)
            t0 = time.perf_counter()
            for _ in range(3):
                p.loglik_batch(vals, models)
            return (time.perf_counter() - t0) / 3 * 1e3
        res.append(_with_env("GS_SOLVE_VARIANT", var, run))
        print(f"vecchia_c3: variant {var} -> {res[-1]:.3f} ms", flush=True)
    set_row("vecchia", 2, 32, 0, int(np.argmin(res)))


def write_table(rows, out: Path):
    lines = ["// Launch schedules measured on an MI355X (gfx950) by geostat_hip/tune.py:",
             "// {family, D, NB, M, lean, ext, group-size variant, waves per CU}."]
    for r in rows:
        lines.append('{"%s", %d, %d, %d, %d, %d, %d, %d},' % tuple(r))
    out.write_text("\n".join(lines) + "\n")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(DEFAULT_OUT))
    ap.add_argument("--quick", action="store_true", help="smaller workloads")
    args = ap.parse_args()
    rows = tune_krige(args.quick)
    tune_octant_grid(rows, args.quick)
    rows += tune_vecchia(args.quick)
    tune_bench_shapes(rows)
    write_table(rows, Path(args.out))
    print(f"wrote {args.out} ({len(rows)} entries)", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
