"""A predict call captured as a HIP graph (gs_kriging_graph_create /
gs_graph_launch): the graph is replayed several times, with the predictor's
own eager calls and other device work between the replays and new targets
written into the captured buffer before each, and every replay must be
bit-identical to the eager device path on the same targets.

Cases: the tile search in Hilbert order (k-nearest, 2-D), uneven octant
neighbourhoods (size buckets, 3-D universal kriging), FP32-refined solves
with their FP64 fallback list, a search-free predictor, a global
neighbourhood, an external-drift predictor (every target fails), and the
refusal of neighbourhoods above 64 points."""
import numpy as np
import pytest
import torch

import geostat_hip as gh
from geostat_hip import workloads

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _outs(T):
    return (torch.full((T,), -7.0, dtype=torch.float64, device=DEV),
            torch.full((T,), -7.0, dtype=torch.float64, device=DEV),
            torch.full((T,), -7.0, dtype=torch.float64, device=DEV),
            torch.full((T,), -7, dtype=torch.int32, device=DEV))


def _eager(k, d_t, precision):
    o = _outs(d_t.shape[0])
    k.predict_many_device(d_t, *o, precision=precision)
    torch.cuda.synchronize()
    return [x.cpu().numpy() for x in o]


def _same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x, y, equal_nan=True)


def _replays(k, target_sets, precision, reps=4):
    """Capture on target_sets[0], then replay on each later set (written in
    place into the captured buffer), checking every replay."""
    T = target_sets[0].shape[0]
    d_t = torch.from_numpy(np.ascontiguousarray(target_sets[0])).to(DEV)
    outs = _outs(T)
    want = _eager(k, d_t.clone(), precision)
    g = k.capture_graph(d_t, *outs, precision=precision)
    torch.cuda.synchronize()
    _same([x.cpu().numpy() for x in outs], want)       # the capture's own eager pass
    s = torch.cuda.Stream(DEV)
    for r in range(reps):
        t = target_sets[1 + r % (len(target_sets) - 1)]
        d_t.copy_(torch.from_numpy(np.ascontiguousarray(t)))
        want = _eager(k, d_t.clone(), precision)
        # other work of the predictor and of the program between replays
        _eager(k, d_t[: max(1, T // 3)].clone(), precision)
        (outs[0] * 0.0).sum().item()
        for x in outs:
            x.fill_(-3)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            g.launch(s.cuda_stream)
        s.synchronize()
        print(f"  replay {r}: T={T}", flush=True)
        _same([x.cpu().numpy() for x in outs], want)
    g.close()


def _target_sets(seed, T, ext, n_sets=3, recipe=0):
    return [workloads.points(seed + i, T, ext, recipe=recipe)[0] for i in range(n_sets)]


@pytest.mark.parametrize("precision", ["fp64", "fp32r"])
def test_graph_knn_2d(precision):
    c, v = workloads.points(201, 50_000, [1000.0, 1000.0])
    m = gh.VariogramModel(0.05, [gh.Structure(gh.ModelKind.MATERN15, 1.0, 80.0,
                                              gh.Anisotropy(30.0, 0.5, 0.5, 10.0, 0.0))])
    k = gh.Kriging(c, v, m, gh.KrigingConfig(max_neighbors=32))
    print(f"graph knn 2-D {precision}", flush=True)
    _replays(k, _target_sets(202, 100_000, [1000.0, 1000.0]), precision)


def test_graph_octant_3d_buckets():
    c, v = workloads.points(211, 40_000, [1000.0, 1000.0, 200.0], recipe=1)
    m = gh.VariogramModel(0.1, [gh.Structure(gh.ModelKind.SPHERICAL, 1.0, 150.0)])
    cfg = gh.KrigingConfig(method=gh.KrigingMethod.Universal(1), max_per_octant=8,
                           search_radius=90.0, min_neighbors=4)
    k = gh.Kriging(c, v, m, cfg)
    print("graph octant 3-D", flush=True)
    _replays(k, _target_sets(212, 70_000, [1000.0, 1000.0, 200.0], recipe=1), "fp64")


def test_graph_small_searchless_global_and_drift():
    c, v = workloads.points(221, 40, [100.0, 100.0])
    m = gh.VariogramModel(0.0, [gh.Structure(gh.ModelKind.EXPONENTIAL, 1.0, 30.0)])
    k = gh.Kriging(c, v, m, gh.KrigingConfig(max_neighbors=60))
    print("graph search-free", flush=True)
    _replays(k, _target_sets(222, 1000, [100.0, 100.0]), "fp32r", reps=2)
    c, v = workloads.points(223, 300, [100.0, 100.0])
    k = gh.Kriging(c, v, m, gh.KrigingConfig(method=gh.KrigingMethod.Ordinary()))
    print("graph global", flush=True)
    _replays(k, _target_sets(224, 2000, [100.0, 100.0]), "fp64", reps=2)
    drift = np.stack([c[:, 0] / 100.0], axis=1)
    k = gh.Kriging.with_external_drift(
        c, v, m, gh.KrigingConfig(method=gh.KrigingMethod.ExternalDrift(1), max_neighbors=16),
        drift)
    print("graph external drift", flush=True)
    _replays(k, _target_sets(225, 500, [100.0, 100.0]), "fp64", reps=2)


def test_graph_rejections():
    c, v = workloads.points(231, 5000, [1000.0, 1000.0])
    m = gh.VariogramModel(0.0, [gh.Structure(gh.ModelKind.EXPONENTIAL, 1.0, 100.0)])
    k = gh.Kriging(c, v, m, gh.KrigingConfig(max_neighbors=100))
    d_t = torch.zeros((10, 2), dtype=torch.float64, device=DEV)
    outs = _outs(10)
    with pytest.raises(gh.InvalidParameter, match="above 64"):
        k.capture_graph(d_t, *outs)
    k = gh.Kriging(c, v, m, gh.KrigingConfig(max_neighbors=16))
    with pytest.raises(gh.InvalidParameter):
        k.capture_graph(d_t, *outs, precision="fp16")
    with pytest.raises(gh.InvalidParameter, match="T >= 1"):
        k.capture_graph(d_t[:0], *[x[:0] for x in outs])
