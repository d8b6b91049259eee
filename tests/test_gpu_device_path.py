"""The device-resident predict path (gs_kriging_predict_many_device), the
one the benchmark times: targets and outputs are torch CUDA tensors and the
launches go onto a caller stream.

Bars: bit-equal to the host-buffer path (the same kernels on the same
targets; the order of the search and of the solves does not depend on where
the buffers live), and through it within 1e-10 of the oracle on a sample.
"""
import numpy as np
import pytest
import torch

import geostat_hip as gh
from geostat_hip import workloads

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def _device_predict(k, t, stream=None, with_optional=True):
    dev = torch.device("cuda:0")
    T = t.shape[0]
    d_t = torch.from_numpy(np.ascontiguousarray(t)).to(dev)
    v = torch.full((T,), -7.0, dtype=torch.float64, device=dev)
    var = torch.full((T,), -7.0, dtype=torch.float64, device=dev)
    lag = torch.full((T,), -7.0, dtype=torch.float64, device=dev) if with_optional else None
    st = torch.full((T,), -7, dtype=torch.int32, device=dev) if with_optional else None
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    with torch.cuda.stream(s):
        k.predict_many_device(d_t, v, var, lag, st, stream=s.cuda_stream)
    s.synchronize()
    out = [v.cpu().numpy(), var.cpu().numpy()]
    if with_optional:
        out += [lag.cpu().numpy(), st.cpu().numpy()]
    return out


def _rel(a, b):
    assert np.array_equal(np.isnan(a), np.isnan(b)), "NaN patterns differ"
    m = ~np.isnan(b)
    return float(np.abs(a[m] - b[m]).max() / max(np.abs(b[m]).max(), 1e-300))


def test_device_path_full_size_equals_host_path_and_oracle(oracle):
    w = workloads.ok2d_large()
    k = gh.Kriging(w.coords, w.values, w.model, w.config)
    t = w.targets()
    print(f"device path: ok2d_large T={len(t)}", flush=True)
    v, var, lag, st = _device_predict(k, t)
    hv, hvar, hlag, hst = k.predict_many(t)
    assert np.array_equal(st, hst) and (st == 0).all()
    assert np.array_equal(v, hv) and np.array_equal(var, hvar) and np.array_equal(lag, hlag)
    sel = t[::400]                  # 10,000 targets against the oracle
    rc, ov, ovar, _, ost, _ = oracle.predict_many(w.coords, w.values, w.model, w.config, sel)
    assert rc == 0 and np.array_equal(st[::400], ost)
    assert _rel(v[::400], ov) < TOL64 and _rel(var[::400], ovar) < TOL64


def test_device_path_side_stream_octant_and_optional_outputs():
    w = workloads.uk3d_octant(n=250_000, nx=100, ny=100, nz=20)
    k = gh.Kriging(w.coords, w.values, w.model, w.config)
    t = w.targets()
    print(f"device path: uk3d_octant n=250000 T={len(t)} side stream", flush=True)
    side = torch.cuda.Stream(device=torch.device("cuda:0"))
    v, var, lag, st = _device_predict(k, t, stream=side)
    hv, hvar, hlag, hst = k.predict_many(t)
    assert np.array_equal(st, hst)
    assert np.array_equal(v, hv) and np.array_equal(var, hvar)
    assert np.array_equal(np.isnan(lag), np.isnan(hlag))
    assert np.array_equal(lag[~np.isnan(lag)], hlag[~np.isnan(hlag)])
    # lagrange and status are optional
    v2, var2 = _device_predict(k, t[:5000], with_optional=False)
    assert np.array_equal(v2, hv[:5000]) and np.array_equal(var2, hvar[:5000])


def test_device_path_failed_targets_and_empty_call():
    c, vals = workloads.points(21, 4000, [1000.0, 1000.0])
    m = gh.VariogramModel(0.1, [gh.Structure(gh.ModelKind.SPHERICAL, 1.0, 300.0)])
    k = gh.Kriging(c, vals, m, gh.KrigingConfig(max_neighbors=16, min_neighbors=4,
                                                search_radius=40.0))
    t = np.array([[500.0, 500.0], [np.nan, 3.0], [5000.0, 5000.0], [c[7, 0], c[7, 1]]])
    print("device path: NaN / out-of-radius / on-data targets", flush=True)
    v, var, lag, st = _device_predict(k, t)
    hv, hvar, hlag, hst = k.predict_many(t)
    assert np.array_equal(st, hst)
    assert st[0] == 0 and st[1] != 0 and st[2] != 0 and st[3] == 0
    assert np.isnan(v[1]) and np.isnan(v[2]) and np.isnan(var[1]) and np.isnan(var[2])
    assert np.array_equal(v[[0, 3]], hv[[0, 3]]) and np.array_equal(var[[0, 3]], hvar[[0, 3]])
    assert abs(v[3] - vals[7]) < 1e-8 and var[3] < 1e-8
    # an empty call enqueues nothing and succeeds
    e = _device_predict(k, np.zeros((0, 2)))
    assert all(a.size == 0 for a in e)
