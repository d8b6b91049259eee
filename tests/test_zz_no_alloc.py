"""gs_kriging_predict_many_noalloc (the entry point that allocates nothing)
against gs_kriging_predict_many_device, through the C ABI: bit-equal
outputs on two calls in a row, the second with fewer targets (the scratch
keeps the first call's contents), for k-nearest and octant neighbourhoods
and both precisions; oversized calls and global neighbourhoods are refused
before any work is enqueued."""
import ctypes as C

import numpy as np
import pytest
import torch

import geostat_hip as gh
from geostat_hip import _abi, workloads

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _bufs(T):
    return (torch.full((T,), -7.0, dtype=torch.float64, device=DEV),
            torch.full((T,), -7.0, dtype=torch.float64, device=DEV),
            torch.full((T,), -7.0, dtype=torch.float64, device=DEV),
            torch.full((T,), -7, dtype=torch.int32, device=DEV))


def _call(fn, k, d_t, prec, outs, stream):
    T = d_t.shape[0]
    return fn(k._h, C.c_void_p(d_t.data_ptr()), T, gh._PREC[prec],
              *[C.c_void_p(o.data_ptr()) for o in outs], C.c_void_p(stream.cuda_stream))


def _compare(k, t, prec):
    lib = _abi.lib()
    s = torch.cuda.Stream(device=DEV)
    d_t = torch.from_numpy(np.ascontiguousarray(t)).to(DEV)
    want, got = _bufs(len(t)), _bufs(len(t))
    torch.cuda.synchronize()
    assert _call(lib.gs_kriging_predict_many_device, k, d_t, prec, want, s) == 0, _abi.last_error()
    assert _call(lib.gs_kriging_predict_many_noalloc, k, d_t, prec, got, s) == 0, _abi.last_error()
    s.synchronize()
    for w, g in zip(want, got):
        assert torch.equal(torch.isnan(w), torch.isnan(g))
        m = ~torch.isnan(w)
        assert torch.equal(w[m], g[m])
    assert (got[3] >= 0).all()


@pytest.mark.parametrize("prec", ["fp64", "fp32r"])
def test_two_calls_second_smaller_knn(prec):
    w = workloads.ok2d_large()
    k = gh.Kriging(w.coords, w.values, w.model, w.config)
    t = w.targets()
    print(f"no-alloc {prec}: ok2d_large 200000 then 37000 targets", flush=True)
    _compare(k, t[::20], prec)
    _compare(k, t[5::108], prec)


def test_two_calls_second_smaller_octant():
    w = workloads.uk3d_octant(n=250_000, nx=100, ny=100, nz=10)
    k = gh.Kriging(w.coords, w.values, w.model, w.config)
    t = w.targets()
    print(f"no-alloc: uk3d_octant {len(t)} then {len(t[1::7])} targets", flush=True)
    _compare(k, t, "fp64")
    _compare(k, t[1::7], "fp64")


def test_refusals_enqueue_nothing():
    lib = _abi.lib()
    s = torch.cuda.Stream(device=DEV)
    w = workloads.ok2d_large()
    k = gh.Kriging(w.coords, w.values, w.model, w.config)
    T = 2_000_000                      # beyond the 64 MiB scratch
    d_t = torch.zeros((T, 2), dtype=torch.float64, device=DEV)
    outs = _bufs(T)
    torch.cuda.synchronize()
    assert _call(lib.gs_kriging_predict_many_noalloc, k, d_t, "fp64", outs, s) == gh.InvalidParameter.code
    assert "split the targets" in _abi.last_error()
    s.synchronize()
    assert (outs[0] == -7.0).all() and (outs[3] == -7).all()
    c, v = workloads.points(3, 500, [100.0, 100.0])
    kg = gh.Kriging(c, v, w.model, gh.KrigingConfig())     # global neighbourhood
    small = torch.zeros((10, 2), dtype=torch.float64, device=DEV)
    assert _call(lib.gs_kriging_predict_many_noalloc, kg, small, "fp64", _bufs(10), s) == \
        gh.InvalidParameter.code
