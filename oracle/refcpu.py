"""ctypes wrapper of the CPU oracle (oracle/build/librefcpu.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the reference the GPU results are
compared with or the timed CPU baseline, never as part of the product path.

Model / config arguments are read by attribute (kind, sill, range, anis,
nugget, structures, method, ...), so the same description objects the
product API takes can be handed to the oracle.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "build" / "librefcpu.so"
NONE = -1
MAX_STRUCT = 8


class rc_aniso(C.Structure):
    _fields_ = [("azimuth_deg", C.c_double), ("ratio", C.c_double), ("ratio_z", C.c_double),
                ("dip_deg", C.c_double), ("rake_deg", C.c_double)]


class rc_structure(C.Structure):
    _fields_ = [("kind", C.c_int32), ("param", C.c_double), ("sill", C.c_double),
                ("range", C.c_double), ("has_anis", C.c_int32), ("anis", rc_aniso)]


class rc_model(C.Structure):
    _fields_ = [("nugget", C.c_double), ("n_structures", C.c_int32),
                ("s", rc_structure * MAX_STRUCT)]


class rc_kconfig(C.Structure):
    _fields_ = [("method", C.c_int32), ("simple_mean", C.c_double), ("degree", C.c_int32),
                ("max_neighbors", C.c_int64), ("min_neighbors", C.c_int64),
                ("max_per_octant", C.c_int64), ("has_radius", C.c_int32), ("radius", C.c_double),
                ("has_anis_search", C.c_int32), ("anis_search", rc_aniso)]


class rc_rng(C.Structure):
    _fields_ = [("s", C.c_uint64 * 4), ("has_spare", C.c_int32), ("spare", C.c_double)]


_DP = C.POINTER(C.c_double)
_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_U32P = C.POINTER(C.c_uint32)
_lib = None


def build() -> Path:
    """Compiles the oracle (plain C) with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        h = C.CDLL(str(LIB))
        sig = {
            "rc_rng_new": (None, [C.POINTER(rc_rng), C.c_uint64]),
            "rc_rng_next_u64": (C.c_uint64, [C.POINTER(rc_rng)]),
            "rc_rng_uniform": (C.c_double, [C.POINTER(rc_rng)]),
            "rc_rng_normal": (C.c_double, [C.POINTER(rc_rng)]),
            "rc_bessel_k": (C.c_double, [C.c_double, C.c_double]),
            "rc_gamma_fn": (C.c_double, [C.c_double]),
            "rc_shape_g": (C.c_double, [C.c_int32, C.c_double, C.c_double, C.c_double]),
            "rc_gamma_dh": (C.c_double, [C.POINTER(rc_model), C.c_int32, _DP]),
            "rc_total_sill": (C.c_double, [C.POINTER(rc_model)]),
            "rc_model_validate": (C.c_int32, [C.POINTER(rc_model), C.c_int32]),
            "rc_lu_solve": (C.c_int32, [_DP, C.c_int32, _DP]),
            "rc_cholesky_solve": (C.c_int32, [_DP, C.c_int32, _DP]),
            "rc_kdtree_knn": (C.c_int64, [C.c_int32, _DP, C.c_int64, _DP, C.c_int64, C.c_int64,
                                          C.c_int32, C.c_double, _I32P, _I32P]),
            "rc_kdtree_layout": (C.c_int32, [C.c_int32, _DP, C.c_int64, _U32P]),
            "rc_kriging_predict_many": (C.c_int32, [C.c_int32, _DP, _DP, C.c_int64,
                                                    C.POINTER(rc_model), C.POINTER(rc_kconfig),
                                                    _DP, C.c_int64, C.c_int32, _DP, _DP, _DP,
                                                    _I32P, _I64P]),
            "rc_kriging_predict_many_ex": (C.c_int32, [C.c_int32, _DP, _DP, C.c_int64,
                                                       C.POINTER(rc_model),
                                                       C.POINTER(rc_kconfig), _DP, C.c_int32,
                                                       _DP, _DP, C.c_int64, _DP, _DP, C.c_int64,
                                                       C.c_int32, C.c_int32, _DP, _DP, _DP,
                                                       _I32P, _I64P]),
            "rc_kriging_neighbors": (C.c_int32, [C.c_int32, _DP, _DP, C.c_int64,
                                                 C.POINTER(rc_model), C.POINTER(rc_kconfig), _DP,
                                                 C.c_int64, C.c_int64, _I32P, _I32P, C.c_int32]),
            "rc_maxmin_order": (C.c_int32, [C.c_int32, _DP, C.c_int64, _I64P]),
            "rc_maxmin_order_brute": (C.c_int32, [C.c_int32, _DP, C.c_int64, _I64P]),
            "rc_vecchia_plan": (C.c_int32, [C.c_int32, _DP, C.c_int64, C.c_int64, _I64P, _I64P,
                                            _I32P, _I32P]),
            "rc_vecchia_loglik_plan": (C.c_int32, [C.c_int32, _DP, _DP, C.c_int64,
                                                   C.POINTER(rc_model), _I64P, _I32P, _I32P,
                                                   C.c_int64, _DP]),
            "rc_vecchia_loglik_grad_plan": (C.c_int32, [C.c_int32, _DP, _DP, C.c_int64,
                                                        C.c_int32, C.c_double, C.c_double,
                                                        C.c_double, _I64P, _I32P, _I32P,
                                                        C.c_int64, _DP, _DP]),
            "rc_vecchia_mle_plan": (C.c_int32, [C.c_int32, _DP, _DP, C.c_int64, C.c_int32,
                                                C.c_double, _I64P, _I32P, _I32P, C.c_int64,
                                                _DP, _DP, _I64P]),
            "rc_dense_mvn_loglik": (C.c_int32, [C.c_int32, _DP, _DP, C.c_int64,
                                                C.POINTER(rc_model), _DP]),
            "rc_duplicate_pair": (C.c_int32, [C.c_int32, _DP, C.c_int64, _I64P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def _d(a):
    return a.ctypes.data_as(_DP)


def _arr(a, dtype=np.float64):
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))


def aniso(a) -> rc_aniso:
    return rc_aniso(a.azimuth_deg, a.ratio, getattr(a, "ratio_z", 1.0), getattr(a, "dip_deg", 0.0),
                    getattr(a, "rake_deg", 0.0))


def model(m) -> rc_model:
    out = rc_model()
    out.nugget = m.nugget
    out.n_structures = len(m.structures)
    for i, s in enumerate(m.structures):
        kind, param = s.kind
        out.s[i].kind = kind
        out.s[i].param = param
        out.s[i].sill = s.sill
        out.s[i].range = s.range
        out.s[i].has_anis = 0 if s.anis is None else 1
        if s.anis is not None:
            out.s[i].anis = aniso(s.anis)
    return out


def kconfig(c) -> rc_kconfig:
    out = rc_kconfig()
    out.method = c.method.kind
    out.simple_mean = c.method.mean
    out.degree = c.method.degree
    out.max_neighbors = NONE if c.max_neighbors is None else c.max_neighbors
    out.min_neighbors = NONE if c.min_neighbors is None else c.min_neighbors
    out.max_per_octant = NONE if c.max_per_octant is None else c.max_per_octant
    out.has_radius = 0 if c.search_radius is None else 1
    out.radius = 0.0 if c.search_radius is None else c.search_radius
    out.has_anis_search = 0 if c.anisotropic_search is None else 1
    if c.anisotropic_search is not None:
        out.anis_search = aniso(c.anisotropic_search)
    return out


class Rng:
    """rng.rs Rng restated."""

    def __init__(self, seed: int):
        self._r = rc_rng()
        lib().rc_rng_new(C.byref(self._r), seed)

    def next_u64(self) -> int:
        return lib().rc_rng_next_u64(C.byref(self._r))

    def uniform(self) -> float:
        return lib().rc_rng_uniform(C.byref(self._r))

    def normal(self) -> float:
        return lib().rc_rng_normal(C.byref(self._r))


def shape_g(kind, h, a):
    k, p = kind
    return lib().rc_shape_g(k, p, h, a)


def gamma_dh(m, dh):
    dh = _arr(dh)
    return lib().rc_gamma_dh(C.byref(model(m)), dh.shape[0], _d(dh))


def bessel_k(nu, x):
    return lib().rc_bessel_k(nu, x)


def gamma_fn(x):
    return lib().rc_gamma_fn(x)


def lu_solve(a, b):
    a = _arr(a).copy()
    b = _arr(b).copy()
    st = lib().rc_lu_solve(_d(a), a.shape[0], _d(b))
    return st, b


def cholesky_solve(a, b):
    a = _arr(a).copy()
    b = _arr(b).copy()
    st = lib().rc_cholesky_solve(_d(a), a.shape[0], _d(b))
    return st, b


def kdtree_knn(coords, targets, k, radius=None):
    coords = _arr(coords)
    targets = _arr(targets).reshape(-1, coords.shape[1])
    T = targets.shape[0]
    idx = np.empty((T, k), dtype=np.int32)
    cnt = np.empty(T, dtype=np.int32)
    lib().rc_kdtree_knn(coords.shape[1], _d(coords), coords.shape[0], _d(targets), T, k,
                        0 if radius is None else 1, 0.0 if radius is None else radius,
                        idx.ctypes.data_as(_I32P), cnt.ctypes.data_as(_I32P))
    return idx, cnt


def kdtree_layout(coords):
    coords = _arr(coords)
    out = np.empty(coords.shape[0], dtype=np.uint32)
    lib().rc_kdtree_layout(coords.shape[1], _d(coords), coords.shape[0],
                           out.ctypes.data_as(_U32P))
    return out


def predict_many(coords, values, m, cfg, targets, threads=None):
    """Kriging::new + predict_many restated.  Returns (status, value,
    variance, lagrange, per-target status, dup_pair)."""
    coords = _arr(coords)
    values = _arr(values)
    D = coords.shape[1]
    targets = _arr(targets).reshape(-1, D)
    T = targets.shape[0]
    v, var, lag = np.empty(T), np.empty(T), np.empty(T)
    st = np.empty(T, dtype=np.int32)
    dup = np.zeros(2, dtype=np.int64)
    cm, cc = model(m), kconfig(cfg)
    threads = threads or (os.cpu_count() or 1)
    rc = lib().rc_kriging_predict_many(D, _d(coords), _d(values), coords.shape[0], C.byref(cm),
                                       C.byref(cc), _d(targets), T, threads, _d(v), _d(var),
                                       _d(lag), st.ctypes.data_as(_I32P),
                                       dup.ctypes.data_as(_I64P))
    return rc, v, var, lag, st, dup


def predict_many_ex(coords, values, m, cfg, targets, errors=None, drift=None, tdrift=None,
                    offsets=None, block=False, threads=None):
    """with_measurement_error / with_external_drift + predict_many,
    predict_many_with_drift (tdrift given) or predict_block per target
    (block=True with offsets).  Same returns as predict_many."""
    coords = _arr(coords)
    values = _arr(values)
    D = coords.shape[1]
    targets = _arr(targets).reshape(-1, D)
    T = targets.shape[0]
    v, var, lag = np.empty(T), np.empty(T), np.empty(T)
    st = np.empty(T, dtype=np.int32)
    dup = np.zeros(2, dtype=np.int64)
    cm, cc = model(m), kconfig(cfg)
    threads = threads or (os.cpu_count() or 1)
    e = None if errors is None else _arr(errors)
    dd = None if drift is None else _arr(drift).reshape(coords.shape[0], -1)
    nv = 0 if dd is None else dd.shape[1]
    td = None if tdrift is None else _arr(tdrift).reshape(T, -1)
    off = None if offsets is None else _arr(offsets).reshape(-1, D)
    n_off = 0 if off is None else off.shape[0]
    rc = lib().rc_kriging_predict_many_ex(
        D, _d(coords), _d(values), coords.shape[0], C.byref(cm), C.byref(cc),
        None if e is None else _d(e), nv, None if dd is None else _d(dd), _d(targets), T,
        None if td is None else _d(td), None if off is None else _d(off), n_off,
        1 if block else 0, threads, _d(v), _d(var), _d(lag), st.ctypes.data_as(_I32P),
        dup.ctypes.data_as(_I64P))
    return rc, v, var, lag, st, dup


def neighbors(coords, values, m, cfg, targets, max_out, threads=None):
    coords = _arr(coords)
    values = _arr(values)
    D = coords.shape[1]
    targets = _arr(targets).reshape(-1, D)
    T = targets.shape[0]
    idx = np.empty((T, max_out), dtype=np.int32)
    cnt = np.empty(T, dtype=np.int32)
    cm, cc = model(m), kconfig(cfg)
    threads = threads or (os.cpu_count() or 1)
    rc = lib().rc_kriging_neighbors(D, _d(coords), _d(values), coords.shape[0], C.byref(cm),
                                    C.byref(cc), _d(targets), T, max_out,
                                    idx.ctypes.data_as(_I32P), cnt.ctypes.data_as(_I32P), threads)
    return rc, idx, cnt


def maxmin_order(coords, brute=False):
    coords = _arr(coords)
    out = np.empty(coords.shape[0], dtype=np.int64)
    f = lib().rc_maxmin_order_brute if brute else lib().rc_maxmin_order
    f(coords.shape[1], _d(coords), coords.shape[0], out.ctypes.data_as(_I64P))
    return out


def vecchia_plan(coords, m, order=None):
    coords = _arr(coords)
    n = coords.shape[0]
    o_out = np.empty(n, dtype=np.int64)
    nb = np.empty((n, m), dtype=np.int32)
    cnt = np.empty(n, dtype=np.int32)
    o_in = None if order is None else _arr(order, np.int64)
    rc = lib().rc_vecchia_plan(coords.shape[1], _d(coords), n, m,
                               None if o_in is None else o_in.ctypes.data_as(_I64P),
                               o_out.ctypes.data_as(_I64P), nb.ctypes.data_as(_I32P),
                               cnt.ctypes.data_as(_I32P))
    return rc, o_out, nb, cnt


def vecchia_loglik_plan(coords, values, m_, order, nb, cnt):
    coords = _arr(coords)
    values = _arr(values)
    out = C.c_double()
    cm = model(m_)
    nb = _arr(nb, np.int32)
    rc = lib().rc_vecchia_loglik_plan(coords.shape[1], _d(coords), _d(values), coords.shape[0],
                                      C.byref(cm), _arr(order, np.int64).ctypes.data_as(_I64P),
                                      nb.ctypes.data_as(_I32P),
                                      _arr(cnt, np.int32).ctypes.data_as(_I32P), nb.shape[1],
                                      C.byref(out))
    return rc, out.value


def vecchia_loglik_grad_plan(coords, values, kind, nugget, psill, rng, order, nb, cnt):
    """loglik_grad_inner (vecchia.rs:428-566): (rc, loglik, [d/dnugget,
    d/dpsill, d/drange]) for one isotropic structure of `kind`."""
    coords = _arr(coords)
    values = _arr(values)
    nb = _arr(nb, np.int32)
    ll = C.c_double()
    grad = np.zeros(3)
    rc = lib().rc_vecchia_loglik_grad_plan(
        coords.shape[1], _d(coords), _d(values), coords.shape[0], int(kind), float(nugget),
        float(psill), float(rng), _arr(order, np.int64).ctypes.data_as(_I64P),
        nb.ctypes.data_as(_I32P), _arr(cnt, np.int32).ctypes.data_as(_I32P), nb.shape[1],
        C.byref(ll), _d(grad))
    return rc, ll.value, grad


def vecchia_mle_plan(coords, values, kind, param, order, nb, cnt):
    """mle_fit_with_plan (vecchia.rs:952-1081): (rc, (nugget, psill, range),
    loglik, evaluations)."""
    coords = _arr(coords)
    values = _arr(values)
    nb = _arr(nb, np.int32)
    fit = np.zeros(3)
    ll = C.c_double()
    ev = C.c_int64()
    rc = lib().rc_vecchia_mle_plan(
        coords.shape[1], _d(coords), _d(values), coords.shape[0], int(kind), float(param),
        _arr(order, np.int64).ctypes.data_as(_I64P), nb.ctypes.data_as(_I32P),
        _arr(cnt, np.int32).ctypes.data_as(_I32P), nb.shape[1], _d(fit), C.byref(ll),
        C.byref(ev))
    return rc, fit, ll.value, ev.value


def dense_mvn_loglik(coords, values, m_):
    coords = _arr(coords)
    values = _arr(values)
    out = C.c_double()
    cm = model(m_)
    rc = lib().rc_dense_mvn_loglik(coords.shape[1], _d(coords), _d(values), coords.shape[0],
                                   C.byref(cm), C.byref(out))
    return rc, out.value


def duplicate_pair(coords):
    coords = _arr(coords)
    pair = np.zeros(2, dtype=np.int64)
    found = lib().rc_duplicate_pair(coords.shape[1], _d(coords), coords.shape[0],
                                    pair.ctypes.data_as(_I64P))
    return (int(pair[0]), int(pair[1])) if found else None


# ---- validation.rs ---------------------------------------------------------
def shuffle(rng: "Rng", xs: list) -> None:
    """Rng::shuffle (rng.rs:76-82): Fisher-Yates from the end."""
    for i in range(len(xs) - 1, 0, -1):
        j = rng.next_u64() % (i + 1)
        xs[i], xs[j] = xs[j], xs[i]


def fold_assignment(n: int, k: int, seed: int):
    """fold_assignment (validation.rs:199-212)."""
    order = list(range(n))
    shuffle(Rng(seed), order)
    fold_of = [0] * n
    members = [[] for _ in range(k)]
    for j, i in enumerate(order):
        fold_of[i] = j % k
        members[j % k].append(i)
    return fold_of, members


def spatial_block_assignment(coords, blocks_per_dim):
    """spatial_block_assignment (validation.rs:294-323)."""
    coords = _arr(coords)
    D = coords.shape[1]
    mn = list(coords[0])
    mx = list(coords[0])
    for p in coords:
        for d in range(D):
            mn[d] = min(mn[d], p[d])
            mx[d] = max(mx[d], p[d])
    nb = int(np.prod(blocks_per_dim))
    fold_of = [0] * len(coords)
    members = [[] for _ in range(nb)]
    for i, p in enumerate(coords):
        idx, stride = 0, 1
        for d in range(D):
            span = max(mx[d] - mn[d], 2.2250738585072014e-308)
            frac = min(max((p[d] - mn[d]) / span, 0.0), 1.0 - 1e-12)
            bd = min(int(frac * blocks_per_dim[d]), blocks_per_dim[d] - 1)
            idx += bd * stride
            stride *= blocks_per_dim[d]
        fold_of[i] = idx
        members[idx].append(i)
    return fold_of, members


def cv_with_folds(coords, values, m, cfg, fold_of, members, drift=None):
    """cv_with_folds (validation.rs:214-256): (rc, predicted, variance); rc is
    the first failing fold's error (build or per-member status)."""
    coords = _arr(coords)
    values = _arr(values)
    n = len(values)
    pred, var = np.zeros(n), np.zeros(n)
    fold_of = np.asarray(fold_of)
    for f, mem in enumerate(members):
        if not mem:
            continue
        tr = np.nonzero(fold_of != f)[0]
        if drift is None:
            rc, v, vv, _, st, _ = predict_many(coords[tr], values[tr], m, cfg, coords[mem])
        else:
            dd = _arr(drift)
            rc, v, vv, _, st, _ = predict_many_ex(coords[tr], values[tr], m, cfg, coords[mem],
                                                  drift=dd[tr], tdrift=dd[mem])
        if rc != 0:
            return rc, pred, var
        for j, i in enumerate(mem):
            if st[j] != 0:
                return int(st[j]), pred, var
            pred[i] = v[j]
            var[i] = vv[j]
    return 0, pred, var
