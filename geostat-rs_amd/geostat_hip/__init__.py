"""geostat_hip — Python mirror of geostat-core's predictor / likelihood API
over the MI355X C ABI (include/geostat_hip.h).

Names, argument meaning and error behaviour follow the reference
(franciscoparrao/geostat-rs, crates/geostat-core/src/):

* ``Kriging(coords, values, model, config)``  — Kriging::new (kriging.rs:166)
* ``Kriging.predict / predict_many / predict_grid`` — kriging.rs:494, 648, 783
* ``Kriging.neighbors``                        — kriging.rs:441-490
* ``vecchia_plan`` / ``VecchiaPlan.loglik``    — vecchia.rs:232, 285
* ``vecchia_loglik``                           — vecchia.rs:874
* ``VariogramModel`` / ``Structure`` / ``Anisotropy`` / ``ModelKind`` — variogram/model.rs
* ``GeostatError`` and subclasses              — error.rs:15-40

Build-time problems raise; per-target problems give NaN estimates
(kriging.rs:130-134, 648-651).  Every call runs on the GPU through
libgeostat_hip.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import math
import re
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _abi
from ._abi import GS_NONE, PREC_FP32_REFINED, PREC_FP64

__all__ = [
    "GeostatError", "InsufficientData", "DimensionMismatch", "SingularSystem",
    "InvalidParameter", "NoNeighbors", "DuplicatePoints", "DeviceError",
    "ModelKind", "Anisotropy", "Structure", "VariogramModel", "KrigingMethod",
    "KrigingConfig", "KrigingEstimate", "Context", "Kriging", "Grid2D", "Grid3D",
    "block_offsets", "Server", "Ticket", "VecchiaPlan", "VecchiaFit", "vecchia_plan", "vecchia_loglik", "vecchia_mle", "PREC_FP64",
    "PREC_FP32_REFINED",
]


# ---------------------------------------------------------------- errors --
class GeostatError(ValueError):
    """GeostatError (error.rs:15-40).  A ValueError, as the reference's
    Python binding raises every GeostatError as one."""
    code = -1


class InsufficientData(GeostatError):
    code = 1


class DimensionMismatch(GeostatError):
    code = 2


class SingularSystem(GeostatError):
    code = 3


class InvalidParameter(GeostatError):
    code = 4


class NoNeighbors(GeostatError):
    code = 5


class DuplicatePoints(GeostatError):
    code = 6

    def __init__(self, i: int, j: int, msg: str = ""):
        super().__init__(msg or f"duplicate points {i}, {j}")
        self.i, self.j = i, j


class DeviceError(GeostatError):
    code = 100


_ERRORS = {c.code: c for c in (InsufficientData, DimensionMismatch, SingularSystem,
                               InvalidParameter, NoNeighbors, DeviceError)}


def _check(status: int, dup=None) -> None:
    if status == 0:
        return
    msg = _abi.last_error()
    if status == DuplicatePoints.code:
        if dup is None:   # the message carries the pair: "duplicate points i, j"
            found = re.search(r"duplicate points (\d+), (\d+)", msg)
            dup = (int(found.group(1)), int(found.group(2))) if found else (-1, -1)
        raise DuplicatePoints(int(dup[0]), int(dup[1]), msg)
    raise _ERRORS.get(status, DeviceError)(msg)


# ------------------------------------------------------------------ model --
class ModelKind:
    """ModelKind (model.rs:47-95); parametrised kinds are (kind, param) pairs."""
    SPHERICAL = (0, 0.0)
    EXPONENTIAL = (1, 0.0)
    GAUSSIAN = (2, 0.0)
    MATERN15 = (3, 0.0)
    MATERN25 = (4, 0.0)
    CIRCULAR = (6, 0.0)
    HOLE = (8, 0.0)
    WAVE = (9, 0.0)

    @staticmethod
    def Matern(nu: float):
        return (5, float(nu))

    @staticmethod
    def Stable(alpha: float):
        return (7, float(alpha))

    @staticmethod
    def Power(theta: float):
        return (10, float(theta))


@dataclass
class Anisotropy:
    """Anisotropy (model.rs:321-341)."""
    azimuth_deg: float
    ratio: float
    ratio_z: float = 1.0
    dip_deg: float = 0.0
    rake_deg: float = 0.0

    def _c(self) -> _abi.gs_anisotropy:
        return _abi.gs_anisotropy(self.azimuth_deg, self.ratio, self.ratio_z, self.dip_deg,
                                  self.rake_deg)


@dataclass
class Structure:
    """Structure (model.rs:438-449)."""
    kind: tuple
    sill: float
    range: float
    anis: Optional[Anisotropy] = None

    @classmethod
    def new(cls, kind, sill, rng):
        return cls(kind, sill, rng)

    @classmethod
    def with_anisotropy(cls, kind, sill, rng, azimuth_deg, ratio):
        return cls(kind, sill, rng, Anisotropy(azimuth_deg, ratio))

    @classmethod
    def with_rotation(cls, kind, sill, rng, azimuth_deg, dip_deg, rake_deg, ratio, ratio_z):
        return cls(kind, sill, rng, Anisotropy(azimuth_deg, ratio, ratio_z, dip_deg, rake_deg))


_KIND_NAME = {0: "spherical", 1: "exponential", 2: "gaussian", 3: "matern15", 4: "matern25",
              5: "matern", 6: "circular", 7: "stable", 8: "hole", 9: "wave", 10: "power"}
_KIND_CODE = {v: k for k, v in _KIND_NAME.items()}


def _json_num(x: float) -> str:
    """Shortest round-trip decimal, in serde_json's spelling (1.0, 1e-7)."""
    x = float(x)
    if not math.isfinite(x):
        raise InvalidParameter("non-finite model parameter cannot be written as JSON")
    r = repr(x)
    if "e" in r:
        mant, ex = r.split("e")
        return f"{mant}e{int(ex)}"
    return r


@dataclass
class VariogramModel:
    """VariogramModel (model.rs:543-549): nugget + nested structures."""
    nugget: float
    structures: list = field(default_factory=list)

    def total_sill(self) -> float:
        s = -0.0
        for st in self.structures:
            s += st.sill
        return self.nugget + s

    # ---- JSON (serde form of VariogramModel, model.rs:47-48, 320-341,
    # 438-449, 543-549): kinds in lowercase, parametrised kinds as
    # {"matern": nu}; anisotropy fields at their defaults are omitted
    @classmethod
    def from_json(cls, text: str) -> "VariogramModel":
        """VariogramModel.from_json (python/src/lib.rs:131-138)."""
        import json
        try:
            d = json.loads(text)
            structs = []
            for st in d["structures"]:
                k = st["kind"]
                if isinstance(k, str):
                    kind = (_KIND_CODE[k], 0.0)
                else:
                    (name, param), = k.items()
                    if name not in ("matern", "stable", "power"):
                        raise KeyError(name)
                    kind = (_KIND_CODE[name], float(param))
                a = st.get("anis")
                an = None if a is None else Anisotropy(
                    float(a["azimuth_deg"]), float(a["ratio"]), float(a.get("ratio_z", 1.0)),
                    float(a.get("dip_deg", 0.0)), float(a.get("rake_deg", 0.0)))
                structs.append(Structure(kind, float(st["sill"]), float(st["range"]), an))
            m = cls(float(d["nugget"]), structs)
        except (KeyError, TypeError, ValueError, AttributeError) as e:
            raise InvalidParameter(f"cannot parse model JSON: {e}") from None
        m.validate()
        return m

    def to_json(self) -> str:
        """VariogramModel.to_json (python/src/lib.rs:140-143)."""
        parts = []
        for st in self.structures:
            code, param = st.kind
            name = _KIND_NAME[code]
            kind = (f'{{"{name}":{_json_num(param)}}}' if code in (5, 7, 10) else f'"{name}"')
            body = f'"kind":{kind},"sill":{_json_num(st.sill)},"range":{_json_num(st.range)}'
            if st.anis is not None:
                a = st.anis
                an = f'"azimuth_deg":{_json_num(a.azimuth_deg)},"ratio":{_json_num(a.ratio)}'
                if a.ratio_z != 1.0:
                    an += f',"ratio_z":{_json_num(a.ratio_z)}'
                if a.dip_deg != 0.0:
                    an += f',"dip_deg":{_json_num(a.dip_deg)}'
                if a.rake_deg != 0.0:
                    an += f',"rake_deg":{_json_num(a.rake_deg)}'
                body += f',"anis":{{{an}}}'
            parts.append("{" + body + "}")
        return f'{{"nugget":{_json_num(self.nugget)},"structures":[{",".join(parts)}]}}'

    def validate(self) -> None:
        """VariogramModel::new's checks (model.rs:554-618), as the C library
        runs them."""
        if len(self.structures) > _abi.GS_MAX_STRUCTURES:
            raise InvalidParameter("unsupported on GPU: more than 8 nested structures")
        cm = self._c()
        _check(_abi.lib().gs_model_validate(C.byref(cm)))

    def _c(self) -> _abi.gs_model:
        if len(self.structures) > _abi.GS_MAX_STRUCTURES:
            raise InvalidParameter("unsupported on GPU: more than 8 nested structures")
        m = _abi.gs_model()
        m.nugget = self.nugget
        m.n_structures = len(self.structures)
        for i, s in enumerate(self.structures):
            cs = m.structures[i]
            cs.kind, cs.param = s.kind
            cs.sill = s.sill
            cs.range = s.range
            cs.has_anis = 1 if s.anis is not None else 0
            if s.anis is not None:
                cs.anis = s.anis._c()
        return m


# ----------------------------------------------------------------- config --
@dataclass
class KrigingMethod:
    """KrigingMethod (kriging.rs:25-44): kind 0 simple, 1 ordinary, 2
    universal, 3 external drift (degree holds the covariate count)."""
    kind: int = 1
    mean: float = 0.0
    degree: int = 1

    @staticmethod
    def Simple(mean: float):
        return KrigingMethod(0, mean=mean)

    @staticmethod
    def Ordinary():
        return KrigingMethod(1)

    @staticmethod
    def Universal(degree: int):
        return KrigingMethod(2, degree=degree)

    @staticmethod
    def ExternalDrift(n_vars: int):
        return KrigingMethod(3, degree=n_vars)


@dataclass
class KrigingConfig:
    """KrigingConfig (kriging.rs:73-115); None = not set."""
    method: KrigingMethod = field(default_factory=KrigingMethod.Ordinary)
    max_neighbors: Optional[int] = None
    search_radius: Optional[float] = None
    min_neighbors: Optional[int] = None
    max_per_octant: Optional[int] = None
    anisotropic_search: Optional[Anisotropy] = None

    def _c(self) -> _abi.gs_kconfig:
        c = _abi.gs_kconfig()
        c.method = self.method.kind
        c.simple_mean = self.method.mean
        c.degree = self.method.degree
        c.max_neighbors = GS_NONE if self.max_neighbors is None else int(self.max_neighbors)
        c.min_neighbors = GS_NONE if self.min_neighbors is None else int(self.min_neighbors)
        c.max_per_octant = GS_NONE if self.max_per_octant is None else int(self.max_per_octant)
        c.has_search_radius = 0 if self.search_radius is None else 1
        c.search_radius = 0.0 if self.search_radius is None else float(self.search_radius)
        c.has_anisotropic_search = 0 if self.anisotropic_search is None else 1
        if self.anisotropic_search is not None:
            c.anisotropic_search = self.anisotropic_search._c()
        return c


@dataclass
class KrigingEstimate:
    """KrigingEstimate (kriging.rs:118-128)."""
    value: float
    variance: float
    lagrange: Optional[float]


# ---------------------------------------------------------------- context --
class Context:
    """A device plus its default stream (gs_ctx)."""
    _default = None

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(_abi.lib().gs_ctx_create(device, C.byref(h)))
        self._h = h
        self._release = _abi.lib().gs_ctx_destroy   # kept: module globals may be gone at exit
        self.device = device

    @classmethod
    def default(cls) -> "Context":
        if cls._default is None:
            cls._default = Context(0)
        return cls._default

    def synchronize(self) -> None:
        _check(_abi.lib().gs_ctx_synchronize(self._h))

    def __del__(self):
        h = getattr(self, "_h", None)
        release = getattr(self, "_release", None)
        if h and release is not None:
            release(h)
            self._h = None


def _f64(a, shape_last=None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    return a


_PREC = {"fp64": PREC_FP64, "fp32r": PREC_FP32_REFINED, PREC_FP64: PREC_FP64,
         PREC_FP32_REFINED: PREC_FP32_REFINED}


# ---------------------------------------------------------------- kriging --
class Kriging:
    """Kriging::<D, VariogramModel>::new (kriging.rs:166) bound to a device.

    coords: (n, D) array; values: (n,).  Data and the search index stay
    resident in device memory until the object is freed.
    """

    def __init__(self, coords, values, model: VariogramModel,
                 config: Optional[KrigingConfig] = None, ctx: Optional[Context] = None,
                 _errors=None, _drift=None):
        coords = _f64(coords)
        values = _f64(values).reshape(-1)
        if coords.ndim != 2 or coords.shape[1] not in (2, 3):
            raise DimensionMismatch(f"coords must be (n, 2|3), got {coords.shape}")
        if coords.shape[0] != values.shape[0]:
            raise DimensionMismatch(f"{coords.shape[0]} coordinates vs {values.shape[0]} values")
        self.ctx = ctx or Context.default()
        self.dim = coords.shape[1]
        self.n = coords.shape[0]
        self.model = model
        self.config = config or KrigingConfig()
        self.n_vars = 0
        cm = model._c()
        cc = self.config._c()
        h = C.c_void_p()
        dup = np.zeros(2, dtype=np.int64)
        lib = _abi.lib()
        if _errors is not None:
            e = _f64(_errors).reshape(-1)
            power = any(st_.kind[0] == 10 for st_ in model.structures)
            if e.shape[0] != self.n and not (self.config.method.kind == 3 or power):
                # (external drift and Power are rejected first, before the
                # variances are read: kriging.rs:187-204)
                raise DimensionMismatch(f"{e.shape[0]} error variances vs {self.n} data points")
            st = lib.gs_kriging_new_with_measurement_error(
                self.ctx._h, self.dim, _abi.dptr(coords), _abi.dptr(values), self.n,
                C.byref(cm), C.byref(cc), _abi.dptr(e), C.byref(h), _abi.i64ptr(dup))
        elif _drift is not None:
            d = _f64(_drift)
            if d.ndim != 2 or d.shape[0] != self.n:
                raise DimensionMismatch(f"{d.shape[0] if d.ndim else 0} drift rows vs {self.n} "
                                        "data points")
            if self.config.method.kind == 3 and d.shape[1] != self.config.method.degree:
                raise DimensionMismatch(f"drift rows have {d.shape[1]} values, expected "
                                        f"{self.config.method.degree}")
            self.n_vars = d.shape[1]
            st = lib.gs_kriging_new_with_external_drift(
                self.ctx._h, self.dim, _abi.dptr(coords), _abi.dptr(values), self.n,
                C.byref(cm), C.byref(cc), self.n_vars, _abi.dptr(d), C.byref(h),
                _abi.i64ptr(dup))
        else:
            st = lib.gs_kriging_new(self.ctx._h, self.dim, _abi.dptr(coords), _abi.dptr(values),
                                    self.n, C.byref(cm), C.byref(cc), C.byref(h),
                                    _abi.i64ptr(dup))
        _check(st, dup)
        self._h = h
        self._release = _abi.lib().gs_kriging_free   # kept: module globals may be gone at exit

    @classmethod
    def with_measurement_error(cls, coords, values, model: VariogramModel,
                               config: KrigingConfig, errors, ctx: Optional[Context] = None):
        """with_measurement_error (kriging.rs:181-218): errors[i] is the
        error variance of datum i (data-data diagonal only)."""
        return cls(coords, values, model, config, ctx, _errors=errors)

    @classmethod
    def with_external_drift(cls, coords, values, model: VariogramModel,
                            config: KrigingConfig, drift_data, ctx: Optional[Context] = None):
        """with_external_drift (kriging.rs:223-260): drift_data (n, n_vars)."""
        return cls(coords, values, model, config, ctx, _drift=drift_data)

    def __del__(self):
        h = getattr(self, "_h", None)
        release = getattr(self, "_release", None)
        if h and release is not None:
            release(h)
            self._h = None

    @property
    def max_neighbors(self) -> int:
        return int(_abi.lib().gs_kriging_max_neighbors(self._h))

    def predict_many(self, targets, precision="fp64"):
        """predict_many (kriging.rs:648-652): arrays (value, variance,
        lagrange, status); failed targets are NaN with a non-zero status."""
        t = _f64(targets).reshape(-1, self.dim)
        T = t.shape[0]
        v = np.empty(T)
        var = np.empty(T)
        lag = np.empty(T)
        st = np.empty(T, dtype=np.int32)
        _check(_abi.lib().gs_kriging_predict_many(self._h, _abi.dptr(t), T, _PREC[precision],
                                                  _abi.dptr(v), _abi.dptr(var), _abi.dptr(lag),
                                                  _abi.i32ptr(st)))
        return v, var, lag, st

    def predict_many_with_drift(self, targets, drifts, precision="fp64"):
        """predict_many_with_drift (kriging.rs:656-672): drifts (T, n_vars)."""
        t = _f64(targets).reshape(-1, self.dim)
        T = t.shape[0]
        d = _f64(drifts).reshape(T, -1) if T else _f64(drifts)
        if d.shape[0] != T:
            raise DimensionMismatch(f"{T} targets vs {d.shape[0]} drift rows")
        if self.n_vars and T and d.shape[1] != self.n_vars:
            raise DimensionMismatch(f"target drift has {d.shape[1]} values, expected {self.n_vars}")
        v, var, lag = np.empty(T), np.empty(T), np.empty(T)
        st = np.empty(T, dtype=np.int32)
        _check(_abi.lib().gs_kriging_predict_many_with_drift(
            self._h, _abi.dptr(t), T, _abi.dptr(d), _PREC[precision], _abi.dptr(v),
            _abi.dptr(var), _abi.dptr(lag), _abi.i32ptr(st)))
        return v, var, lag, st

    def predict_block_many(self, centers, offsets, precision="fp64"):
        """predict_block (kriging.rs:678-777) at every centre: arrays (value,
        variance, lagrange, status)."""
        t = _f64(centers).reshape(-1, self.dim)
        T = t.shape[0]
        off = _f64(offsets).reshape(-1, self.dim)
        v, var, lag = np.empty(T), np.empty(T), np.empty(T)
        st = np.empty(T, dtype=np.int32)
        _check(_abi.lib().gs_kriging_predict_block_many(
            self._h, _abi.dptr(t), T, _abi.dptr(off) if off.shape[0] else None, off.shape[0],
            _PREC[precision], _abi.dptr(v), _abi.dptr(var), _abi.dptr(lag), _abi.i32ptr(st)))
        return v, var, lag, st

    def predict_block(self, center, offsets, precision="fp64") -> KrigingEstimate:
        """predict_block (kriging.rs:678-777): raises the per-target error."""
        v, var, lag, st = self.predict_block_many(np.asarray(center, dtype=np.float64)
                                                  .reshape(1, -1), offsets, precision)
        if st[0] != 0:
            raise _ERRORS.get(int(st[0]), GeostatError)(f"block failed with status {st[0]}")
        lg = None if self.config.method.kind == 0 else float(lag[0])
        return KrigingEstimate(float(v[0]), float(var[0]), lg)

    def predict_block_grid(self, grid: "Grid2D", block_size, discr, precision="fp64"):
        """predict_block_grid (kriging.rs:790-803): (values, variances)."""
        offsets = block_offsets(block_size, discr)
        v, var, _, _ = self.predict_block_many(grid.centers(), offsets, precision)
        return v, var

    def predict(self, target, precision="fp64") -> KrigingEstimate:
        """predict (kriging.rs:494-501): raises the per-target error."""
        v, var, lag, st = self.predict_many(np.asarray(target, dtype=np.float64).reshape(1, -1),
                                            precision)
        if st[0] != 0:
            raise _ERRORS.get(int(st[0]), GeostatError)(f"target failed with status {st[0]}")
        lg = None if self.config.method.kind == 0 else float(lag[0])
        return KrigingEstimate(float(v[0]), float(var[0]), lg)

    def predict_grid(self, grid: "Grid2D", precision="fp64"):
        """predict_grid (kriging.rs:783-786): (values, variances) in
        storage order iy*nx + ix."""
        v, var, _, _ = self.predict_many(grid.centers(), precision)
        return v, var

    def _check_device_args(self, targets, value, variance, lagrange, status):
        """Shape, dtype, layout and device checks of the device-path buffers,
        made before anything is enqueued: the kernels index them as
        contiguous float64 (T, dim) / (T,) arrays and int32 statuses on the
        handle's GPU."""
        import torch
        dev = self.ctx.device

        def tensor(x, name, dtype):
            if not isinstance(x, torch.Tensor):
                raise InvalidParameter(f"{name} must be a torch tensor on cuda:{dev}")
            if x.dtype != dtype:
                raise InvalidParameter(f"{name} must be {dtype}, got {x.dtype}")
            if x.device.type != "cuda" or x.device.index != dev:
                raise InvalidParameter(f"{name} must be on cuda:{dev}, got {x.device}")
            if not x.is_contiguous():
                raise InvalidParameter(f"{name} must be contiguous")

        tensor(targets, "targets", torch.float64)
        if targets.dim() != 2 or targets.shape[1] != self.dim:
            raise DimensionMismatch(f"targets must be (T, {self.dim}), got {tuple(targets.shape)}")
        T = int(targets.shape[0])
        for x, name, dtype in ((value, "value", torch.float64),
                               (variance, "variance", torch.float64),
                               (lagrange, "lagrange", torch.float64),
                               (status, "status", torch.int32)):
            if x is None and name in ("lagrange", "status"):
                continue
            tensor(x, name, dtype)
            if x.numel() < T:
                raise InvalidParameter(f"{name} holds {x.numel()} entries, need {T}")
        return T

    def predict_many_device(self, targets, value, variance, lagrange=None, status=None,
                            precision="fp64", stream=None):
        """Enqueues predict_many on torch tensors resident on the handle's GPU
        (targets (T, dim) float64, value / variance / lagrange float64 with at
        least T entries, status int32) on `stream` (a hipStream_t as an int,
        e.g. a torch stream's cuda_stream; None or 0 is the null stream,
        torch's default stream); returns immediately.  Bad buffers raise
        before any work is enqueued."""
        if precision not in _PREC:
            raise InvalidParameter(f"unknown precision {precision!r}")
        T = self._check_device_args(targets, value, variance, lagrange, status)
        ptr = (lambda x: None if x is None else x.data_ptr())
        _check(_abi.lib().gs_kriging_predict_many_device(
            self._h, ptr(targets), T, _PREC[precision], ptr(value), ptr(variance),
            ptr(lagrange), ptr(status), stream))

    def capture_graph(self, targets, value, variance, lagrange=None, status=None,
                      precision="fp64") -> "PredictGraph":
        """predict_many_device captured as a HIP graph over these tensors
        (gs_kriging_graph_create): runs the call once, then every
        PredictGraph.launch repeats it on whatever `targets` holds, as one
        graph launch.  The tensors must stay alive (the graph keeps
        references) and keep their shapes; launches of one graph must not
        overlap."""
        if precision not in _PREC:
            raise InvalidParameter(f"unknown precision {precision!r}")
        T = self._check_device_args(targets, value, variance, lagrange, status)
        ptr = (lambda x: None if x is None else x.data_ptr())
        h = C.c_void_p()
        _check(_abi.lib().gs_kriging_graph_create(
            self._h, ptr(targets), T, _PREC[precision], ptr(value), ptr(variance),
            ptr(lagrange), ptr(status), C.byref(h)))
        return PredictGraph(h, self, (targets, value, variance, lagrange, status))

    def set_timing(self, enable: bool) -> None:
        """Records HIP events around the search and solve phases of every
        subsequent predict call (measurement hook)."""
        _check(_abi.lib().gs_kriging_set_timing(self._h, 1 if enable else 0))

    def timing_summary(self):
        """(calls, total search ms, total solve ms) since the last summary."""
        calls = C.c_int64()
        a = C.c_double()
        b = C.c_double()
        _check(_abi.lib().gs_kriging_timing_summary(self._h, C.byref(calls), C.byref(a),
                                                    C.byref(b)))
        return int(calls.value), float(a.value), float(b.value)

    def neighbors(self, targets):
        """Kriging::neighbors (kriging.rs:441-490) for each target: (idx, count)
        with idx an (T, max_neighbors) int32 array padded with -1."""
        t = _f64(targets).reshape(-1, self.dim)
        T = t.shape[0]
        k = max(self.max_neighbors, 1)
        idx = np.empty((T, k), dtype=np.int32)
        cnt = np.empty(T, dtype=np.int32)
        _check(_abi.lib().gs_kriging_neighbors(self._h, _abi.dptr(t), T, k, _abi.i32ptr(idx),
                                               _abi.i32ptr(cnt)))
        return idx, cnt


class PredictGraph:
    """A captured predict call (Kriging.capture_graph)."""

    def __init__(self, h, owner, buffers):
        self._h = h
        self._owner = owner          # the predictor and buffers outlive the graph
        self._buffers = buffers

    def launch(self, stream=None) -> None:
        """Enqueues the captured call on `stream` (a hipStream_t as an int;
        None or 0 is the null stream); returns immediately."""
        if not self._h:
            raise InvalidParameter("graph is closed")
        _check(_abi.lib().gs_graph_launch(self._h, stream))

    def close(self) -> None:
        if getattr(self, "_h", None):
            _abi.lib().gs_graph_free(self._h)
            self._h = None

    def __del__(self):
        self.close()


def block_offsets(size, n) -> np.ndarray:
    """block_offsets (kriging.rs:817-838): regular n[0] x n[1] discretization
    of a size[0] x size[1] block, x fastest."""
    if n[0] == 0 or n[1] == 0:
        raise InvalidParameter("block discretization needs at least 1 point per axis")
    if not (size[0] > 0.0) or not (size[1] > 0.0):
        raise InvalidParameter(f"block size must be positive, got {list(size)}")
    out = np.empty((n[0] * n[1], 2))
    k = 0
    for iy in range(n[1]):
        for ix in range(n[0]):
            out[k, 0] = ((ix + 0.5) / n[0] - 0.5) * size[0]
            out[k, 1] = ((iy + 0.5) / n[1] - 0.5) * size[1]
            k += 1
    return out


# ------------------------------------------------------------------ grids --
@dataclass
class Grid2D:
    """Grid2D (grid.rs:11-96): cell centres x0 + (ix + 0.5) dx, order iy*nx+ix."""
    x0: float
    y0: float
    dx: float
    dy: float
    nx: int
    ny: int

    @classmethod
    def from_bbox(cls, mn, mx, nx, ny):
        if nx == 0 or ny == 0:
            raise InvalidParameter("grid must have at least one cell per axis")
        if not (mx[0] > mn[0]) or not (mx[1] > mn[1]):
            raise InvalidParameter("degenerate bbox")
        return cls(mn[0], mn[1], (mx[0] - mn[0]) / nx, (mx[1] - mn[1]) / ny, nx, ny)

    def n_cells(self) -> int:
        return self.nx * self.ny

    def centers(self) -> np.ndarray:
        ix = np.arange(self.nx, dtype=np.float64)
        iy = np.arange(self.ny, dtype=np.float64)
        cx = self.x0 + (ix + 0.5) * self.dx
        cy = self.y0 + (iy + 0.5) * self.dy
        out = np.empty((self.ny, self.nx, 2))
        out[:, :, 0] = cx[None, :]
        out[:, :, 1] = cy[:, None]
        return out.reshape(-1, 2)


@dataclass
class Grid3D:
    """Grid3D (grid.rs:127-202): x fastest, then y, then z."""
    x0: float
    y0: float
    z0: float
    dx: float
    dy: float
    dz: float
    nx: int
    ny: int
    nz: int

    @classmethod
    def from_bbox(cls, mn, mx, nx, ny, nz):
        if nx == 0 or ny == 0 or nz == 0:
            raise InvalidParameter("grid must have at least one cell per axis")
        for d in range(3):
            if not (mx[d] > mn[d]):
                raise InvalidParameter("degenerate bbox")
        return cls(mn[0], mn[1], mn[2], (mx[0] - mn[0]) / nx, (mx[1] - mn[1]) / ny,
                   (mx[2] - mn[2]) / nz, nx, ny, nz)

    def n_cells(self) -> int:
        return self.nx * self.ny * self.nz

    def centers(self) -> np.ndarray:
        cx = self.x0 + (np.arange(self.nx, dtype=np.float64) + 0.5) * self.dx
        cy = self.y0 + (np.arange(self.ny, dtype=np.float64) + 0.5) * self.dy
        cz = self.z0 + (np.arange(self.nz, dtype=np.float64) + 0.5) * self.dz
        out = np.empty((self.nz, self.ny, self.nx, 3))
        out[..., 0] = cx[None, None, :]
        out[..., 1] = cy[None, :, None]
        out[..., 2] = cz[:, None, None]
        return out.reshape(-1, 3)


# ---------------------------------------------------------------- Vecchia --
class VecchiaPlan:
    """VecchiaPlan (vecchia.rs:223-228) resident on the device."""

    def __init__(self, coords, m: int, order=None, ctx: Optional[Context] = None):
        coords = _f64(coords)
        if coords.ndim != 2 or coords.shape[1] not in (2, 3):
            raise DimensionMismatch(f"coords must be (n, 2|3), got {coords.shape}")
        self.ctx = ctx or Context.default()
        self.dim = coords.shape[1]
        self.n = coords.shape[0]
        self.m = int(m)
        o = None
        if order is not None:
            o = np.ascontiguousarray(np.asarray(order, dtype=np.int64))
            if o.shape[0] != self.n:
                raise DimensionMismatch(f"order has {o.shape[0]} entries for {self.n} points")
        h = C.c_void_p()
        _check(_abi.lib().gs_vecchia_plan_new(self.ctx._h, self.dim, _abi.dptr(coords), self.n,
                                              self.m, None if o is None else _abi.i64ptr(o),
                                              C.byref(h)))
        self._h = h
        self._release = _abi.lib().gs_vecchia_plan_free   # kept: module globals may be gone at exit

    def __del__(self):
        h = getattr(self, "_h", None)
        release = getattr(self, "_release", None)
        if h and release is not None:
            release(h)
            self._h = None

    def arrays(self):
        """(order, neighbours[n, m] padded with -1, counts)."""
        order = np.empty(self.n, dtype=np.int64)
        nb = np.empty((self.n, self.m), dtype=np.int32)
        cnt = np.empty(self.n, dtype=np.int32)
        _check(_abi.lib().gs_vecchia_plan_get(self._h, _abi.i64ptr(order), _abi.i32ptr(nb),
                                              _abi.i32ptr(cnt)))
        return order, nb, cnt

    @property
    def order(self):
        return self.arrays()[0]

    def loglik_batch(self, values, models: Sequence[VariogramModel], precision="fp64"):
        """Batched loglik_with_plan (vecchia.rs:285-336): (loglik, status)
        arrays; status 3 where the reference returns Err(SingularSystem)."""
        v = _f64(values).reshape(-1)
        if v.shape[0] != self.n:
            raise DimensionMismatch(f"{v.shape[0]} values for {self.n} points")
        arr = (_abi.gs_model * len(models))(*[mm._c() for mm in models])
        out = np.empty(len(models))
        st = np.empty(len(models), dtype=np.int32)
        _check(_abi.lib().gs_vecchia_loglik_batch(self._h, _abi.dptr(v), arr, len(models),
                                                  _PREC[precision], _abi.dptr(out),
                                                  _abi.i32ptr(st)))
        return out, st

    def loglik_grad_batch(self, values, models: Sequence[VariogramModel], precision="fp64"):
        """Batched loglik_grad_with_plan (vecchia.rs:416-566): (loglik,
        grad (n_models, 3) as d/dnugget, d/dpsill, d/drange, status); status
        GS_ERR_INVALID_PARAMETER where the reference returns None (kinds
        without a closed-form range derivative, several structures,
        anisotropy), 3 where a conditional is singular."""
        v = _f64(values).reshape(-1)
        if v.shape[0] != self.n:
            raise DimensionMismatch(f"{v.shape[0]} values for {self.n} points")
        arr = (_abi.gs_model * len(models))(*[mm._c() for mm in models])
        ll = np.empty(len(models))
        grad = np.empty((len(models), 3))
        st = np.empty(len(models), dtype=np.int32)
        _check(_abi.lib().gs_vecchia_loglik_grad_batch(self._h, _abi.dptr(v), arr, len(models),
                                                       _PREC[precision], _abi.dptr(ll),
                                                       _abi.dptr(grad), _abi.i32ptr(st)))
        return ll, grad, st

    def loglik(self, values, model: VariogramModel, precision="fp64") -> float:
        out, st = self.loglik_batch(values, [model], precision)
        if st[0] != 0:
            raise SingularSystem("singular conditioning system")
        return float(out[0])

    def mle(self, values, kind) -> "VecchiaFit":
        """mle_fit_with_plan (vecchia.rs:952-1081): maximum-likelihood fit
        of nugget + one structure of `kind` over this plan."""
        v = _f64(values).reshape(-1)
        if v.shape[0] != self.n:
            raise DimensionMismatch(f"{v.shape[0]} values for {self.n} points")
        code, param = kind
        fit = _abi.gs_model()
        ll = C.c_double()
        _check(_abi.lib().gs_vecchia_mle(self._h, _abi.dptr(v), int(code), float(param),
                                         C.byref(fit), C.byref(ll)))
        s = fit.structures[0]
        model = VariogramModel(fit.nugget, [Structure((int(s.kind), float(s.param)), s.sill,
                                                      s.range)])
        return VecchiaFit(model, float(ll.value))


@dataclass
class VecchiaFit:
    """VecchiaFit (vecchia.rs:904-911): fitted model and its maximized
    Vecchia log-likelihood."""
    model: VariogramModel
    loglik: float


def vecchia_mle(coords, values, kind, m: int, order=None,
                ctx: Optional[Context] = None) -> VecchiaFit:
    """vecchia_mle (vecchia.rs:921-932): plan + maximum-likelihood fit."""
    code = kind[0]
    if code == ModelKind.Power(1.0)[0]:
        raise InvalidParameter("Power models cannot be fitted by Vecchia MLE")
    coords = _f64(coords)
    if code == ModelKind.CIRCULAR[0] and coords.shape[1] > 2:
        raise InvalidParameter("Circular is not a valid covariance in 3 dimensions")
    return VecchiaPlan(coords, m, order, ctx).mle(values, kind)


def maxmin_order(coords) -> np.ndarray:
    """maxmin_order (vecchia.rs:144-215): the maxmin ordering of the points
    (host computation)."""
    c = _f64(coords)
    if c.ndim != 2 or c.shape[1] not in (2, 3):
        raise DimensionMismatch(f"coords must be (n, 2|3), got {c.shape}")
    out = np.empty(c.shape[0], dtype=np.int64)
    _check(_abi.lib().gs_maxmin_order(c.shape[1], _abi.dptr(c), c.shape[0], _abi.i64ptr(out)))
    return out


def maxmin_order_gpu(coords, ctx: Optional[Context] = None, info: bool = False):
    """maxmin_order (vecchia.rs:144-215) computed on the GPU; the same order
    as maxmin_order.  With info=True also returns (rounds, threshold
    retries, points ordered on the GPU before the host finished)."""
    ctx = ctx or Context.default()
    c = _f64(coords)
    if c.ndim != 2 or c.shape[1] not in (2, 3):
        raise DimensionMismatch(f"coords must be (n, 2|3), got {c.shape}")
    out = np.empty(c.shape[0], dtype=np.int64)
    inf = np.zeros(3, dtype=np.int64)
    _check(_abi.lib().gs_maxmin_order_gpu(ctx._h, c.shape[1], _abi.dptr(c), c.shape[0],
                                           _abi.i64ptr(out), _abi.i64ptr(inf)))
    return (out, tuple(int(v) for v in inf)) if info else out


def vecchia_plan(coords, m: int, order=None, ctx: Optional[Context] = None) -> VecchiaPlan:
    """vecchia_plan (vecchia.rs:232-282)."""
    return VecchiaPlan(coords, m, order, ctx)


def vecchia_loglik(coords, values, model: VariogramModel, m: int, order=None,
                   ctx: Optional[Context] = None) -> float:
    """vecchia_loglik (vecchia.rs:874-884)."""
    ctx = ctx or Context.default()
    coords = _f64(coords)
    values = _f64(values).reshape(-1)
    o = None if order is None else np.ascontiguousarray(np.asarray(order, dtype=np.int64))
    cm = model._c()
    out = C.c_double()
    _check(_abi.lib().gs_vecchia_loglik(ctx._h, coords.shape[1], _abi.dptr(coords),
                                        _abi.dptr(values), coords.shape[0], C.byref(cm), int(m),
                                        None if o is None else _abi.i64ptr(o), C.byref(out)))
    return float(out.value)


# ----------------------------------------------------------- request loop --
class Ticket:
    """A submitted request: holds its buffers until Server.wait returns."""

    def __init__(self, tid: int, kind: str, keep: tuple, outputs: tuple):
        self.id = tid
        self.kind = kind
        self._keep = keep
        self.outputs = outputs
        self.queue_ms = None
        self.service_ms = None


class Server:
    """Request loop over resident predictors and plans (gs_server_*):
    n_slots concurrent slots, admission by device-memory footprint."""

    def __init__(self, n_slots: int = 4, mem_budget: int = 16 << 30,
                 ctx: Optional[Context] = None):
        self.ctx = ctx or Context.default()
        h = C.c_void_p()
        _check(_abi.lib().gs_server_create(self.ctx._h, int(n_slots), int(mem_budget),
                                           C.byref(h)))
        self._h = h
        self._release = _abi.lib().gs_server_destroy   # kept: module globals may be gone at exit

    def close(self) -> None:
        h = getattr(self, "_h", None)
        release = getattr(self, "_release", None)
        if h and release is not None:
            release(h)
            self._h = None

    def __del__(self):
        self.close()

    def submit_krige(self, kriging: "Kriging", targets, precision="fp64") -> Ticket:
        t = _f64(targets).reshape(-1, kriging.dim)
        T = t.shape[0]
        v, var, lag = np.empty(T), np.empty(T), np.empty(T)
        st = np.empty(T, dtype=np.int32)
        tid = C.c_int64()
        _check(_abi.lib().gs_server_submit_krige(self._h, kriging._h, _abi.dptr(t), T,
                                                 _PREC[precision], _abi.dptr(v), _abi.dptr(var),
                                                 _abi.dptr(lag), _abi.i32ptr(st), C.byref(tid)))
        return Ticket(tid.value, "krige", (kriging, t), (v, var, lag, st))

    def submit_vecchia(self, plan: "VecchiaPlan", values, models, precision="fp64") -> Ticket:
        v = _f64(values).reshape(-1)
        arr = (_abi.gs_model * len(models))(*[mm._c() for mm in models])
        out = np.empty(len(models))
        st = np.empty(len(models), dtype=np.int32)
        tid = C.c_int64()
        _check(_abi.lib().gs_server_submit_vecchia(self._h, plan._h, _abi.dptr(v), arr,
                                                   len(models), _PREC[precision], _abi.dptr(out),
                                                   _abi.i32ptr(st), C.byref(tid)))
        return Ticket(tid.value, "vecchia", (plan, v, arr), (out, st))

    def wait(self, ticket: Ticket):
        """Blocks until the request is done; returns its outputs ((value,
        variance, lagrange, status) or (loglik, status)) and records
        ticket.queue_ms / ticket.service_ms."""
        q, sv = C.c_double(), C.c_double()
        rc = _abi.lib().gs_server_wait(self._h, ticket.id, C.byref(q), C.byref(sv))
        ticket.queue_ms, ticket.service_ms = q.value, sv.value
        ticket._keep = None
        _check(rc)
        return ticket.outputs
