"""The host runtime built with AddressSanitizer and UBSan (build/asan/
libgeostat_hip.so: gs_host.cpp and gs_mle.cpp instrumented on the host side,
the kernels unchanged), run in a child process with the clang ASan runtime
preloaded, halting on the first report.  Without a GPU this covers the host
computations: the maxmin traversal on random, clustered, lattice and
duplicated point sets in 2-D and 3-D (against the oracle), model validation,
and the argument checks of the entry points that reject before touching a
device."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "geostat-rs_amd"
sys.path.insert(0, str(PKG))
import build_lib  # noqa: E402

HOST_SCRIPT = r"""
import sys
sys.path[:0] = [{root!r}, {pkg!r}]
import numpy as np
import geostat_hip as gh
from geostat_hip import _abi
from oracle import refcpu
assert str(_abi.LIB_PATH).endswith("asan/libgeostat_hip.so"), _abi.LIB_PATH
maps = open("/proc/self/maps").read()
assert "libclang_rt.asan" in maps, "ASan runtime not loaded"
r = np.random.default_rng(3)
sets = [r.uniform(0, 1, (3000, 2)), r.uniform(0, 1, (1500, 3))]
k = 40
g = np.meshgrid(np.arange(k, dtype=float), np.arange(k, dtype=float), indexing="ij")
lat = np.column_stack([a.ravel() for a in g]); r.shuffle(lat); sets.append(lat)
c6 = r.uniform(0, 1, (6, 2))
sets.append(np.unique(c6[r.integers(6, size=2500)] + 0.02 * r.standard_normal((2500, 2)), axis=0))
d = r.uniform(0, 1, (500, 2)); sets.append(r.permutation(np.concatenate([d, d])))
for c in sets:
    assert np.array_equal(gh.maxmin_order(c), np.asarray(refcpu.maxmin_order(c), dtype=np.int64))
assert gh.maxmin_order(np.zeros((0, 2))).size == 0 and list(gh.maxmin_order([[1.0, 2.0]])) == [0]
for bad in ([[0.0, np.nan], [1.0, 1.0]],):
    try:
        gh.maxmin_order(bad); raise SystemExit("accepted a NaN coordinate")
    except gh.InvalidParameter:
        pass
L = _abi.lib()
ok = gh.VariogramModel(0.1, [gh.Structure(gh.ModelKind.MATERN25, 1.0, 0.3)])
assert L.gs_model_validate(ok._c()) == 0
bad = gh.VariogramModel(0.1, [gh.Structure(gh.ModelKind.SPHERICAL, -1.0, 0.3)])
assert L.gs_model_validate(bad._c()) != 0 and L.gs_last_error()
assert L.gs_abi_version() == 2
print("host-sanitized ok")
"""


def _san_env() -> dict:
    """The sanitized library, with the ASan runtime first among the preloaded
    libraries (ahead of any the environment already preloads)."""
    pre = os.environ.get("LD_PRELOAD", "")
    return {"GEOSTAT_HIP_LIB": str(build_lib.ASAN_LIB),
            "LD_PRELOAD": str(build_lib.asan_runtime()) + (":" + pre if pre else ""),
            "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0:protect_shadow_gap=0",
            "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}


def _run(script: str, timeout: int):
    env = dict(os.environ)
    env.update(_san_env())
    return subprocess.run([sys.executable, "-c", script], env=env, capture_output=True,
                          text=True, timeout=timeout)


@pytest.fixture(scope="module")
def sanitized_lib():
    if not build_lib.ASAN_LIB.exists() or (
            build_lib.ASAN_LIB.stat().st_mtime < build_lib.LIB.stat().st_mtime):
        build_lib.build_sanitized()
    return build_lib.ASAN_LIB


def test_host_runtime_under_sanitizers(sanitized_lib):
    syms = subprocess.run(["nm", "-D", str(sanitized_lib)], capture_output=True, text=True).stdout
    assert "__asan_report" in syms and "__ubsan_handle" in syms   # instrumented
    r = _run(HOST_SCRIPT.format(root=str(ROOT), pkg=str(PKG)), 600)
    assert r.returncode == 0 and "host-sanitized ok" in r.stdout, (r.stdout[-2000:], r.stderr[-6000:])
