"""Builds libgeostat_hip.so (gfx950) and libgs_synth.so in-tree.

The solve kernels are templates over (dimension, padded system size,
drift terms); every instantiation gets its own generated translation unit
so the fully unrolled kernels compile in parallel.  Output goes to
geostat-rs_amd/build/ (git-ignored; the built .so files travel with the
tree to the GPU host).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
BUILD = HERE / "build"
GEN = BUILD / "gen"
OBJ = BUILD / "obj"
LIB = BUILD / "libgeostat_hip.so"
SYNTH = BUILD / "libgs_synth.so"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 (v_pk_*_f32) stays off for every kernel in this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
EXACT = ["-ffp-contract=off"]     # search + host: no a*b+c fusion (bit-exact distances)

# (D, NB, M) instantiations of the kriging kernel: M = drift terms
# (0 simple, 1 ordinary, 1+D universal deg 1, 1+D+D(D+1)/2 universal deg 2).
# Group sizes G (lanes per system) compiled for each padded size NB; the
# runtime tuner picks one per configuration (gs_internal.hpp group_size).
# The lean model class (spherical / exponential / Gaussian / closed-form
# Matern) gets both group sizes; the general class only the first.
GROUPS = {16: (2, 4), 32: (8, 16), 64: (32, 64)}
KRIGE = [(d, nb, m, g, lean, 0) for d, ms in ((2, (0, 1, 3, 6)), (3, (0, 1, 4, 10)))
         for nb in (16, 32, 64) for m in ms for lean in (1, 0)
         for g in (GROUPS[nb] if lean else GROUPS[nb][:1])]
# Extended systems (measurement error, external drift with 1-3 covariates,
# block kriging): first group size only.
KRIGE += [(d, nb, m, GROUPS[nb][0], lean, 1)
          for d, ms in ((2, (0, 1, 2, 3, 4, 6)), (3, (0, 1, 2, 3, 4, 10)))
          for nb in (16, 32, 64) for m in ms for lean in (1, 0)]
VECCHIA = [(d, nb, g, lean) for d in (2, 3) for nb in (16, 32, 64) for lean in (1, 0)
           for g in (GROUPS[nb] if lean else GROUPS[nb][:1])]
# FP32-refined kernels: lean models; kriging with M = 0 (simple) or 1
# (ordinary); Vecchia conditionals.
KRIGE32R = [(d, nb, m) for d in (2, 3) for nb in (16, 32, 64) for m in (0, 1)]
VECCHIA32R = [(d, nb) for d in (2, 3) for nb in (16, 32, 64)]
# Vecchia likelihood on the matrix cores: single lean structure, by
# (dimension, anisotropy, kind 0..4)
VMFMA = [(d, an, k) for d in (2, 3) for an in (0, 1) for k in range(5)]
# Vecchia gradient on the matrix cores: (dimension, kind)
VGMFMA = [(d, k) for d in (2, 3) for k in range(5)]
# simple / ordinary kriging on the matrix cores: (dimension, anisotropy, kind, M)
KMFMA = [(d, an, k, m) for d in (2, 3) for an in (0, 1) for k in range(5) for m in (0, 1)]
# 33-64-neighbour kriging on the matrix cores: (dimension, covariance
# signature, M).  Signatures: 0-4 one isotropic lean structure, 8-12 one
# anisotropic, 16 / 17 exponential + spherical sharing an anisotropic /
# isotropic frame; M: simple, ordinary, linear and (2-D) quadratic drift.
K64_SIGS = (0, 1, 2, 3, 4, 8, 9, 10, 11, 12, 16, 17)
K64 = [(2, g, m) for g in K64_SIGS for m in (0, 1, 3, 6)] + \
      [(3, g, m) for g in K64_SIGS for m in (0, 1, 4)]
# Vecchia gradient: lean single-structure models, conditioning sets <= 32,
# two row blocks per lane
VGRAD = [(d, nb, nb // 2) for d in (2, 3) for nb in (16, 32)]


def _gen_sources() -> list[Path]:
    GEN.mkdir(parents=True, exist_ok=True)
    out = []
    for d, nb, m, g, lean, x in KRIGE:
        p = GEN / (f"krige_d{d}_nb{nb}_m{m}_g{g}_{'lean' if lean else 'gen'}"
                   f"{'_x' if x else ''}.hip")
        txt = (f'#include "{CSRC}/gs_solve_impl.hpp"\n'
               f"namespace gs {{\ntemplate hipError_t launch_krige_inst<{d}, {nb}, {m}, {g}, "
               f"{'true' if lean else 'false'}, {'true' if x else 'false'}>"
               f"(const SolveParams &, int64_t, hipStream_t);\n}}\n")
        if not p.exists() or p.read_text() != txt:
            p.write_text(txt)
        out.append(p)
    for d, nb, g, lean in VECCHIA:
        p = GEN / f"vecchia_d{d}_nb{nb}_g{g}_{'lean' if lean else 'gen'}.hip"
        txt = (f'#include "{CSRC}/gs_solve_impl.hpp"\n'
               f"namespace gs {{\ntemplate hipError_t launch_vecchia_inst<{d}, {nb}, {g}, "
               f"{'true' if lean else 'false'}>"
               f"(const VecchiaParams &, int64_t, hipStream_t);\n}}\n")
        if not p.exists() or p.read_text() != txt:
            p.write_text(txt)
        out.append(p)
    for d, an, k in VMFMA:
        p = GEN / f"vmfma_d{d}_a{an}_k{k}.hip"
        txt = (f'#include "{CSRC}/gs_vecchia_mfma_impl.hpp"\n'
               f"namespace gs {{\ntemplate hipError_t launch_vecchia_mfma_inst<{d}, "
               f"{'true' if an else 'false'}, {k}>"
               f"(const VecchiaParams &, int64_t, hipStream_t);\n}}\n")
        if not p.exists() or p.read_text() != txt:
            p.write_text(txt)
        out.append(p)
    for d, k in VGMFMA:
        p = GEN / f"vgmfma_d{d}_k{k}.hip"
        txt = (f'#include "{CSRC}/gs_vecchia_grad_mfma_impl.hpp"\n'
               f"namespace gs {{\ntemplate hipError_t launch_vecchia_grad_mfma_inst<{d}, {k}>"
               f"(const VecchiaParams &, int64_t, hipStream_t);\n}}\n")
        if not p.exists() or p.read_text() != txt:
            p.write_text(txt)
        out.append(p)
    for d, an, k, m in KMFMA:
        p = GEN / f"kmfma_d{d}_a{an}_k{k}_m{m}.hip"
        txt = (f'#include "{CSRC}/gs_vecchia_mfma_impl.hpp"\n'
               f"namespace gs {{\ntemplate hipError_t launch_krige_mfma_inst<{d}, "
               f"{'true' if an else 'false'}, {k}, {m}>"
               f"(const SolveParams &, int64_t, hipStream_t);\n}}\n")
        if not p.exists() or p.read_text() != txt:
            p.write_text(txt)
        out.append(p)
    for d, sig, m in K64:
        p = GEN / f"k64mfma_d{d}_s{sig}_m{m}.hip"
        txt = (f'#include "{CSRC}/gs_krige_mfma64_impl.hpp"\n'
               f"namespace gs {{\ntemplate hipError_t launch_krige_mfma64_inst<{d}, {sig}, {m}>"
               f"(const SolveParams &, int64_t, hipStream_t);\n}}\n")
        if not p.exists() or p.read_text() != txt:
            p.write_text(txt)
        out.append(p)
    for d, nb, g in VGRAD:
        p = GEN / f"vgrad_d{d}_nb{nb}_g{g}.hip"
        txt = (f'#include "{CSRC}/gs_solve_impl.hpp"\n'
               f"namespace gs {{\ntemplate hipError_t launch_vecchia_grad_inst<{d}, {nb}, {g}>"
               f"(const VecchiaParams &, int64_t, hipStream_t);\n}}\n")
        if not p.exists() or p.read_text() != txt:
            p.write_text(txt)
        out.append(p)
    for d, nb, m in KRIGE32R:
        p = GEN / f"krige32r_d{d}_nb{nb}_m{m}.hip"
        txt = (f'#include "{CSRC}/gs_fp32r_impl.hpp"\n'
               f"namespace gs {{\ntemplate hipError_t launch_krige_fp32r_inst<{d}, {nb}, {m}>"
               f"(const SolveParams &, int64_t, hipStream_t);\n}}\n")
        if not p.exists() or p.read_text() != txt:
            p.write_text(txt)
        out.append(p)
    for d, nb in VECCHIA32R:
        p = GEN / f"vecchia32r_d{d}_nb{nb}.hip"
        txt = (f'#include "{CSRC}/gs_fp32r_impl.hpp"\n'
               f"namespace gs {{\ntemplate hipError_t launch_vecchia_fp32r_inst<{d}, {nb}>"
               f"(const VecchiaParams &, int64_t, hipStream_t);\n}}\n")
        if not p.exists() or p.read_text() != txt:
            p.write_text(txt)
        out.append(p)
    return out


def _deps_mtime(src: Path) -> float:
    """Newest header a unit depends on: a generated solve unit includes
    only the kernel templates; every other unit may include any header."""
    kernel = ["gs_solve_impl.hpp", "gs_internal.hpp", "gs_device.hpp"]
    if src.parent == GEN:
        names = kernel + (["gs_fp32r_impl.hpp"] if "32r" in src.stem else []) + \
            (["gs_vecchia_mfma_impl.hpp"] if src.stem.startswith(("vmfma", "kmfma", "vgmfma", "k64mfma")) else []) + \
            (["gs_krige_mfma64_impl.hpp"] if src.stem.startswith("k64mfma") else []) + \
            (["gs_vecchia_grad_mfma_impl.hpp"] if src.stem.startswith("vgmfma") else [])
        hdrs = [CSRC / n for n in names]
    else:
        hdrs = [p for p in CSRC.iterdir() if p.suffix in (".hpp", ".h", ".inc")]
        hdrs.append(HERE.parent / "include" / "geostat_hip.h")
    return max(p.stat().st_mtime for p in hdrs)


def _compile(src: Path, extra: list[str], incremental: bool = False) -> Path:
    obj = OBJ / (src.stem + ".o")
    if incremental and obj.exists() and obj.stat().st_mtime > max(_deps_mtime(src), src.stat().st_mtime):
        return obj
    remarks = src.suffix == ".hip"
    cmd = [HIPCC, *COMMON, *extra, *(["-Rpass-analysis=kernel-resource-usage"] if remarks else []),
           "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if remarks:   # the compiler's register / spill / occupancy report per kernel
        (OBJ / (src.stem + ".rem")).write_text(r.stderr)
    return obj


_REMARK = re.compile(r"remark:\s+(Function Name|VGPRs|AGPRs|TotalSGPRs|ScratchSize \[bytes/lane\]|"
                     r"Occupancy \[waves/SIMD\]|SGPRs Spill|VGPRs Spill|LDS Size \[bytes/block\]):"
                     r"\s+(\S+)")


def resource_report() -> list[dict]:
    """Per kernel of the last build: VGPRs, AGPRs, SGPRs, spills, scratch
    bytes per lane, LDS bytes and occupancy, from the compiler's
    kernel-resource-usage remarks (kernels only, not device functions)."""
    rows, cur = [], None
    for rem in sorted(OBJ.glob("*.rem")):
        for line in rem.read_text().splitlines():
            m = _REMARK.search(line)
            if not m:
                continue
            key, val = m.groups()
            if key == "Function Name":
                cur = {"unit": rem.stem, "kernel": val}
                rows.append(cur)
            elif cur is not None:
                cur[key.split(" [")[0]] = int(val) if val.lstrip("-").isdigit() else val
    return [r for r in rows if "Occupancy" in r]


def build(jobs: int | None = None, verbose: bool = True, incremental: bool = False) -> Path:
    """Compiles every HIP source for gfx950 and links the library
    (incremental=True skips objects newer than every source and header)."""
    OBJ.mkdir(parents=True, exist_ok=True)
    units: list[tuple[Path, list[str]]] = [
        (CSRC / "gs_search.hip", EXACT),
        (CSRC / "gs_tile_search.hip", EXACT),
        (CSRC / "gs_solve.hip", []),
        (CSRC / "gs_host.cpp", EXACT),
        (CSRC / "gs_mle.cpp", EXACT),
        (CSRC / "gs_global.hip", EXACT),
        (CSRC / "gs_big.hip", EXACT),
        (CSRC / "gs_plan.hip", EXACT),
        (CSRC / "gs_maxmin.hip", EXACT),
        (CSRC / "gs_vecchia_mfma.hip", []),
    ]
    units += [(p, []) for p in _gen_sources()]
    jobs = jobs or max(1, min(16, os.cpu_count() or 4))
    objs: list[Path] = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(_compile, s, e, incremental): s for s, e in units}
        for f in cf.as_completed(futs):
            objs.append(f.result())
            if verbose:
                print(f"  built {futs[f].name}", flush=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *sorted(map(str, objs)), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # synthetic-data generator (host C, no GPU code)
    tmp = SYNTH.with_suffix(".so.tmp")
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-std=c11",
           "-o", str(tmp), str(CSRC / "gs_synth.c"), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"synth build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, SYNTH)
    return LIB


# Host-sanitized variant (build/asan/libgeostat_hip.so): the host runtime
# (gs_host.cpp, gs_mle.cpp: validation, search indexes, maxmin, plans,
# staging, request loop) built with AddressSanitizer and UBSan for the host
# side only; the kernels are the regular objects.  Loaded with the clang ASan
# runtime preloaded (tests/test_host_sanitized.py; GEOSTAT_HIP_LIB selects it).
ASAN = BUILD / "asan"
ASAN_LIB = ASAN / "libgeostat_hip.so"
HOST_SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=all"]
HOST_UNITS = ("gs_host.cpp", "gs_mle.cpp")


def asan_runtime() -> Path:
    """The clang ASan runtime that must be preloaded with the sanitized library."""
    out = subprocess.run([str(Path(HIPCC).parent.parent / "lib" / "llvm" / "bin" / "clang"),
                          "-print-file-name=libclang_rt.asan-x86_64.so"],
                         capture_output=True, text=True).stdout.strip()
    p = Path(out)
    if not p.is_absolute() or not p.exists():
        hits = sorted(Path("/opt/rocm/lib/llvm/lib/clang").glob("*/lib/linux/libclang_rt.asan-x86_64.so"))
        if not hits:
            raise RuntimeError("clang ASan runtime not found")
        p = hits[-1]
    return p


def build_sanitized(verbose: bool = False) -> Path:
    """Links build/asan/libgeostat_hip.so from the regular kernel objects and
    host-sanitized gs_host / gs_mle objects (build() first)."""
    (ASAN / "obj").mkdir(parents=True, exist_ok=True)
    objs = []
    for o in sorted(OBJ.glob("*.o")):
        if o.stem + ".cpp" in HOST_UNITS:
            continue
        objs.append(str(o))
    for unit in HOST_UNITS:
        src = CSRC / unit
        obj = ASAN / "obj" / (src.stem + ".o")
        cmd = [HIPCC, *COMMON, *EXACT, *HOST_SAN, "-g", "-c", str(src), "-o", str(obj)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        if verbose:
            print(f"  built {unit} (host sanitizers)", flush=True)
        objs.append(str(obj))
    tmp = ASAN_LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *HOST_SAN, "-shared-libsan",
           "-o", str(tmp), *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, ASAN_LIB)
    return ASAN_LIB


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


def print_report(only_spills: bool = True) -> None:
    cols = ("VGPRs", "AGPRs", "TotalSGPRs", "SGPRs Spill", "VGPRs Spill", "ScratchSize",
            "LDS Size", "Occupancy")
    print("unit kernel " + " ".join(c.replace(" ", "_") for c in cols))
    for r in resource_report():
        if only_spills and not (r.get("SGPRs Spill") or r.get("VGPRs Spill") or r.get("ScratchSize")):
            continue
        print(r["unit"], r["kernel"][:60], *(r.get(c, "-") for c in cols))


if __name__ == "__main__":
    if "--report" in sys.argv:       # resource usage of the last build's kernels
        print_report(only_spills="--all" not in sys.argv)
        sys.exit(0)
    if "--clean" in sys.argv:
        clean()
    print(build(incremental="--incremental" in sys.argv))
    if "--sanitized" in sys.argv:
        print(build_sanitized(verbose=True))
