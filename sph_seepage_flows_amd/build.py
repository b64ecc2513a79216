"""Builds libsf_seepage.so (gfx950 HIP kernels + C-ABI) in-tree with hipcc.

The library is self-contained: HIP runtime only, no torch, no Python.
`python -m sph_seepage_flows_amd.build` or `build()` from __graft_entry__.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "_lib"
BUILD = PKG / "_build"
LIB = LIBDIR / "libsf_seepage.so"
INCLUDE = PKG.parent / "include"

SOURCES = ["sf_searcher.hip", "sf_solver.hip", "sf_adjoint.hip", "sf_views.hip", "sf_grid.hip", "sf_api.cpp", "sf_diff.cpp", "sf_emitter.cpp",
           "sf_host.cpp", "sf_scene.cpp"]
ARCH = os.environ.get("SF_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (set HIPCC)")


def common_flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        # no packed FP32 instructions anywhere
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        # IEEE division / sqrt where source code writes them (grid hash parity)
        "-fhip-fp32-correctly-rounded-divide-sqrt",
        "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
        f"-I{INCLUDE}",
    ]


def _compile(src: Path, obj: Path) -> tuple[str, int, str]:
    lang = ["-x", "hip"] if src.suffix == ".hip" else []
    cmd = [hipcc(), *common_flags(), *lang, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return " ".join(cmd), r.returncode, r.stdout + r.stderr


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    LIBDIR.mkdir(parents=True, exist_ok=True)
    objs = []
    jobs = []
    with cf.ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        for s in SOURCES:
            src = CSRC / s
            obj = BUILD / (s + ".o")
            objs.append(obj)
            jobs.append(ex.submit(_compile, src, obj))
        for j in jobs:
            cmd, rc, out = j.result()
            if verbose or rc:
                print(cmd)
                if out.strip():
                    print(out)
            if rc:
                raise RuntimeError(f"hipcc failed ({rc}): {cmd}\n{out}")
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-lz", "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}{r.stderr}")
    os.replace(tmp, LIB)
    # the main.cpp-style example driver on the drop-in C++ layer
    drv = LIBDIR / "seepage_main"
    cmd = [hipcc(), "-O2", "-std=c++17", f"-I{INCLUDE}", str(PKG / "cpp" / "seepage_main.cpp"), f"-L{LIBDIR}",
           "-lsf_seepage", "-Wl,-rpath,$ORIGIN", "-o", str(drv)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}{r.stderr}")
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
