"""Builds libvvr.so (the HIP kernels for gfx950 and the C ABI) in-tree.

    python -m volview_amd.build          # or: from volview_amd.build import build; build()

Every HIP source is compiled for gfx950 with packed-FP32 instructions disabled and
without floating-point contraction;
host sources are C++17, compiled by the host compiler. Objects go to volview_amd/_build/, the library to
volview_amd/libvvr.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libvvr.so"
ARCH = os.environ.get("VVR_OFFLOAD_ARCH", "gfx950")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
CXX = os.environ.get("CXX", "g++")

HIP_SOURCES = ["kernels_volume.hip", "kernels_render.hip", "kernels_wavefront.hip", "kernels_lighting.hip",
               "kernels_denoise.hip", "kernels_post.hip", "kernels_unet.hip"]
HOST_SOURCES = ["transfer.cpp", "frame.cpp", "phantom.cpp", "renderer.cpp", "server.cpp", "environment.cpp", "unet.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", f"-I{PKG.parent / 'include'}"]
DEVICE = [
    f"--offload-arch={ARCH}",
    # Packed FP32 (v_pk_*_f32) stays off: some of its forms have given wrong results.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # No contraction of a * b + c into one FMA: every operation is rounded on its
    # own, in source order, as in the FP32 oracle, so the path tracer's random
    # walks take the oracle's decisions (explicit fmaf calls stay fused).
    "-ffp-contract=off",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the MI355X kernels cannot be built")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")


def build(verbose: bool = False, defines: tuple = (), out: Path | None = None) -> Path:
    """Compiles every source and links the library (to `out`, default libvvr.so);
    `defines` are extra -D flags for the HIP sources."""
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = BUILD / (src + ".o")
        jobs.append([hipcc, "-x", "hip", *COMMON, *DEVICE, *[f"-D{d}" for d in defines], "-c", str(CSRC / src),
                     "-o", str(obj)])
    for src in HOST_SOURCES:
        obj = BUILD / (src + ".o")
        jobs.append([CXX, *COMMON, "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", "-c", str(CSRC / src), "-o", str(obj)])
    workers = min(len(jobs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for cmd in jobs:
            if verbose:
                print(" ".join(cmd), flush=True)
        list(ex.map(_run, jobs))
    objs = [str(BUILD / (s + ".o")) for s in HIP_SOURCES + HOST_SOURCES]
    target = Path(out) if out is not None else LIB
    tmp = target.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs, "-lpthread",
          f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, target)
    return target


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
