"""In-tree build of the native libraries.

Produces, next to this file:
  libisti.so        HIP kernels for gfx950 + the C ABI of include/isti.h
  libisti_scene.so  seeded synthetic keyframe generator (host C++, OpenMP)
  libisti_io.so     reader of the reference's keyframe recordings (host C++, zlib)
  splinter          the host driver executable (C++, links the three libraries)

Run as ``python -m imagebasedspatiotemporalinterpolation_amd.build``. Every
call recompiles (the sources are small). hipcc cross-compiles gfx950 code
without a GPU.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"

ARCH = os.environ.get("ISTI_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# Numerics contract: no FMA contraction, IEEE fp32 division and sqrt (the
# splat keys embed raw depth bits, so the kernels must round exactly like the
# CPU restatement). Packed-fp32 instructions are disabled.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall"]

LIB = PKG / "libisti.so"
SCENE_LIB = PKG / "libisti_scene.so"
IO_LIB = PKG / "libisti_io.so"
SPLINTER = PKG / "splinter"


def _run(cmd, verbose):
    if verbose:
        print(" ".join(str(c) for c in cmd), flush=True)
    res = subprocess.run([str(c) for c in cmd], capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(str(c) for c in cmd)}")
    return res


def build_native(verbose: bool = False) -> None:
    """Compile libisti.so (HIP), libisti_scene.so and the splinter driver."""
    obj_dir = PKG / "_build"
    obj_dir.mkdir(exist_ok=True)
    cmds, objs = [], []
    for src in ("resolve_kernels.hip", "isti_kernels.hip", "unet.hip", "quality.hip", "isti_noalloc.hip",
                "codec.hip", "isti_api.cpp"):
        obj = obj_dir / (src.rsplit(".", 1)[0] + ".o")
        extra = ["-x", "hip"] if src.endswith(".hip") else []
        cmds.append([HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", *extra, "-c", CSRC / src, "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1)) as pool:
        for f in [pool.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIB)

    tmp = SCENE_LIB.with_suffix(".so.tmp")
    _run([CXX, *HOST_FLAGS, "-shared", f"-I{INCLUDE}", CSRC / "scene.cpp", "-o", tmp], verbose)
    os.replace(tmp, SCENE_LIB)

    tmp = IO_LIB.with_suffix(".so.tmp")
    _run([CXX, *HOST_FLAGS, "-shared", f"-I{INCLUDE}", CSRC / "keyframe_io.cpp", "-lz", "-o", tmp], verbose)
    os.replace(tmp, IO_LIB)

    tmp = SPLINTER.with_name("splinter.tmp")
    _run([HIPCC, "-O2", "-std=c++17", "-fopenmp", f"-I{INCLUDE}", CSRC / "splinter.cpp",
          f"-L{PKG}", "-listi", "-listi_scene", "-listi_io", f"-Wl,-rpath,$ORIGIN", "-o", tmp], verbose)
    os.replace(tmp, SPLINTER)


def build_oracle(verbose: bool = False) -> None:
    """Compile the CPU oracle (test infrastructure) with its own Makefile."""
    _run(["make", "-B", "-C", ROOT / "oracle"], verbose)


def build(verbose: bool = False) -> None:
    build_native(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build(verbose=True)
