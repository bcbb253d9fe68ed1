"""Build recipe for the engine's native libraries (in-tree, gfx950 only).

    python -m heavyrt_amd.build          # libhrt.so (HIP kernels + C-ABI)
                                         # libhrt_scene.so (C++ ISceneObject adapter)

The HIP sources are compiled with -ffp-contract=off (no fused multiply-add:
the reference's arithmetic rounds after every operation) and correctly
rounded fp32 division, without fast-math; packed fp32 instructions are
disabled for the device code.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
INCLUDE = os.path.join(ROOT, "include")

LIB_HRT = os.path.join(HERE, "libhrt.so")
LIB_SCENE = os.path.join(HERE, "libhrt_scene.so")

HIP_SOURCES = ["sort.hip", "lbvh.hip", "ploc.hip", "trace.hip", "trace4.hip", "trace_tlas.hip", "batch.hip", "serve.hip",
               "rtao.hip", "eye.hip", "capi.hip", "noalloc.hip", "probe.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hrt(verbose: bool = True) -> str:
    objs = []
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-I", INCLUDE, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB_HRT + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB_HRT)
    return LIB_HRT


def build_scene_adapter() -> str:
    """C++ ISceneObject adapter (host compiler only) over libhrt.so."""
    src = os.path.join(HOST, "HipSceneRT.cpp")
    tmp = LIB_SCENE + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", INCLUDE, "-o", tmp, src,
          "-L", HERE, "-lhrt", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, LIB_SCENE)
    return LIB_SCENE


HARNESS_SRC = os.path.join(ROOT, "tests", "cpp", "scene_harness.cpp")
HARNESS_BIN = os.path.join(ROOT, "tests", "cpp", "_build", "scene_harness")


def build_harness() -> str:
    """Renderer-style C++ caller of ISceneObject used by the GPU tests."""
    os.makedirs(os.path.dirname(HARNESS_BIN), exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", "-fopenmp", "-Wall", "-I", INCLUDE, "-o", HARNESS_BIN, HARNESS_SRC,
          "-L", HERE, "-lhrt_scene", "-lhrt", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/../../../heavyrt_amd"])
    return HARNESS_BIN


def build_all() -> None:
    build_hrt()
    build_scene_adapter()
    build_harness()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
