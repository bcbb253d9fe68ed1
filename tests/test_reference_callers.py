"""The reference's own renderers compile unchanged against this build's headers.

01_eye_rays and 02_rtao (examples/01_eye_rays/eye_ray*.cpp,
examples/02_rtao/rtao*.cpp) and the scene loaders they include
(examples/loader/{cmesh,hydraxml,gltf_loader}.cpp) are the callers of
ISceneObject (core/CrossRT.h:64-196).  Each is compiled to an object file with
include/heavyrt first on the include path, so CrossRT.h and LiteMath.h come
from this build and every other header from the reference tree where it lies.
Nothing is copied; the objects go to a temporary directory.  Skipped where the
reference tree is absent.
"""
import os
import subprocess

import pytest

from conftest import ROOT

REF = "/root/reference"
SOURCES = [
    "examples/01_eye_rays/eye_ray.cpp",
    "examples/01_eye_rays/eye_ray_host.cpp",
    "examples/02_rtao/rtao.cpp",
    "examples/02_rtao/rtao_host.cpp",
    "examples/loader/cmesh.cpp",
    "examples/loader/hydraxml.cpp",
    "examples/loader/gltf_loader.cpp",
]


def include_flags():
    dirs = [os.path.join(ROOT, "include", "heavyrt")]
    dirs += [os.path.join(REF, d) for d in ("examples", "examples/loader", "core/common", "external/tiny_gltf",
                                            "examples/01_eye_rays", "examples/02_rtao")]
    return [f"-I{d}" for d in dirs]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "examples")), reason="reference tree not present")
@pytest.mark.parametrize("src", SOURCES)
def test_reference_renderer_compiles_unchanged(src, tmp_path):
    path = os.path.join(REF, src)
    assert os.path.exists(path), path
    obj = tmp_path / (os.path.basename(src) + ".o")
    r = subprocess.run(["g++", "-std=c++17", "-O0", "-c", *include_flags(), path, "-o", str(obj)],
                       capture_output=True, text=True)
    errors = [line for line in r.stderr.splitlines() if "error" in line]
    assert r.returncode == 0, "\n".join(errors[:20]) or r.stderr[-2000:]
    assert obj.stat().st_size > 0
