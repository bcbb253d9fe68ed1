"""libhrt.so loads and exports every entry point include/hrt.h declares (CPU only)."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

import heavyrt_amd as H

HEADER = os.path.join(ROOT, "include", "hrt.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hrt_[a-z0-9_]+)\s*\(", text)))


def test_header_lists_match_python_binding():
    assert declared_functions() == sorted(H.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(H.LIB_PATH):
        pytest.fail("libhrt.so not built (python -m heavyrt_amd.build)")
    out = subprocess.run(["nm", "-D", "--defined-only", H.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (hrt_\w+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing


def test_library_loads_and_reports_errors_without_a_gpu():
    L = H.lib()
    for name in declared_functions():
        assert hasattr(L, name)
    # null out-pointer is rejected before any device call
    assert L.hrt_init(0, None) == 1
    assert b"null" in L.hrt_last_error()
    assert L.hrt_scene_create(None, None) == 1
    assert L.hrt_geom_count(None) == 0


def test_scene_adapter_exports_factory():
    """libhrt_scene.so provides the reference interface's factory (CrossRT.h:195-196)."""
    lib = os.path.join(os.path.dirname(H.LIB_PATH), "libhrt_scene.so")
    if not os.path.exists(lib):
        pytest.fail("libhrt_scene.so not built (python -m heavyrt_amd.build)")
    out = subprocess.run(["nm", "-DC", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    assert "CreateSceneRT(char const*, char const*, char const*)" in out
    assert "DeleteSceneRT(ISceneObject*)" in out
    assert "HipSceneRT_NearestHitBatch" in out


def test_crossrt_header_compiles(tmp_path):
    src = ('#include "heavyrt/CrossRT.h"\n#include "heavyrt/HipSceneRT.h"\n'
           'static_assert(sizeof(CRT_Hit) == 32, "CRT_Hit layout");\n'
           'int main(){ LiteMath::float4x4 m; LiteMath::float4 v(1,2,3,1); LiteMath::float4 r = m * v;'
           ' return r.x == 1.0f ? 0 : 1; }\n')
    p = tmp_path / "x.cpp"
    p.write_text(src)
    exe = str(tmp_path / "x")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(p), "-o", exe],
                   check=True)
    assert subprocess.run([exe]).returncode == 0


def test_header_compiles_as_c(tmp_path):
    src = '#include "hrt.h"\nint main(void){ hrt_hit h; (void)h; return (int)sizeof(hrt_hit) - 32; }\n'
    tmp = str(tmp_path / "hdr_check.c")
    with open(tmp, "w") as fh:
        fh.write(src)
    exe = tmp[:-2]
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), tmp, "-o", exe],
                   check=True)
    assert subprocess.run([exe]).returncode == 0
