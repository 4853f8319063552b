"""The C ABI boundary, checked without a GPU: librtamd.so loads, exports every
entry point include/rt_api.h declares, and the ctypes mirrors have the C
layouts (sizes and offsets from the C compiler itself)."""
import os
import re
import subprocess
import tempfile

import ctypes as C
import pytest

from rustracer_amd import _ffi as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rt_api.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rt_[a-z0-9_]+)\s*\(", text)))


def test_header_functions_are_all_in_the_ctypes_table():
    assert set(declared_functions()) == set(F.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(F.LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_api_version_without_device():
    assert F.lib().rt_api_version() == F.API_VERSION == 3


STRUCTS = ["rt_xform", "rt_medium", "rt_object", "rt_material", "rt_background", "rt_scene_desc",
           "rt_ray", "rt_hit", "rt_camera", "rt_sample_params", "rt_photon_params", "rt_wavefront_opts", "rt_batch_view",
           "rt_param", "rt_image", "rt_accum_opts", "rt_pixel_state", "rt_resolve_opts"]


def test_struct_layouts_match_c_compiler():
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for s in STRUCTS:
        src.append(f'printf("{s} %zu\\n", sizeof({s}));')
        for name, _ in getattr(F, s)._fields_:
            src.append(f'printf("{s}.{name} %zu\\n", offsetof({s}, {name}));')
    src.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(c, "w").write("\n".join(src))
        subprocess.run(["gcc", "-o", exe, c], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    want = dict(line.split() for line in out if line.strip())
    for s in STRUCTS:
        cls = getattr(F, s)
        assert C.sizeof(cls) == int(want[s]), s
        for name, _ in cls._fields_:
            assert getattr(cls, name).offset == int(want[f"{s}.{name}"]), f"{s}.{name}"


def test_numpy_ray_and_hit_dtypes_match_structs():
    assert F.ray_dtype().itemsize == C.sizeof(F.rt_ray)
    assert F.hit_dtype().itemsize == C.sizeof(F.rt_hit)


def test_scene_descriptor_flattens_every_shape_kind():
    from tests.helpers import mixed_scene
    sc = mixed_scene()
    d = sc.descriptor()
    assert d.num_objects == len(sc.objects)
    kinds = {d.objects[i].shape for i in range(d.num_objects)}
    assert {F.SHAPE_SPHERE, F.SHAPE_QUAD, F.SHAPE_BOX, F.SHAPE_MESH, F.SHAPE_CONVEX} <= kinds
    assert d.num_tris > 0 and d.pool_len % 4 == 0 and d.num_xforms >= 5


def test_renderer_fails_loudly_without_library(monkeypatch):
    monkeypatch.setattr(F, "LIB_PATH", "/nonexistent/librtamd.so")
    monkeypatch.setattr(F, "_lib", None)
    from rustracer_amd import renderer
    with pytest.raises(ImportError):
        renderer.Renderer(0)
