"""The C-ABI library loads and exports every function include/*.h declares;
argument validation rejects bad input without touching a GPU."""
import ctypes as C
import re
from pathlib import Path

import pytest

import smallproject_amd as sp
from smallproject_amd import _lib

INCLUDE = Path(__file__).resolve().parent.parent / "include"


def declared_functions():
    names = set()
    for h in INCLUDE.glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(vrw?_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_api():
    names = declared_functions()
    assert "vr_render_frame" in names and "vrw_render" in names and len(names) > 40


def test_every_declared_function_is_exported():
    L = sp.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_python_binding_covers_the_api():
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_struct_layouts():
    import oracle as O
    # 2+4*3 floats, 4 u32, 4*4 floats, 4 i32, 5 u32
    assert C.sizeof(_lib.VrRenderParams) == C.sizeof(O.Params) == 172
    assert C.sizeof(_lib.VrConfig) == 8 * 4 + 256 * 4
    assert C.sizeof(_lib.VrLight) == 12


def test_default_config_is_the_reference():
    cfg = sp.default_config(1600, 1024)
    assert (cfg.width, cfg.height) == (1600, 1024)
    assert cfg.quirks == _lib.QUIRKS_REFERENCE and cfg.gfxexp == 1 and cfg.light_mode == _lib.LIGHTS_AREA
    assert cfg.brick_capacity == 1 << 20
    assert [cfg.emit_lut[i] for i in (0, 1, 255)] == [0.0, 1.0, 255.0]


def test_null_and_invalid_arguments_are_rejected():
    L = sp.lib()
    assert L.vr_create(None, None) == _lib.VR_ERR_ARG
    bad = sp.default_config(0, 10)
    h = C.c_void_p()
    assert L.vr_create(C.byref(bad), C.byref(h)) == _lib.VR_ERR_ARG
    bad = sp.default_config(10, 10)
    bad.emit_lut[0] = 1.0  # voxels without emission must stay dark
    assert L.vr_create(C.byref(bad), C.byref(h)) == _lib.VR_ERR_ARG
    assert L.vr_render_frame(None, None) == _lib.VR_ERR_ARG
    assert L.vr_destroy(None) == _lib.VR_ERR_ARG
    assert L.vrw_create(None, None) == _lib.VR_ERR_ARG


def test_log_lut():
    lut = sp.log_lut(0.1, 10.0)
    assert lut[0] == 0.0
    assert lut[1] == pytest.approx(0.1, rel=1e-6) and lut[255] == pytest.approx(10.0, rel=1e-6)
    assert all(lut[i] < lut[i + 1] for i in range(1, 255))
