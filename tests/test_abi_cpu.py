"""The C ABI library on the host (no GPU needed).

Checks that libscanline.so loads and exports every entry point include/scanline.h
declares, that the Python binding declares the same set, that argument errors come back
as status codes, that the renderer refuses to run without a GPU (there is no CPU
fallback), and that the seeded scene generators are deterministic and follow their
specifications.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import scanline
from scanline import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "scanline.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sl_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = header_functions()
    assert "sl_create" in names and "sl_rasterize" in names and len(names) >= 20


def test_library_exports_every_header_symbol():
    L = _abi.lib()
    missing = [n for n in header_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert sorted(_abi.EXPORTS) == header_functions()


def test_version_string():
    v = _abi.lib().sl_version().decode()
    assert v and "scanline" in v.lower()


def test_null_arguments_are_status_codes():
    L = _abi.lib()
    assert L.sl_create(0, None) == _abi.SL_E_ARG
    assert L.sl_destroy(None) == _abi.SL_OK  # like deleting a null pointer
    assert L.sl_set_samples(None, 8) == _abi.SL_E_ARG
    assert L.sl_rasterize(None, None) == _abi.SL_E_ARG
    assert L.sl_add_vg(None, None) == _abi.SL_E_ARG
    assert L.sl_last_error(None).decode()
    assert L.sl_step_name(99) == b""


def test_no_cpu_fallback():
    """Without a GPU, creating a renderer fails with SL_E_NODEVICE."""
    L = _abi.lib()
    h = ctypes.c_void_p()
    rc = L.sl_create(0, ctypes.byref(h))
    if rc == _abi.SL_OK:
        L.sl_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == _abi.SL_E_NODEVICE
    with pytest.raises(scanline.ScanlineError):
        scanline.VGRasterizer(0)


def test_generator_is_deterministic():
    a = scanline.generate(scanline.scene.KIND_MIXED, 256, 256, 50, 1234)
    b = scanline.generate(scanline.scene.KIND_MIXED, 256, 256, 50, 1234)
    c = scanline.generate(scanline.scene.KIND_MIXED, 256, 256, 50, 1235)
    for f in ("vertex", "curve_vertex_pos", "curve_type", "curve_path_id", "curve_arc_w", "fill_rule",
              "fill_color"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f))
    assert a.vertex.shape != c.vertex.shape or not np.array_equal(a.vertex, c.vertex)


def test_generator_rejects_bad_arguments():
    L = _abi.lib()
    obj = ctypes.c_void_p()
    assert L.sl_scene_generate(9, 64, 64, 10, 1, 0, 0, ctypes.byref(obj)) == _abi.SL_E_ARG
    assert L.sl_scene_generate(1, 0, 64, 10, 1, 0, 0, ctypes.byref(obj)) == _abi.SL_E_ARG


def test_cubic_chain_scene_follows_its_specification():
    """512x512, 200 closed chains of 4-12 cubics, control points in the canvas, nonzero,
    opaque colours."""
    s, W, H, spp = scanline.config_scene(1)
    assert (W, H, spp) == (512, 512, 8)
    assert s.n_paths == 200
    assert np.all(s.curve_type == scanline.CURVE_CUBIC)
    per_path = np.bincount(s.curve_path_id, minlength=200)
    assert per_path.min() >= 4 and per_path.max() <= 12
    v = s.vertex
    assert v.min() >= 0 and v.max() < 512
    assert np.all(s.fill_rule == scanline.FILL_NONZERO)
    assert np.all((s.fill_color >> 24) == 255)
    # closed chains: each cubic ends where the next one of its path starts
    for p in range(5):
        cs = np.flatnonzero(s.curve_path_id == p)
        starts = v[s.curve_vertex_pos[cs]]
        ends = v[s.curve_vertex_pos[cs] + 3]
        np.testing.assert_array_equal(ends, np.roll(starts, -1, axis=0))


def test_mixed_scene_follows_its_specification():
    s = scanline.generate(scanline.scene.KIND_MIXED, 2048, 2048, 2000, scanline.scene.SEED_BASE ^ 2)
    frac = np.bincount(s.curve_type, minlength=5)[2:5] / len(s.curve_type)
    assert frac[0] > 0.3  # lines: 40 % of segments plus the star polygons
    assert 0.2 < frac[1] < 0.4 and 0.2 < frac[2] < 0.4
    rules = s.fill_rule.astype(int)
    assert abs(rules.mean() - 0.5) < 0.01  # alternating nonzero / even-odd
    a = (s.fill_color >> 24).astype(int)
    assert a.min() >= round(0.2 * 255) - 1 and a.max() <= 255
