"""RVG scene loading (the reference's rvg.cpp semantics), on the CPU.

The reference ships no .rvg scenes, so tests/golden/sample.rvg is a hand-written file that
exercises every command of the loader; the expected SoA below follows rvg.cpp line by line
(cited in scanline/rvg.py), and the oracle renders it as a sanity check.
"""
import os

import numpy as np
import pytest

import pyoracle
import scanline
from scanline import rvg

HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLE = os.path.join(HERE, "golden", "sample.rvg")
f32 = np.float32


def test_sample_structure():
    vg, viewport, window = rvg.load(SAMPLE)
    a = vg.arrays()
    assert viewport == ((0, 0), (64, 48)) and window == ((0, 0), (64, 48))
    # the skipped element leaves one empty path behind it (a new path starts only after an
    # element that added vertices): 4 paths, the last without curves
    assert a.n_paths == 4
    assert list(a.fill_rule) == [0, 1, 0, 0]
    types = list(a.curve_type)
    L, Q, C, R = (scanline.CURVE_LINEAR, scanline.CURVE_QUADRATIC, scanline.CURVE_CUBIC,
                  scanline.CURVE_RATIONAL)
    # square: 4 lines (Z marks the contour closed but adds no line); element 2: Q, C + closing line; element 3: two
    # half-circle arcs (A with w = 0), a rational arc, one line, already closed
    assert types == [L, L, L, L, Q, C, L, R, R, R, L]
    assert list(a.curve_path_id) == [0] * 4 + [1] * 3 + [2] * 4
    v = a.vertex
    # closing line of element 2: from the cubic's end back to its M point
    c = 6
    p = a.curve_vertex_pos[c]
    np.testing.assert_array_equal(v[p:p + 2], [[34, 36], [30, 6]])
    # element 3: the half circle from (8,36) to (20,36) bulging by (0,-6), then shifted by (2, 1.5)
    p = a.curve_vertex_pos[7]
    mid = f32(0.5) * (f32(8) + f32(20)) + f32(0), f32(0.5) * (f32(36) + f32(36)) + f32(-6)
    d = f32(0.5) * (f32(20) - f32(8))
    np.testing.assert_array_equal(v[p:p + 3], [[8 + 2, 36 + 1.5], [mid[0] - d + 2, mid[1] + 1.5],
                                               [mid[0] + 2, mid[1] + 1.5]])
    assert np.isclose(a.curve_arc_w[7], 0.70710678) and np.isclose(a.curve_arc_w[9], 0.5)
    p = a.curve_vertex_pos[9]
    np.testing.assert_array_equal(v[p + 1], [26 / 0.5 + 2, 30 / 0.5 + 1.5])
    # colours: truncated to 8 bits, alpha scaled by the opacity; rgb(...) keeps only
    # len - 6 characters after "rgb(", so "0.8" loses its last digit and blue is 0
    col = a.fill_color.astype(np.uint32)
    assert col[0] == scanline.rgba(255, 0, 0, 255)
    assert col[1] == scanline.rgba(0, 127, 255, 127)
    assert col[2] == scanline.rgba(51, 102, 0, 255)


def test_a128_and_gradients():
    text = open(SAMPLE).read()
    vg, _, _ = rvg.loads(text, a128=True)
    col = vg.arrays().fill_color.astype(np.uint32)
    assert (col[1] >> 24) == int(np.trunc(f32(int(np.trunc(f32(127) / f32(255) * f32(127)))) / f32(127) * f32(255)))
    grad = text.replace("solid rgba(0,0.5,1,1)", "linear 0,0 1,1 0,rgba(0,0,0,1) 1,rgba(1,1,1,1) pad 1,0,0,0,1,0")
    with pytest.raises(rvg.RVGError):
        rvg.loads(grad, gradients="raise")
    vg, _, _ = rvg.loads(grad, gradients="skip")
    assert vg.arrays().fill_color[1] == 0
    # rendered: the fill word names the path's table slot (path index + 1)
    vg, _, _ = rvg.loads(grad)
    assert vg.arrays().fill_color[1] == 2
    # a gradient line without its transform numbers is refused
    with pytest.raises(rvg.RVGError):
        rvg.loads(text.replace("solid rgba(0,0.5,1,1)", "linear 0,0 1,1 pad"))


def test_unsupported_command_raises():
    text = open(SAMPLE).read().replace("L 28,4", "l 28,4")
    with pytest.raises(rvg.RVGError):
        rvg.loads(text)


def test_oracle_renders_sample():
    vg, _, _ = rvg.load(SAMPLE)
    ref = pyoracle.run(vg.arrays(), 64, 48, 8, srgb=False)
    img = ref.get("image").reshape(48, 64, 4)
    # inside the opaque red square (y up: row index = y)
    assert tuple(img[16, 16]) == (255, 0, 0, 255)
    # outside every path: white
    assert tuple(img[46, 2]) == (255, 255, 255, 255)
