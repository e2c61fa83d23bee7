"""Gradient paint on the CPU: the texture builder and the oracle's gradient compositing.

The builder (scanline/gradient.py) restates RVG::load's table and ramp construction
(rvg.cpp:490-758); the oracle restates the output shaders' gradient evaluation
(output_32.vert.glsl:55-97, output_32.frag.glsl:45-51).  Both are checked against closed
forms worked out independently of their code:
- a linear gradient from (x0, y0) to (x1, y0) has ramp parameter (x - x0) / (x1 - x0), and a
  radial one centred at its focus has parameter |p - centre| / r;
- the ramp of two stops holds trunc(255 (c0 + (c1 - c0) i / (n - 1))), and its length is the
  next power of two above 16 / (smallest stop gap);
- a rendered opaque gradient rectangle shows, at pixel (x, y), the ramp at the parameter of
  the point (x + 1.5, y + 0.5) in coverage output and (x + 0.5, y + 0.5) in multisample
  output (the shaders' line ends sit at x + 1 and x + 0), within 1.5 codes (ramp texels are
  truncated to 8 bits, the framebuffer rounds).
The GL sampler's own fixed-point filtering is not restated (parity with the reference's
GPU sampler is unpinned); the renderer matches this restatement bit for bit
(tests/test_gpu_gradient.py).
"""
import os

import numpy as np
import pytest

import pyoracle
from scanline import VGContainer, rgba, rvg
from scanline.gradient import GradientTextures, lrgb_to_srgb, srgb_to_lrgb
from scanline.scene import FT_GRADIENT

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "gradient.rvg")
f32 = np.float32


def words(table, slot):
    t = np.asarray(table, dtype=np.int64).astype(np.uint32)[12 * slot:12 * slot + 12]
    return t.view(np.float32), t.view(np.int32)


def test_linear_table_closed_form():
    g = GradientTextures(srgb=False)
    g.add("linear", 10, 20, 50, 20, 0, [(0, 0, 0, 0, 1), (1, 1, 1, 1, 1)], 1.0, [1, 0, 0, 0, 1, 0])
    ramp, table = g.textures()
    w, wi = words(table, 0)
    # g.x = (x - 10) / 40 * 2 - 1, g.y = 0, focal (-1, 0, 0): c = (g.x + 1) / 2
    np.testing.assert_allclose(w[:6], [0.05, 0, -1.5, 0, 0, 0], rtol=1e-6, atol=0)
    np.testing.assert_array_equal(w[8:11], f32([-1, 0, 0]))
    assert wi[6] == 0 and w[7] == f32(31) / f32(1024) and wi[11] == 0
    for x in (10.0, 20.0, 37.5, 50.0):
        gx = f32(w[0] * f32(x)) + w[2]
        d = gx + 1
        assert abs(float(d * d / (2 * d) if d else 0) - (x - 10) / 40) < 1e-6


def test_radial_table_closed_form():
    g = GradientTextures(srgb=False)
    # paint transform: scale 2 then move by (6, -4); centre (3, 5) = focus, radius 4
    g.add("radial", 3, 5, 3, 5, 4, [(0, 1, 1, 1, 1), (1, 0, 0, 0, 1)], 1.0, [2, 0, 6, 0, 2, -4])
    _, table = g.textures()
    w, _ = words(table, 0)
    # gradient space: ((x - 6) / 2 - 3) / 4 and ((y + 4) / 2 - 5) / 4
    np.testing.assert_allclose(w[:6], [0.125, 0, -1.5, 0, 0.125, -0.75], rtol=0, atol=1e-7)
    np.testing.assert_array_equal(w[8:11], f32([0, 0, 1]))
    # ramp parameter at a point: its distance from the transformed centre (12, 6) over 8
    for x, y in ((12, 6), (20, 6), (12, 14), (17, 12)):
        gx = float(w[0]) * x + float(w[1]) * y + float(w[2])
        gy = float(w[3]) * x + float(w[4]) * y + float(w[5])
        assert abs(np.hypot(gx, gy) - np.hypot(x - 12, y - 6) / 8) < 1e-6


def test_two_stop_ramp_and_resolution():
    g = GradientTextures(srgb=False)
    g.add("linear", 0, 0, 1, 0, 0, [(0, 1, 0, 0.2, 1), (1, 0, 0.5, 1, 0.5)], 0.8, [1, 0, 0, 0, 1, 0])
    ramp, _ = g.textures()
    r = ramp.astype(np.uint32)[:32]
    assert ramp.size == 1024 and not ramp[32:].any()
    t = np.arange(32) / 31.0
    for ch, (a, b) in enumerate([(1, 0), (0, 0.5), (0.2, 1), (0.8, 0.4)]):
        got = (r >> (8 * ch)) & 0xFF
        exact = 255 * (a + (b - a) * t)
        assert np.all((got <= exact + 1e-3) & (got >= exact - 1 - 1e-3)), ch   # trunc of the float value
        assert got[0] == int(255 * a + 1e-6) and got[-1] == int(255 * b + 1e-6)
    # the resolution: next power of two above 16 / smallest gap
    for stops, n in (([0, 1], 32), ([0, 0.3, 1], 64), ([0, 0.1, 1], 256), ([0, 0.01, 1], 1024), ([0.5], 32)):
        g = GradientTextures(srgb=False)
        g.add("linear", 0, 0, 1, 0, 0, [(s, 0, 0, 0, 1) for s in stops], 1.0, [1, 0, 0, 0, 1, 0])
        _, table = g.textures()
        assert words(table, 0)[0][7] == f32(n - 1) / f32(1024), stops


def test_srgb_ramp_interpolates_in_linear_light():
    g = GradientTextures(srgb=True)
    g.add("linear", 0, 0, 1, 0, 0, [(0, 0, 0, 0, 1), (1, 1, 1, 1, 1)], 1.0, [1, 0, 0, 0, 1, 0])
    ramp, _ = g.textures()
    r = ramp.astype(np.uint32)[:32] & 0xFF
    t = np.arange(32) / 31.0
    exact = 255 * np.where(t <= 0.0031308, 12.92 * t, 1.055 * t ** (1 / 2.4) - 0.055)
    assert np.all(np.abs(r - exact) <= 1.0)
    assert float(srgb_to_lrgb(lrgb_to_srgb(0.25))) == pytest.approx(0.25, abs=1e-6)


def test_packing_by_decreasing_length():
    g = GradientTextures(srgb=False)
    g.add("linear", 0, 0, 1, 0, 0, [(0, 0, 0, 0, 1), (1, 1, 1, 1, 1)], 1.0, [1, 0, 0, 0, 1, 0])          # n = 32
    g.add_solid()
    g.add("linear", 0, 0, 1, 0, 0, [(0, 0, 0, 0, 1), (0.3, 1, 1, 1, 1), (1, 0, 0, 0, 1)], 1.0, [1, 0, 0, 0, 1, 0])  # 64
    ramp, table = g.textures()
    (w0, i0), (w1, i1), (w2, i2) = words(table, 0), words(table, 1), words(table, 2)
    assert (i2[6], i0[6], i1[6]) == (0, 64, 96)
    assert w1[7] == f32(-1) / f32(1024) and not i1[:6].any()
    r = ramp.astype(np.uint32)
    assert (r[64] & 0xFF) == 0 and (r[64 + 31] & 0xFF) == 255       # the 32-texel ramp after the 64-texel one
    assert (r[0] & 0xFF) == 0 and (r[63] & 0xFF) == 0


def test_rvg_gradient_fixture():
    vg, _, _ = rvg.load(FIXTURE, srgb=False)
    a = vg.arrays()
    assert a.n_paths == 3
    assert vg._ftype[0] == FT_GRADIENT and vg._ftype[2] == FT_GRADIENT
    assert list(a.fill_color) == [1, rgba(0, 153, 0, int(f32(255) * f32(0.7))), 3]
    assert a.gradient_table.size == 36 and a.gradient_ramp.size == 1024
    w0, i0 = words(a.gradient_table, 0)
    np.testing.assert_allclose(w0[:6], [0.025, 0, -1.2, 0, 0, 0], rtol=1e-6, atol=0)
    assert i0[6] == 64 and w0[7] == f32(31) / f32(1024)
    w2, i2 = words(a.gradient_table, 2)
    assert i2[6] == 0 and w2[7] == f32(63) / f32(1024)
    np.testing.assert_allclose(w2[8:10], [4 / 14, -2 / 14], rtol=1e-6)
    # the radial ramp's last stop alpha is 0.5 * 0.5 = 0.25
    assert (int(a.gradient_ramp.astype(np.uint32)[63]) >> 24) == int(f32(0.25) * 255)


def _gradient_rect(x0, x1, kind="linear", srgb=False, transform=None, rect=(8, 4, 88, 28)):
    vg = VGContainer()
    vg.addPath(0, rgba(255, 0, 255, 0), FT_GRADIENT)
    X0, Y0, X1, Y1 = rect
    vg.addPolygon([(X0, Y0), (X1, Y0), (X1, Y1), (X0, Y1)])
    g = GradientTextures(srgb=srgb)
    if kind == "linear":
        g.add("linear", x0, 0, x1, 0, 0, [(0, 0, 0, 0, 1), (1, 1, 1, 1, 1)], 1.0, [1, 0, 0, 0, 1, 0])
    else:
        g.add("radial", x0, x1, x0, x1, 40, [(0, 1, 1, 1, 1), (1, 0, 0, 0, 1)], 1.0, [1, 0, 0, 0, 1, 0])
    vg.gradient_ramp, vg.gradient_table = (np.asarray(v, dtype=np.int64).tolist() for v in g.textures())
    return vg


@pytest.mark.parametrize("spp", [8, 32])
@pytest.mark.parametrize("ms", [False, True])
def test_oracle_linear_gradient_closed_form(spp, ms):
    vg = _gradient_rect(8, 88)
    img = pyoracle.run(vg.arrays(), 96, 32, spp, srgb=False, ms=ms).get("image").reshape(32, 96, 4)
    off = 0.5 if ms else 1.5
    ys, xs = np.mgrid[5:27, 9:87]
    c = np.clip((xs + off - 8) / 80.0, 0, 1)
    got = img[ys, xs, 0].astype(np.float64)
    assert np.abs(got - 255 * c).max() <= 1.5
    assert (img[ys, xs, 3] == 255).all() and (img[ys, xs, 0] == img[ys, xs, 2]).all()
    # outside the rectangle: white
    assert (img[30, :, :3] == 255).all()


def test_oracle_radial_gradient_closed_form():
    vg = _gradient_rect(48, 16, kind="radial", rect=(4, 0, 92, 32))
    img = pyoracle.run(vg.arrays(), 96, 32, 16, srgb=False).get("image").reshape(32, 96, 4)
    ys, xs = np.mgrid[1:31, 5:91]
    d = np.hypot(xs + 1.5 - 48, ys + 0.5 - 16)
    v = 255 * (1 - np.clip(d / 40, 0, 1))
    assert np.abs(img[ys, xs, 0] - v).max() <= 1.5


def test_oracle_gradient_follows_the_transform():
    """Scaled by 2 about the origin, the scene shows object point ((x + 1.5) / 2, ...) at
    window pixel x: the gradient moves with its path."""
    vg = _gradient_rect(4, 44, rect=(4, 2, 44, 14))
    m = np.eye(4, dtype=np.float32)
    m[0, 0] = m[1, 1] = 2.0
    img = pyoracle.run(vg.arrays(), 96, 32, 32, srgb=False, transform=m).get("image").reshape(32, 96, 4)
    ys, xs = np.mgrid[5:27, 9:87]
    c = np.clip(((xs + 1.5) / 2 - 4) / 40.0, 0, 1)
    assert np.abs(img[ys, xs, 0] - 255 * c).max() <= 1.5


def test_oracle_srgb_gradient():
    """sRGB correction: the ramp holds sRGB codes of a linear-light blend, texels are decoded
    before filtering and the framebuffer encodes again, so the image shows the sRGB code of
    the linear parameter."""
    vg = _gradient_rect(8, 88, srgb=True)
    img = pyoracle.run(vg.arrays(), 96, 32, 32, srgb=True).get("image").reshape(32, 96, 4)
    ys, xs = np.mgrid[5:27, 9:87]
    c = np.clip((xs + 1.5 - 8) / 80.0, 0, 1)
    enc = 255 * np.where(c <= 0.0031308, 12.92 * c, 1.055 * c ** (1 / 2.4) - 0.055)
    assert np.abs(img[ys, xs, 0] - enc).max() <= 2.0


def test_oracle_renders_the_fixture_with_its_paints():
    vg, _, _ = rvg.load(FIXTURE, srgb=False)
    img = pyoracle.run(vg.arrays(), 96, 64, 32, srgb=False).get("image").reshape(64, 96, 4)
    # left end of the linear gradient is red, right end blue (points clear of the green square)
    assert tuple(img[10, 9, :3])[0] > 240 and img[10, 9, 2] < 15
    assert img[10, 86, 0] < 15 and img[10, 86, 2] > 240
    # the radial paint is drawn at half opacity over white: no pixel of it is white or black
    patch = img[44:52, 44:52, :3].astype(int)
    assert (patch < 250).any() and (patch > 100).all()
    # a word naming a slot whose ramp scale is not positive paints nothing
    a = vg.arrays()
    a.fill_color[0] = 2  # the solid path's empty slot
    blank = pyoracle.run(a, 96, 64, 32, srgb=False).get("image").reshape(64, 96, 4)
    assert (blank[10, 9:30, :3] == 255).all()


def test_svg_linear_gradient_closed_form():
    """An SVG linear gradient from x = 14 to 94 in user space; the viewBox origin (4, 2)
    and the y flip move it to x = 10 .. 90 in the scene, where pixel (x, y) shows
    ((x + 1.5) - 10) / 80 of the way from red to blue."""
    from scanline import svg
    vg = svg.load(os.path.join(HERE, "golden", "gradient.svg"), srgb=False)
    a = vg.arrays()
    assert list(a.fill_color) == [1, 2] and a.gradient_table.size == 24
    img = pyoracle.run(a, 96, 64, 32, srgb=False).get("image").reshape(64, 96, 4)
    ys, xs = np.mgrid[37:59, 11:89]       # rows y = 66 - 30 .. 66 - 6, inside the rectangle
    c = np.clip((xs + 1.5 - 10) / 80.0, 0, 1)
    assert np.abs(img[ys, xs, 0] - 255 * (1 - c)).max() <= 1.5
    assert np.abs(img[ys, xs, 2] - 255 * c).max() <= 1.5
