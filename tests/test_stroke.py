"""Stroke-to-fill on the CPU (scanline/stroke.py restating vg_container.cpp:555-1353 and
bezier_curve.cpp:186-600).

Checked against what the geometry fixes, not against the conversion's own code:
- every approximated offset curve stays within a small distance of the true offset: dense
  samples of each piece lie at distance h from the base curve (float64 distance to dense
  base samples), and consecutive pieces share their end points;
- every stroke outline is one closed contour (each curve starts where the previous ends);
- rendered by the oracle, a stroke of a horizontal line on integer coordinates is a
  pixel-exact rectangle, a stroked circle covers the annulus area pi((r+h)^2 - (r-h)^2),
  and a bevel-joined corner covers the two bands plus the bevel triangle.
"""
import os

import numpy as np
import pytest

import pyoracle
from scanline import VGContainer, rgba, svg
from scanline import stroke as S

HERE = os.path.dirname(os.path.abspath(__file__))


def dist_to_curve(c, pts, n=4000):
    t = np.linspace(0, 1, n)
    base = np.array([c.point(x) for x in t])
    d = np.sqrt(((pts[:, None, :] - base[None, :, :]) ** 2).sum(-1))
    return d.min(axis=1)


def min_curvature_radius(c):
    t = np.linspace(0, 1, 4001)
    d1 = np.array([c.deriv(x) for x in t])
    d2 = np.gradient(d1, t, axis=0)
    k = np.abs(d1[:, 0] * d2[:, 1] - d1[:, 1] * d2[:, 0]) / np.maximum(np.hypot(d1[:, 0], d1[:, 1]) ** 3, 1e-30)
    return 1.0 / k.max()


@pytest.mark.parametrize("ct", [S.QUAD, S.CUBIC, S.RATIONAL])
def test_offset_pieces_lie_at_the_offset_distance(ct):
    """Curves whose radius of curvature exceeds twice the offset (the offset then has no
    cusps, and every true offset point is at distance h from the base curve)."""
    rng = np.random.default_rng(100 + ct)
    checked = 0
    for _ in range(200):
        n = {S.QUAD: 3, S.CUBIC: 4, S.RATIONAL: 3}[ct]
        c = S.Curve(ct, rng.uniform(0, 100, (n, 2)), rng.uniform(0.4, 2.0))
        h = float(rng.uniform(0.5, 3.0))
        if min_curvature_radius(c) < 2 * h:
            continue
        pieces = S.offset_curves(c, h)
        for a, b in zip(pieces[:-1], pieces[1:]):
            np.testing.assert_allclose(a.p[-1], b.p[0], atol=1e-9)
        np.testing.assert_allclose(pieces[0].p[0], c.offset_point(h, 0.0), atol=1e-9)
        np.testing.assert_allclose(pieces[-1].p[-1], c.offset_point(h, 1.0), atol=1e-9)
        samples = np.array([p.point(t) for p in pieces for t in np.linspace(0, 1, 17)])
        err = np.abs(dist_to_curve(c, samples, 6000) - h)
        assert err.max() < 0.04, (ct, h, err.max())   # the 1/32 px target plus sampling
        checked += 1
        if checked == 25:
            break
    assert checked == 25


def outline(vg):
    return vg.strokeToFill()


def test_outlines_are_closed_contours():
    vg = VGContainer(128, 128)
    vg.addStrokePath(rgba(0, 0, 0), 3.0)
    vg.addCubic((10, 10), (60, 0), (20, 80), (90, 90))
    vg.addQuadratic((90, 90), (110, 40), (70, 20))
    vg.addLinear((70, 20), (40, 60))
    o = outline(vg)
    a = o.arrays()
    assert a.n_paths == 1
    ends = []
    for c in range(a.n_curves):
        p = a.curve_vertex_pos[c]
        n = int(a.curve_type[c]) & 7
        ends.append((a.vertex[p], a.vertex[p + n - 1]))
    for (s0, e0), (s1, e1) in zip(ends, ends[1:] + ends[:1]):
        np.testing.assert_array_equal(e0, s1)


def render(vg, W, H, spp=32):
    return pyoracle.run(vg.arrays(), W, H, spp, srgb=False).get("image").reshape(H, W, 4)


def coverage(img):
    return (255.0 - img[..., 0].astype(np.float64)) / 255.0


def test_stroked_line_is_a_rectangle():
    vg = VGContainer(64, 32)
    vg.addStrokePath(rgba(0, 0, 0), 4.0)
    vg.addLinear((8.5, 16.5), (56.5, 16.5))
    cov = coverage(render(outline(vg), 64, 32))
    # x 8.5 .. 56.5 and y 14.5 .. 18.5: full pixels inside, half pixels on the four edges
    np.testing.assert_allclose(cov[15:18, 9:56], 1.0)
    np.testing.assert_allclose(cov[15:18, 8], 0.5, atol=1 / 255)
    np.testing.assert_allclose(cov[15:18, 56], 0.5, atol=1 / 255)
    np.testing.assert_allclose(cov[14, 9:56], 0.5, atol=1 / 255)
    np.testing.assert_allclose(cov[18, 9:56], 0.5, atol=1 / 255)
    assert cov[13].max() == 0 and cov[19].max() == 0 and cov[:, :8].max() == 0


def test_stroked_circle_covers_the_annulus():
    r, h, cx, cy = 20.0, 2.0, 32.0, 32.0
    w = float(np.sqrt(0.5))
    vg = VGContainer(64, 64)
    vg.addStrokePath(rgba(0, 0, 0), 2 * h)
    quads = [((cx + r, cy), (cx + r, cy + r), (cx, cy + r)), ((cx, cy + r), (cx - r, cy + r), (cx - r, cy)),
             ((cx - r, cy), (cx - r, cy - r), (cx, cy - r)), ((cx, cy - r), (cx + r, cy - r), (cx + r, cy))]
    for q in quads:
        vg.addRational(q[0], q[1], q[2], w)
    cov = coverage(render(outline(vg), 64, 64))
    area = np.pi * ((r + h) ** 2 - (r - h) ** 2)
    assert abs(cov.sum() - area) < 0.01 * area
    assert cov[32, 32] == 0 and cov[32, 32 + 20] == 1.0


def test_bevel_joined_corner():
    vg = VGContainer(64, 64)
    vg.addStrokePath(rgba(0, 0, 0), 4.0)
    vg.addLinear((10, 10), (50, 10))
    vg.addLinear((50, 10), (50, 50))
    cov = coverage(render(outline(vg), 64, 64))
    # two 40 x 4 bands overlapping in a 2 x 2 square at the corner, plus the bevel triangle
    # closing the outer corner (legs 2 and 2): 160 + 160 - 4 + 2
    assert abs(cov.sum() - 318.0) < 1.0


def test_fill_and_stroke_paths_from_svg():
    doc = ('<svg width="64" height="64" viewBox="0 0 64 64">'
           '<path d="M 8 8 L 40 8 L 40 40 L 8 40 Z" fill="#f00" stroke="#00f" stroke-width="3"/>'
           '<path d="M 50 10 L 50 50" fill="none" stroke="#0f0" stroke-width="2"/></svg>')
    plain = svg.loads(doc)
    assert plain.pathNumber() == 2 and plain.fillInfo()[0] == rgba(0, 0, 255)
    vg = svg.loads(doc, stroke_to_fill=True)
    assert vg.source == [(0, "fill"), (0, "stroke"), (1, "stroke")]
    assert vg.fillInfo() == [rgba(255, 0, 0), rgba(0, 0, 255), rgba(0, 255, 0)]
    img = pyoracle.run(vg.arrays(), 64, 64, 32, srgb=False).get("image").reshape(64, 64, 4)
    # the red fill inside, the blue 3-px border over its edge (y flipped by the viewBox)
    assert tuple(img[64 - 24, 24, :3]) == (255, 0, 0)
    assert tuple(img[64 - 8, 24, :3])[2] > 200 and img[64 - 8, 24, 0] < 60
    # the green 2-px line x = 49..51 for y in 10..50
    assert tuple(img[64 - 30, 49, :3]) == (0, 255, 0) and tuple(img[64 - 30, 52, :3]) == (255, 255, 255)


def test_sample_svg_with_stroke_to_fill():
    vg = svg.load(os.path.join(HERE, "golden", "sample.svg"), stroke_to_fill=True)
    kinds = [k for _, k in vg.source]
    assert kinds.count("stroke") == 1 and vg.pathNumber() == len(kinds)
    img = pyoracle.run(vg.arrays(), vg.width, vg.height, 16).get("image")
    assert img.size
