"""SVG loader: the reference's SVG::load semantics (gpu-scanline/src/svg_load.cpp:189-1550)
on hand-written documents, with every expected vertex worked out by hand from the SVG
text and the reference's rules (its quirks included), not by the loader under test.
"""
import math

import numpy as np
import pytest

from scanline import FT_COLOR, FT_GRADIENT, ST_COLOR, rgba, svg
from scanline import CURVE_CUBIC as C, CURVE_LINEAR as L, CURVE_QUADRATIC as Q


def doc(body, vb="0 0 100 100"):
    return '<svg xmlns="http://www.w3.org/2000/svg" viewBox="%s">%s</svg>' % (vb, body)


def curves(vg):
    """[(type, [(x, y), ...]), ...] in curve order, y flipped back (viewBox 0 0 100 100)."""
    out = []
    n = vg.curveNumber()
    for k in range(n):
        p0 = vg._cpos[k]
        nv = {L: 2, Q: 3, C: 4}.get(vg._ctype[k], 3)
        out.append((vg._ctype[k], [(x, 100 - y) for x, y in vg._v[p0:p0 + nv]]))
    return out


def test_absolute_path_and_close():
    vg = svg.loads(doc('<path d="M 10 10 L 20 10 L 20 30 Z" fill="#f00"/>', "0 0 100 50"))
    assert (vg.width, vg.height) == (100, 50)
    # y' = 50 - y; Z adds the closing line from the last vertex to the first
    assert vg._v == [(10, 40), (20, 40), (20, 40), (20, 20), (20, 20), (10, 40)]
    assert vg._color == [rgba(255, 0, 0)] and vg._ftype == [FT_COLOR] and vg._rule == [0]


def test_relative_commands():
    vg = svg.loads(doc('<path d="m 5 5 l 10 0 v 10 h -10 z"/>'))
    assert curves(vg) == [(L, [(5, 5), (15, 5)]), (L, [(15, 5), (15, 15)]), (L, [(15, 15), (5, 15)]),
                          (L, [(5, 15), (5, 5)])]
    assert vg._color == [rgba(0, 0, 0, 255)]  # default fill: black, opaque


def test_moveto_pairs_and_carried_values():
    # extra M pairs only move the pen; a lone number carries over to the next command
    vg = svg.loads(doc('<path d="M 1 2 3 4 L 5 6 M 7 L 8"/>'))
    # M (1,2) then (3,4); L (5,6): line (3,4)-(5,6); "M 7" leaves one value, "L 8" completes
    # the pair (7, 8): the M before it had moved nothing, so a line (5,6)-(7,8) follows
    assert curves(vg) == [(L, [(3, 4), (5, 6)]), (L, [(5, 6), (7, 8)])]


def test_smooth_curves_reflect_last_control_point():
    vg = svg.loads(doc('<path d="M 0 0 Q 10 10 20 0 S 40 10 50 0"/>'))
    # S reflects Q's control point (10,10) about (20,0): (30,-10)
    assert curves(vg) == [(Q, [(0, 0), (10, 10), (20, 0)]), (C, [(20, 0), (30, -10), (40, 10), (50, 0)])]


def test_close_does_not_move_pen_and_new_contours():
    vg = svg.loads(doc('<path d="M 0 0 L 10 0 L 10 10 Z l 5 0 M 50 50 L 60 50"/>'))
    # after Z the pen stays at (10,10): l 5 0 ends at (15,10); M after curves opens a contour
    assert curves(vg) == [(L, [(0, 0), (10, 0)]), (L, [(10, 0), (10, 10)]), (L, [(10, 10), (0, 0)]),
                          (L, [(10, 10), (15, 10)]), (L, [(50, 50), (60, 50)])]
    assert vg._cfirst == [0, 4] and vg._cowner == [0, 0]


def test_number_forms():
    vg = svg.loads(doc('<path d="M1.5-2.5L.5e1,1e-1-3-4"/>'))
    # "-" starts a new number; ".5e1" = 5; "1e-1" = 0.1 (its float); "-3-4" are two numbers
    (t0, p0), (t1, p1) = curves(vg)
    assert p0 == [(1.5, -2.5), (5.0, float(np.float32(100) - (np.float32(100) - np.float32(0.1))))]
    assert p1[1] == (-3.0, -4.0)


def test_quarter_arc_closed_form():
    vg = svg.loads(doc('<path d="M 10 0 A 10 10 0 0 1 0 10"/>'))
    (t, pts), = curves(vg)
    assert t == C
    k = 4.0 / 3.0 * math.tan(math.pi / 8.0) * 10.0  # quarter circle about (0,0)
    want = [(10, 0), (10, k), (k, 10), (0, 10)]
    np.testing.assert_allclose(np.array(pts), np.array(want), atol=2e-5)
    # a half circle splits in two quarter cubics
    vg = svg.loads(doc('<path d="M 10 0 A 10 10 0 0 1 -10 0"/>'))
    cs = curves(vg)
    assert len(cs) == 2 and all(t == C for t, _ in cs)
    np.testing.assert_allclose(np.array(cs[1][1][3]), [-10, 0], atol=2e-5)
    np.testing.assert_allclose(np.array(cs[0][1][3]), [0, 10], atol=2e-5)


def test_transforms():
    body = ('<g transform="translate(3)"><path d="M 1 1 L 2 1"/></g>'
            '<g transform="scale(2, 3) translate(1 1)"><path d="M 1 1 L 2 1"/></g>'
            '<g transform="matrix(1,0,0,1,5,6)"><path d="M 0 0 L 1 0" transform="rotate(30)"/></g>')
    vg = svg.loads(doc(body))
    cs = curves(vg)
    # translate with one value moves both axes
    assert cs[0][1] == [(4, 4), (5, 4)]
    # scale(2,3) * translate(1,1): p -> (2 (x + 1), 3 (y + 1))
    assert cs[1][1] == [(4, 6), (6, 6)]
    # rotate leaves the matrix as it was
    assert cs[2][1] == [(5, 6), (6, 6)]


def test_colours_and_opacity():
    body = ('<path d="M 0 0 L 1 0" fill="#1af"/>'
            '<path d="M 0 0 L 1 0" fill="rgb(1, 2, 300)" fill-opacity="0.5"/>'
            '<path d="M 0 0 L 1 0" fill="none"/>'
            '<path d="M 0 0 L 1 0" fill="#102030" fill-rule="evenodd"/>')
    vg = svg.loads(doc(body))
    # #1af -> 0x11, 0xaa, 0xff; 300 -> (u8) 44; alpha (u8)(255 * 0.5) = 127
    assert vg._color[:2] == [rgba(0x11, 0xAA, 0xFF), rgba(1, 2, 44, 127)]
    assert vg.fillInfo()[2] == 0 and vg._rule == [0, 0, 0, 1]
    assert vg._color[3] == rgba(0x10, 0x20, 0x30)
    a = svg.loads(doc('<path d="M 0 0 L 1 0" fill-opacity="0.5"/>'), a128=True)
    # 127 -> (u8)(127/255*127) = 63 -> (u8)(63/127*255) = 126
    assert a._color == [rgba(0, 0, 0, 126)]


def test_polygon():
    vg = svg.loads(doc('<polygon points="0,0 10,0 10,10" fill="#fff"/>'))
    assert curves(vg) == [(L, [(0, 0), (10, 0)]), (L, [(10, 0), (10, 10)]), (L, [(10, 10), (0, 0)])]


def test_gradient_fill_and_stroke_override():
    grad = ('<defs><linearGradient id="g" x1="0" y1="0" x2="1" y2="0" gradientUnits="userSpaceOnUse">'
            '<stop offset="0" stop-color="#000"/></linearGradient></defs>')
    with pytest.raises(svg.SVGError):
        svg.loads(doc(grad + '<path d="M 0 0 L 1 0" fill="url(#g)"/>'), gradients="raise")
    # the fill word names the path's gradient slot (path index + 1); skipped, no textures
    vg = svg.loads(doc(grad + '<path d="M 0 0 L 1 0" fill="url(#g)"/>'), gradients="skip")
    assert vg._ftype == [FT_GRADIENT] and vg.fillInfo() == [1] and vg.gradient_table == []
    vg = svg.loads(doc(grad + '<path d="M 0 0 L 1 0" fill="url(#g)"/>'))
    assert len(vg.gradient_table) == 12 and len(vg.gradient_ramp) == 1024
    # a colour stroke becomes the fill word, so this path renders without the gradient
    vg = svg.loads(doc(grad + '<path d="M 0 0 L 1 0" fill="url(#g)" stroke="#0f0"/>'))
    assert vg._stype == [ST_COLOR] and vg.fillInfo() == [rgba(0, 255, 0)]
    assert len(vg.svg_gradients) == 1 and vg.svg_gradients[0]["stops"][0]["color"] == (0, 0, 0, 255)


def test_where_the_reference_stops():
    body = '<path d="M 0 0 L 1 0" fill="#f00"/><rect x="0" y="0" width="5" height="5"/><path d="M 0 0 L 2 0"/>'
    with pytest.raises(svg.SVGError) as e:
        svg.loads(doc(body))
    assert e.value.partial is not None and e.value.partial.pathNumber() == 1
    part = svg.loads(doc(body), on_error="partial")
    assert part.pathNumber() == 1 and part._v == [(0, 100), (1, 100)]
    # a named colour stops the load after the path was made but before its paint was set
    part = svg.loads(doc('<path d="M 0 0 L 1 0" fill="red"/>'), on_error="partial")
    assert part.pathNumber() == 1 and part.fillInfo() == [0]
    # a gradient that is not defined yet
    with pytest.raises(svg.SVGError):
        svg.loads(doc('<path d="M 0 0 L 1 0" fill="url(#nope)"/>'))
    # a comma-separated viewBox stops the load at the svg element
    with pytest.raises(svg.SVGError):
        svg.loads(doc('<path d="M 0 0 L 1 0"/>', "0,0,10,10"))


def test_undefined_inputs_raise():
    for body in ('<path d="M 0 0 L 1 0" fill="rgba(1,2,3,0.5)"/>', '<circle cx="5" cy="5" r="3"/>',
                 '<path d="M 0 0 L 1 0" color="#fff"/>', '<path d="M 0 0 Q 1 1 2 0 T 3 3"/>'):
        with pytest.raises(svg.SVGError) as e:
            svg.loads(doc(body))
        assert e.value.partial is None


def test_skipped_elements_and_namespaces():
    body = ('<line x1="0" y1="0" x2="5" y2="5"/><text>t</text>'
            '<svg:path xmlns:svg="http://www.w3.org/2000/svg" d="M 0 0 L 1 1"/>'
            '<g><path d="M 0 0 L 3 0"/></g>')
    vg = svg.loads(doc(body))
    assert vg.pathNumber() == 1 and curves(vg) == [(L, [(0, 0), (3, 0)])]


def test_sample_file():
    import os
    vg = svg.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sample.svg"))
    assert (vg.width, vg.height) == (60, 44)
    a = vg.arrays()
    assert a.n_paths == 7
    # the first path: M 4 4 -> (4 - 2, 44 - (4 - 1)) = (2, 41)
    assert a.vertex[0].tolist() == [2.0, 41.0]
    # the gradient-filled path with a #333 stroke renders in the stroke colour
    assert int(a.fill_color[-1]) == rgba(0x33, 0x33, 0x33)
    # the group's fill-opacity 0.75: alpha (u8)(191.25) = 191
    assert int(a.fill_color[2]) >> 24 == 191
