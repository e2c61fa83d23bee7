"""Stroke-to-fill: stroked paths become filled outline paths the renderer draws as fills.

Restates `VGContainer::strokeToOffsetCurve` (gpu-scanline/src/vg_container.cpp:555-1353)
with its offset curves approximated as `SimpleBezierCurve::segApproxOffsetCurve`
(gpu-scanline/src/bezier_curve.cpp:186-600) does:

* a path with a fill is copied as a fill path, every contour closed with a line when its
  first and last vertices differ (vg_container.cpp:566-652);
* a stroked path adds a path after it, filled with the stroke paint under the nonzero
  rule; each contour becomes one closed outline (vg_container.cpp:655-1349):
  a butt cap from the negative to the positive side at the start, the curves offset by
  half the stroke width to the right-hand normal (dy, -dx)/|d| with bevel lines between
  consecutive curves, a butt cap at the end, and the curves reversed and offset again (the
  other side, walked backwards) with bevel lines between them;
* each offset curve is approximated by curves of its own type (bezier_curve.cpp:548-596):
  recursive halving of the parameter range, each piece replaced by the Tiller-Hanson
  construction (end points offset, inner control points where the offset tangent lines
  meet, bezier_curve.cpp:365-545) until its midpoint is within 1/32 px of the true offset
  midpoint or the depth reaches 8.

The reference's driver never calls this conversion (its `--stroke-to-fill` option is a
commented line of vg_default.cfg), and the active code has slips that leave the outline
open.  This restatement follows the evident intent where they are: the bevel lines of the
negative side join negative-side points (the reference's code joins positive-side points
there), a degenerate Tiller-Hanson intersection of a quadratic or rational falls back to a
line between the offset ends (the reference overwrites that line's end with a NaN), and
cubics with a coincident control point take the tangent of the next distinct one.  Caps
are butt and joins bevel, which is what the reference's active code draws whatever the
path asks for (it forces STROKE_JOIN_BEVEL, vg_container.cpp:910).  Arithmetic is float64,
rounded to float32 vertices.
"""
import math

import numpy as np

from . import _abi

LINEAR, QUAD, CUBIC, RATIONAL = _abi.CURVE_LINEAR, _abi.CURVE_QUADRATIC, _abi.CURVE_CUBIC, _abi.CURVE_RATIONAL
TOLERANCE = 1.0 / 32.0
MAX_DEPTH = 8


class Curve:
    """A Bezier segment: type code, control points (n x 2 float64), rational weight."""

    def __init__(self, ct, pts, w=1.0):
        self.ct = ct
        self.p = np.asarray(pts, dtype=np.float64).reshape(-1, 2)
        self.w = float(w)

    def reversed(self):
        return Curve(self.ct, self.p[::-1].copy(), self.w)

    def point(self, t):
        p, s = self.p, 1.0 - t
        if self.ct == LINEAR:
            return p[0] * s + p[1] * t
        if self.ct == QUAD:
            return p[0] * s * s + p[1] * 2 * s * t + p[2] * t * t
        if self.ct == CUBIC:
            return p[0] * s ** 3 + p[1] * 3 * s * s * t + p[2] * 3 * s * t * t + p[3] * t ** 3
        num = p[0] * s * s + p[1] * (2 * s * t * self.w) + p[2] * t * t
        return num / (s * s + 2 * s * t * self.w + t * t)

    def deriv(self, t):
        p, s = self.p, 1.0 - t
        if self.ct == LINEAR:
            return p[1] - p[0]
        if self.ct == QUAD:
            return 2 * s * (p[1] - p[0]) + 2 * t * (p[2] - p[1])
        if self.ct == CUBIC:
            return 3 * s * s * (p[1] - p[0]) + 6 * s * t * (p[2] - p[1]) + 3 * t * t * (p[3] - p[2])
        w = self.w
        num = p[0] * s * s + p[1] * (2 * s * t * w) + p[2] * t * t
        den = s * s + 2 * s * t * w + t * t
        dnum = -2 * s * p[0] + (2 * w) * (1 - 2 * t) * p[1] + 2 * t * p[2]
        dden = -2 * s + 2 * w * (1 - 2 * t) + 2 * t
        return (dnum * den - num * dden) / (den * den)

    def normal(self, t):
        """Unit right-hand normal (dy, -dx)/|d|; where the derivative vanishes (a control point
        on an end point) the direction comes from the nearest parameter where it does not,
        as the reference's cubic tangent searches (bezier_curve.h:166-193)."""
        d = self.deriv(t)
        n = math.hypot(d[0], d[1])
        step = 1.0 / 1024.0
        while n < 1e-9 and step <= 1.0:
            a, b = min(max(t - step, 0.0), 1.0), min(max(t + step, 0.0), 1.0)
            d = self.point(b) - self.point(a)
            n = math.hypot(d[0], d[1])
            step *= 2
        if n == 0.0:
            return np.zeros(2)
        return np.array([d[1] / n, -d[0] / n])

    def offset_point(self, h, t):
        return self.point(t) + h * self.normal(t)

    def sub(self, t0, t1):
        """The piece over [t0, t1], of the same type (de Casteljau; rationals in homogeneous
        coordinates, renormalised to unit end weights)."""
        if self.ct == RATIONAL:
            hp = np.array([[self.p[0, 0], self.p[0, 1], 1.0],
                           [self.p[1, 0] * self.w, self.p[1, 1] * self.w, self.w],
                           [self.p[2, 0], self.p[2, 1], 1.0]])
            q = _blossom_piece(hp, t0, t1)
            w0, w1, w2 = q[0, 2], q[1, 2], q[2, 2]
            k = math.sqrt(w0 * w2) if w0 * w2 > 0 else 1.0
            return Curve(RATIONAL, [q[0, :2] / w0, q[1, :2] / w1, q[2, :2] / w2], w1 / k)
        return Curve(self.ct, _blossom_piece(self.p, t0, t1), self.w)


def _casteljau_levels(pts, t):
    levels = [list(pts)]
    while len(levels[-1]) > 1:
        prev = levels[-1]
        levels.append([prev[i] * (1 - t) + prev[i + 1] * t for i in range(len(prev) - 1)])
    return levels


def _blossom_piece(p, t0, t1):
    """Control points of a polynomial segment's piece over [t0, t1]: a de Casteljau cut at
    t0 keeps the right part (the last point of every level), a cut of that at the rescaled
    t1 keeps the left part (the first point of every level)."""
    pts = [r.copy() for r in np.asarray(p, dtype=np.float64)]
    right = [lv[-1] for lv in _casteljau_levels(pts, t0)][::-1]
    if t0 >= 1.0:
        return np.array([right[0]] * len(pts))
    u = (t1 - t0) / (1.0 - t0)
    return np.array([lv[0] for lv in _casteljau_levels(right, u)])


def _intersect(p0, d0, p2, d2):
    """Where the line p0 + a d0 meets the line p2 + b d2; None when they are parallel."""
    det = d0[0] * d2[1] - d0[1] * d2[0]
    scale = math.hypot(*d0) * math.hypot(*d2)
    if scale == 0.0 or abs(det) <= 1e-12 * scale:
        return None
    y = p2 - p0
    a = (y[0] * d2[1] - y[1] * d2[0]) / det
    q = p0 + a * d0
    return q if np.all(np.isfinite(q)) else None


def _nonzero(v):
    return math.hypot(v[0], v[1]) > 0.0


def approx_offset(c, h, vf, vl):
    """One Tiller-Hanson approximation of curve c offset by h, with end points vf and vl
    (bezier_curve.cpp:365-545)."""
    p = c.p
    if c.ct == LINEAR:
        return Curve(LINEAR, [vf, vl])
    if c.ct in (QUAD, RATIONAL):
        q = _intersect(vf, p[1] - p[0], vl, p[1] - p[2])
        if q is None:
            return Curve(LINEAR, [vf, vl])
        return Curve(c.ct, [vf, q, vl], c.w)
    # cubic: the offset of the middle control edge guides the inner points
    d0 = p[1] - p[0] if _nonzero(p[1] - p[0]) else p[2] - p[0]
    d3 = p[2] - p[3] if _nonzero(p[2] - p[3]) else p[1] - p[3]
    mid = p[2] - p[1]
    if not _nonzero(mid):
        q = _intersect(vf, d0, vl, d3)
        if q is None:
            return Curve(LINEAR, [vf, vl])
        return Curve(CUBIC, [vf, q, q, vl])
    nm = np.array([mid[1], -mid[0]]) / math.hypot(mid[0], mid[1])
    m1, m2 = p[1] + h * nm, p[2] + h * nm
    q1 = _intersect(vf, d0, m2, m1 - m2)
    q2 = _intersect(vl, d3, m1, m2 - m1)
    return Curve(CUBIC, [vf, m1 if q1 is None else q1, m2 if q2 is None else q2, vl])


def offset_curves(c, h):
    """Curves approximating c offset by h (bezier_curve.cpp:548-596)."""
    if h == 0.0:
        return [c]
    if c.ct == LINEAR:
        return [approx_offset(c, h, c.offset_point(h, 0.0), c.offset_point(h, 1.0))]
    out = []

    def subdiv(a, b, depth):
        piece = approx_offset(c.sub(a, b), h, c.offset_point(h, a), c.offset_point(h, b))
        want = c.offset_point(h, 0.5 * (a + b))
        got = piece.point(0.5)
        if math.hypot(*(want - got)) < TOLERANCE or depth >= MAX_DEPTH:
            out.append(piece)
            return
        m = 0.5 * (a + b)
        subdiv(a, m, depth + 1)
        subdiv(m, b, depth + 1)

    subdiv(0.0, 1.0, 1)
    return out


def contour_outline(curves, half_width):
    """The closed outline of one stroked contour (vg_container.cpp:690-1115): a list of
    Curves, walked positive side forward, negative side backward."""
    h = float(half_width)
    out = []

    def line(a, b):
        if not np.array_equal(a, b):
            out.append(Curve(LINEAR, [a, b]))

    pos0 = [c.offset_point(h, 0.0) for c in curves]
    pos1 = [c.offset_point(h, 1.0) for c in curves]
    neg0 = [c.offset_point(-h, 0.0) for c in curves]
    neg1 = [c.offset_point(-h, 1.0) for c in curves]
    line(neg0[0], pos0[0])                       # start cap (butt)
    for i, c in enumerate(curves):
        out.extend(offset_curves(c, h))
        if i + 1 < len(curves):
            line(pos1[i], pos0[i + 1])           # bevel join
    line(pos1[-1], neg1[-1])                     # end cap (butt)
    for i in range(len(curves) - 1, -1, -1):
        out.extend(offset_curves(curves[i].reversed(), h))
        if i > 0:
            line(neg0[i], neg1[i - 1])           # bevel join of the negative side
    return out


def _container_curve(vg, c):
    ct = vg._ctype[c]
    p0 = vg._cpos[c]
    pts = vg._v[p0:p0 + (ct & 7)]
    return Curve(ct, pts, vg._cw[c] if ct == RATIONAL else 1.0)


def _add(vg, c):
    pts = [(float(np.float32(x)), float(np.float32(y))) for x, y in c.p]
    if c.ct == LINEAR:
        vg.addLinear(*pts)
    elif c.ct == QUAD:
        vg.addQuadratic(*pts)
    elif c.ct == CUBIC:
        vg.addCubic(*pts)
    else:
        vg.addRational(pts[0], pts[1], pts[2], float(np.float32(c.w)))


def stroke_to_fill(vg):
    """The container with every stroked path turned into fill geometry (module notes).
    Returns a new container; `source` on it lists, per new path, (original path, "fill" or
    "stroke")."""
    from .scene import FT_COLOR, FT_GRADIENT, FT_NONE, ST_COLOR, ST_NONE, VGContainer
    out = VGContainer(vg.width, vg.height, vg.x, vg.y)
    out.gradient_ramp, out.gradient_table = list(vg.gradient_ramp), list(vg.gradient_table)
    out.svg_gradients = getattr(vg, "svg_gradients", [])
    out.source = []
    by_path = {}
    for first, end, owner in vg.contours():
        by_path.setdefault(owner, []).append((first, end))
    for pi in range(vg.pathNumber()):
        contours = by_path.get(pi, [])
        if vg._ftype[pi] != FT_NONE:
            npi = out.addPath(vg._rule[pi], vg._color[pi], vg._ftype[pi], vg._findex[pi])
            out.source.append((pi, "fill"))
            for first, end in contours:
                if end <= first:
                    continue
                out.addContour()
                for c in range(first, end):
                    _add(out, _container_curve(vg, c))
                # always close a filled contour (vg_container.cpp:643-651)
                a = _container_curve(vg, first).p[0]
                b = _container_curve(vg, end - 1).p[-1]
                if not np.array_equal(a, b):
                    _add(out, Curve(LINEAR, [b, a]))
        if vg._stype[pi] != ST_NONE:
            grad = vg._stype[pi] != ST_COLOR
            npi = out.addPath(0, vg._scolor[pi], FT_GRADIENT if grad else FT_COLOR,
                              vg._sindex[pi] if grad else 0xFFFFFFFF)
            out.source.append((pi, "stroke"))
            half = 0.5 * float(vg._swidth[pi])
            for first, end in contours:
                curves = [_container_curve(vg, c) for c in range(first, end)]
                if not curves:
                    continue
                out.addContour()
                for c in contour_outline(curves, half):
                    _add(out, c)
    return out
