"""SVG scenes into a VGContainer, with the reference's SVG::load semantics
(gpu-scanline/src/svg_load.cpp:189-1550), which read a subset of SVG in their own way.

What is read, as the reference reads it:

* the document's root element when it is named exactly `svg` (no namespace handling:
  `svg:path` is not a path); its `width`, `height` and `viewBox` attributes only;
* `g` and `defs` descend (`g` with the presentation attributes below), `path` and
  `polygon` make paths, `linearGradient` / `radialGradient` / `stop` register gradients;
  every other element is skipped with its children (svg_load.cpp:1197-1208);
* presentation attributes, inherited down the tree (svg_load.cpp:547-660): `fill`
  (`none`, `url(#id)` of an already defined gradient, `#rgb`, `#rrggbb`, `rgb(r,g,b)`;
  other forms stop the load), `fill-rule`, `fill-opacity` (alpha = (u8)(alpha * opacity)),
  `stroke*`, `transform` (`matrix`, `translate` -- one argument moves both axes --,
  `scale`; `rotate` / `skewX` / `skewY` leave the matrix unchanged, as the reference's
  release build does), `style` (only `stop-color` / `stop-opacity`);
  `opacity`, `visibility` and the dash attributes are read and ignored;
* path data by the reference's state machine (svg_load.cpp:662-1034): commands M L H V C S
  Q A Z in both cases (T is not a command: met after a number it hangs the reference's
  parser, and here raises), numbers of digits, `.`, `e` and a leading `-` (after `e`
  too); a number without its command's full count carries over to
  the next command; M pairs only move the pen; S and Q reflect the last control point
  kept by C / S / Q / A (never reset); Z closes with a line to the contour's first
  vertex when they differ, without moving the pen; arcs become cubics of at most a
  quarter turn (vg_container.cpp:142-266);
* the element's transform applies to its vertices when the element ends, in float;
* at the end every vertex moves by minus the viewBox origin and flips in y about the
  viewBox height; the container's size is the viewBox size truncated to integers.

Where the reference throws -- a `rect` (its reader table sends `rect` to the line reader,
which throws), an unknown colour form, a gradient not yet defined, a stroke cap or join
it does not know, a `style` key other than the two stop keys, an unparseable transform --
its load stops there and keeps what it has read.  Here that raises `SVGError` with the
partial container (post-processed like the reference's) in `.partial`, or, with
`on_error="partial"`, returns it.  Inputs on which the reference reads uninitialised
values (`rgba(...)` and percentage colours, `circle` / `ellipse` without x / y / width /
height, a `color` attribute, which pops the reference's attribute stack, unparseable
numbers) raise `SVGError` with no partial result.

Colour strokes fill the path with the stroke colour (see `VGContainer.fillInfo`), as the
reference's renderer draws its input; with `stroke_to_fill=True` every stroked path is
instead turned into an outline path of its stroke width (`scanline.stroke`), drawn after its
fill.

Gradient fills.  The reference's SVG loader keeps the gradient definitions but builds no
gradient textures (only its RVG loader does, rvg.cpp:490-758), so its renderer reads an
empty gradient table for such paths.  With `gradients="render"` (the default) this loader
builds the textures with the RVG loader's construction (`gradient.GradientTextures`): a
`linearGradient` from (x1, y1) to (x2, y2), a `radialGradient` of centre (cx, cy), focus
(fx, fy) (the centre when absent) and radius r, stops of the stop colours and opacities, the
path's fill-opacity, and a paint transform that maps gradient coordinates to the scene's
final coordinates: the gradientTransform, then the element's transform, then the viewBox
shift and y flip.  `"skip"` leaves gradient fills without textures (they draw nothing) and
`"raise"` refuses them.  `srgb` selects the ramp encoding (the renderer's sRGB setting).
"""
import math
import re
import xml.parsers.expat

import numpy as np

from .gradient import GradientTextures
from .scene import FT_COLOR, FT_GRADIENT, FT_NONE, ST_COLOR, ST_GRADIENT, ST_NONE, VGContainer, rgba
from .vgfile import decimal_to_f32

F = np.float32


class SVGError(ValueError):
    def __init__(self, msg, partial=None):
        super().__init__(msg)
        self.partial = partial


class _Stop(Exception):
    """A condition on which the reference throws: its load stops there."""


# ---------------------------------------------------------------------------------
# XML without namespace processing: element names as written, attributes in order
# ---------------------------------------------------------------------------------
class _Node:
    __slots__ = ("name", "attrs", "children")

    def __init__(self, name, attrs):
        self.name, self.attrs, self.children = name, attrs, []


def _parse_xml(text):
    p = xml.parsers.expat.ParserCreate()
    p.ordered_attributes = True
    doc = _Node("", [])
    stack = [doc]

    def start(name, attrs):
        n = _Node(name, list(zip(attrs[0::2], attrs[1::2])))
        stack[-1].children.append(n)
        stack.append(n)

    def end(_name):
        stack.pop()

    p.StartElementHandler, p.EndElementHandler = start, end
    try:
        p.Parse(text, True)
    except xml.parsers.expat.ExpatError as e:
        raise SVGError("XML parse error: %s" % e)
    return doc


# ---------------------------------------------------------------------------------
# numbers (sscanf "%f" / "%d": leading whitespace, the longest number at the start)
# ---------------------------------------------------------------------------------
_FNUM = re.compile(r"[ \t\n\v\f\r]*([+-]?(?:[0-9]+\.?[0-9]*|\.[0-9]+)(?:[eE][+-]?[0-9]+)?)")
_INUM = re.compile(r"[ \t\n\v\f\r]*([+-]?[0-9]+)")


def _f32(tok):
    return F(decimal_to_f32([tok])[0])


def _scan_floats(s, n):
    """sscanf(s, "%f %f ...") with n conversions: the values read before the first failure."""
    out, pos = [], 0
    for _ in range(n):
        m = _FNUM.match(s, pos)
        if not m:
            break
        out.append(_f32(m.group(1)))
        pos = m.end()
    return out


def _to_float(s, what):
    v = _scan_floats(s, 1)
    if not v:
        raise SVGError("%s: %r is not a number (the reference reads an uninitialised value)" % (what, s))
    return v[0]


# ---------------------------------------------------------------------------------
# 3x3 float matrices in glm's column-major layout, products summed left to right
# ---------------------------------------------------------------------------------
def _identity():
    return [[F(1), F(0), F(0)], [F(0), F(1), F(0)], [F(0), F(0), F(1)]]


def _mat_mul(a, b):
    return [[(a[0][r] * b[c][0] + a[1][r] * b[c][1]) + a[2][r] * b[c][2] for r in range(3)] for c in range(3)]


def _apply(m, x, y):
    x, y, one = F(x), F(y), F(1)
    px = (m[0][0] * x + m[1][0] * y) + m[2][0] * one
    py = (m[0][1] * x + m[1][1] * y) + m[2][1] * one
    pz = (m[0][2] * x + m[1][2] * y) + m[2][2] * one
    return px / pz, py / pz


def _alpha(c):
    return "a" <= c <= "z" or "A" <= c <= "Z"


def _parse_transform(s):
    """parseTransform (svg_load.cpp:310-448)."""
    t = "".join(" " if c in ", \t\n\r" else c for c in s)
    tmat = _identity()
    i, n = 0, len(t)
    while True:
        while i < n and not _alpha(t[i]):
            i += 1
        if i == n:
            return tmat
        b = i
        while i < n and _alpha(t[i]):
            i += 1
        if i == n:
            raise _Stop("transform")
        label = t[b:i]
        if label not in ("matrix", "translate", "scale", "rotate", "skewX", "skewY"):
            raise _Stop("transform")
        while i < n and t[i] != "(":
            i += 1
        if i == n:
            raise _Stop("transform")
        vb = i + 1
        while i < n and t[i] != ")":
            i += 1
        if i == n:
            raise _Stop("transform")
        args = t[vb:i]
        i += 1
        m = _identity()
        if label == "matrix":
            v = _scan_floats(args, 6)
            if len(v) != 6:
                raise _Stop("transform:matrix()")
            m[0][0], m[0][1], m[1][0], m[1][1], m[2][0], m[2][1] = v
        elif label == "translate":
            v = _scan_floats(args, 2)
            if not v:
                raise _Stop("transform:translate()")
            m[2][0], m[2][1] = (v[0], v[0]) if len(v) == 1 else (v[0], v[1])
        elif label == "scale":
            v = _scan_floats(args, 2)
            if not v:
                raise _Stop("transform:scale()")
            m[0][0], m[1][1] = (v[0], v[0]) if len(v) == 1 else (v[0], v[1])
        tmat = _mat_mul(tmat, m)


# ---------------------------------------------------------------------------------
# colours (parseColor, svg_load.cpp:258-307)
# ---------------------------------------------------------------------------------
def _hex_fields(s, width, count):
    out, pos = [], 0
    for _ in range(count):
        k = 0
        while k < width and pos + k < len(s) and s[pos + k] in "0123456789abcdefABCDEF":
            k += 1
        if k == 0:
            raise SVGError("colour %r: bad hex digits (the reference reads uninitialised values)" % s)
        out.append(int(s[pos:pos + k], 16))
        pos += k
    return out


def _color(s):
    """(r, g, b, a) bytes."""
    if s.startswith("rgb"):
        if s.startswith("rgba") or "%" in s:
            raise SVGError("colour %r: the reference reads it as rgb() from the wrong offset or "
                           "stops at '%%' (uninitialised values)" % s)
        body = s[4:]
        k = body.find(")")
        if k >= 0:
            body = body[:k] + " " + body[k + 1:]
        vals, pos = [], 0
        for j in range(3):
            if j:
                if body[pos:pos + 1] != ",":
                    break
                pos += 1
            m = _INUM.match(body, pos)
            if not m:
                break
            vals.append(int(m.group(1)))
            pos = m.end()
        if len(vals) != 3:
            raise SVGError("colour %r: fewer than three integers (uninitialised values)" % s)
        return (vals[0] & 255, vals[1] & 255, vals[2] & 255, 255)
    if s[:1] == "#":
        if len(s) == 4:
            r, g, b = _hex_fields(s[1:], 1, 3)
            return ((r * 17) & 255, (g * 17) & 255, (b * 17) & 255, 255)
        if len(s) == 7:
            r, g, b = _hex_fields(s[1:], 2, 3)
            return (r, g, b, 255)
        raise _Stop("unsupported color format %r" % s)
    raise _Stop("unknown color format %r" % s)


def _u8(v):
    """(uint8_t) of a float: truncation, modulo 256."""
    return int(np.trunc(v)) & 255 if np.isfinite(v) else 0


# ---------------------------------------------------------------------------------
# SVG arcs as cubics (VGContainer::addSVGArc, vg_container.cpp:142-266), in double with
# the reference's float roundings
# ---------------------------------------------------------------------------------
def _add_svg_arc(vg, start, rx_in, ry_in, rot, large, sweep, end):
    if rx_in == 0 and ry_in == 0:
        raise SVGError("arc with zero radii (the reference divides by zero)")
    pi = math.acos(-1.0)
    rad = pi / 180.0

    def vec_angle(ux, uy, vx, vy):
        ta, tb = math.atan2(uy, ux), math.atan2(vy, vx)
        return tb - ta if tb >= ta else 2.0 * pi - (ta - tb)

    sx, sy = F(start[0]), F(start[1])
    ex, ey = F(end[0]), F(end[1])
    RX, RY = F(rx_in), F(ry_in)
    sin_phi = math.sin(float(F(rot)) * rad)
    cos_phi = math.cos(float(F(rot)) * rad)
    dx, dy = float(sx - ex), float(sy - ey)
    x1 = cos_phi * dx / 2.0 + sin_phi * dy / 2.0
    y1 = -sin_phi * dx / 2.0 + cos_phi * dy / 2.0
    rx2, ry2 = RX * RX, RY * RY  # float products
    num = float(rx2 * RY * RY) - float(rx2) * y1 * y1 - float(ry2) * x1 * x1
    rx, ry = RX, RY
    if num < 0.0:
        s = F(math.sqrt(1.0 - num / float(rx2 * RY * RY)))
        rx, ry = rx * s, ry * s
        root = 0.0
    else:
        sign = -1.0 if (large == 1 and sweep == 1) or (large == 0 and sweep == 0) else 1.0
        root = sign * math.sqrt(num / (float(rx2) * y1 * y1 + float(ry2) * x1 * x1))
    rxd, ryd = float(rx), float(ry)
    cxd = root * rxd * y1 / ryd
    cyd = -root * ryd * x1 / rxd
    cx = cos_phi * cxd - sin_phi * cyd + float(sx + ex) / 2.0
    cy = sin_phi * cxd + cos_phi * cyd + float(sy + ey) / 2.0
    theta1 = vec_angle(1.0, 0.0, (x1 - cxd) / rxd, (y1 - cyd) / ryd)
    dtheta = vec_angle((x1 - cxd) / rxd, (y1 - cyd) / ryd, (-x1 - cxd) / rxd, (-y1 - cyd) / ryd)
    if sweep == 0 and dtheta > 0:
        dtheta -= 2.0 * pi
    elif sweep == 1 and dtheta < 0:
        dtheta += 2.0 * pi
    if not math.isfinite(dtheta):
        raise SVGError("arc with coincident end points or no extent (the reference computes NaN)")
    segments = int(math.ceil(abs(dtheta / (pi / 2.0))))
    if segments == 0:
        raise SVGError("arc of zero sweep (the reference then reads the vertices before it)")
    delta = dtheta / segments
    t = 8.0 / 3.0 * math.sin(delta / 4.0) * math.sin(delta / 4.0) / math.sin(delta / 2.0)
    px, py = float(sx), float(sy)
    for _ in range(segments):
        c1, s1 = math.cos(theta1), math.sin(theta1)
        theta2 = theta1 + delta
        c2, s2 = math.cos(theta2), math.sin(theta2)
        epx = cos_phi * rxd * c2 - sin_phi * ryd * s2 + cx
        epy = sin_phi * rxd * c2 + cos_phi * ryd * s2 + cy
        dx1 = t * (-cos_phi * rxd * s1 - sin_phi * ryd * c1)
        dy1 = t * (-sin_phi * rxd * s1 + cos_phi * ryd * c1)
        dxe = t * (cos_phi * rxd * s2 + sin_phi * ryd * c2)
        dye = t * (sin_phi * rxd * s2 - cos_phi * ryd * c2)
        vg.addCubic((F(px), F(py)), (F(px + dx1), F(py + dy1)), (F(epx + dxe), F(epy + dye)), (F(epx), F(epy)))
        theta1 = theta2
        px, py = float(F(epx)), float(F(epy))


# ---------------------------------------------------------------------------------
# path data (ar_path_d, svg_load.cpp:662-1034)
# ---------------------------------------------------------------------------------
def _lower(c):
    return chr(ord(c) + 32) if "A" <= c <= "Z" else c


_NVAL = {"m": 2, "l": 2, "h": 1, "v": 1, "c": 6, "s": 4, "q": 4, "a": 7}


def _path_data(vg, d):
    if not d:
        return
    vg.addPath(0, rgba(0, 0, 0, 255), FT_NONE, 0xFFFFFFFF)
    cfirst_v = vg.vertexNumber()
    vg.addContour()
    contour_curves0 = vg.curveNumber()
    state_cmd = True
    cmd = None
    nstr = ""
    value = [F(0)] * 7
    vi = 0
    relative = False
    lp0 = (F(0), F(0))
    lp1 = (F(0), F(0))
    i, n = 0, len(d)

    def point(k):
        p = (value[k], value[k + 1])
        return (p[0] + lp0[0], p[1] + lp0[1]) if relative else p

    while i <= n:
        cc = d[i] if i < n else "\0"
        lc = _lower(cc)
        if state_cmd:
            relative = "a" <= cc <= "z"
            if lc == "m":
                if vg.curveNumber() != contour_curves0:
                    vg.addContour()
                    contour_curves0 = vg.curveNumber()
                    cfirst_v = vg.vertexNumber()
                cmd, state_cmd = "m", False
            elif lc in "lhvqcsa":
                cmd, state_cmd = lc, False
            elif lc == "z":
                if cfirst_v != vg.vertexNumber():
                    fv, lv = vg._v[cfirst_v], vg._v[-1]
                    if fv != lv:
                        vg.addLinear(lv, fv)
            i += 1
            continue
        # number state
        if (nstr and nstr[-1] in "eE" and cc == "-") or "0" <= cc <= "9" or cc in "eE." or (not nstr and cc == "-"):
            nstr += cc
            i += 1
            continue
        if nstr:
            m = _FNUM.match(nstr)
            if not m:
                raise SVGError("path data: %r is not a number (the reference reads an uninitialised value)" % nstr)
            value[vi] = _f32(m.group(1))
            vi += 1
            nstr = ""
            if vi == _NVAL[cmd]:
                vi = 0
                if cmd == "m":
                    lp0 = point(0)
                elif cmd == "l":
                    p1 = point(0)
                    vg.addLinear(lp0, p1)
                    lp0 = p1
                elif cmd == "h":
                    p1 = (lp0[0] + value[0] if relative else value[0], lp0[1])
                    vg.addLinear(lp0, p1)
                    lp0 = p1
                elif cmd == "v":
                    p1 = (lp0[0], lp0[1] + value[0] if relative else value[0])
                    vg.addLinear(lp0, p1)
                    lp0 = p1
                elif cmd == "c":
                    p1, p2, p3 = point(0), point(2), point(4)
                    vg.addCubic(lp0, p1, p2, p3)
                    lp0, lp1 = p3, p2
                elif cmd == "s":
                    p1 = (lp0[0] + (lp0[0] - lp1[0]), lp0[1] + (lp0[1] - lp1[1]))
                    p2, p3 = point(0), point(2)
                    vg.addCubic(lp0, p1, p2, p3)
                    lp0, lp1 = p3, p2
                elif cmd == "q":
                    p1, p2 = point(0), point(2)
                    vg.addQuadratic(lp0, p1, p2)
                    lp0, lp1 = p2, p1
                elif cmd == "a":
                    end = point(5)
                    _add_svg_arc(vg, lp0, value[0], value[1], value[2], int(np.trunc(value[3])),
                                 int(np.trunc(value[4])), end)
                    lp0 = tuple(F(c) for c in vg._v[-1])
                    lp1 = tuple(F(c) for c in vg._v[-2])
        if cc in " \t\n\r,\0":
            i += 1
        elif lc in "mlhvqcsaz":
            state_cmd = True
        elif cc == "-":
            pass
        else:
            raise SVGError("path data: unexpected %r at %d (the reference loops forever here)" % (cc, i))


# ---------------------------------------------------------------------------------
def _new_attr():
    return dict(fill=FT_COLOR, rule=0, color=(0, 0, 0, 255), findex=0xFFFFFFFF, fopacity=F(1),
                stroke=ST_NONE, scolor=(0, 0, 0, 255), sindex=0xFFFFFFFF, swidth=F(1), sopacity=F(1),
                mat=_identity())


def _copy_attr(a):
    b = dict(a)
    b["mat"] = [list(c) for c in a["mat"]]
    return b


class _Loader:
    def __init__(self, a128, gradients):
        self.a128, self.gradients = a128, gradients
        self.vg = VGContainer()
        self.stack = []
        self.gmap = {}  # gradient id -> index
        self.pmat = {}  # path -> (element transform, fill opacity) when it ended
        self.vg.svg_gradients = []
        self.gradient = None
        self.stop = dict(color=(0, 0, 0, 255), opacity=F(1), offset=F(0))
        self.viewbox = (F(0), F(0), F(0), F(0))

    # presentation attributes (the default attribute reader, svg_load.cpp:1080-1104)
    def _url(self, v):
        ident = v[4:][:-1]
        if ident[:1] == "#":
            ident = ident[1:]
        if ident not in self.gmap:
            raise _Stop("url %r not found" % v)
        return self.gmap[ident]

    def presentation(self, name, v):
        top = self.stack[-1]
        if name == "color":
            raise SVGError("'color' attribute: the reference pops its attribute stack here")
        if name == "fill-rule":
            if v == "nonzero":
                top["rule"] = 0
            elif v == "evenodd":
                top["rule"] = 1
        elif name == "fill-opacity":
            top["fopacity"] = _to_float(v, name)
        elif name == "fill":
            if v == "none":
                top["fill"] = FT_NONE
            elif v[:3] == "url":
                top["findex"] = self._url(v)
                top["fill"] = FT_GRADIENT
            else:
                top["color"] = _color(v)
                top["fill"] = FT_COLOR
        elif name == "stroke-width":
            top["swidth"] = _to_float(v, name)
        elif name == "stroke-opacity":
            top["sopacity"] = _to_float(v, name)
        elif name in ("opacity", "stroke-miterlimit"):
            _to_float(v, name)
        elif name == "stroke-linecap":
            if v not in ("butt", "round", "square"):
                raise _Stop("illegal stroke-linecap %r" % v)
        elif name == "stroke-linejoin":
            if v not in ("miter", "round", "bevel"):
                raise _Stop("illegal stroke-linejoin %r" % v)
        elif name == "stroke":
            if v == "none":
                top["stroke"] = ST_NONE
            elif v[:3] == "url":
                top["sindex"] = self._url(v)
                top["stroke"] = ST_GRADIENT
            else:
                top["scolor"] = _color(v)
                top["stroke"] = ST_COLOR
        elif name == "transform":
            top["mat"] = _mat_mul(top["mat"], _parse_transform(v))
        elif name == "style":
            self.style(v)
        else:
            return False
        return True

    def style(self, v):
        pieces = v.split(";")
        if pieces and pieces[-1] == "":
            pieces.pop()
        for s in pieces:
            kv = [p for p in s.split(":")]
            if kv and kv[-1] == "":
                kv.pop()
            if len(kv) < 2:
                raise SVGError("style %r: a declaration without ':' (the reference reads past its list)" % v)
            if kv[0] == "stop-color":
                self.stop["color"] = _color(kv[1])
            elif kv[0] == "stop-opacity":
                self.stop["opacity"] = _to_float(kv[1], "stop-opacity")
            else:
                raise _Stop("style key %r" % kv[0])

    def read_attrs(self, node, extra=None, path_d=False):
        for name, v in node.attrs:
            if path_d and name == "d":
                _path_data(self.vg, v)
            elif extra and name in extra:
                extra[name](v)
            else:
                self.presentation(name, v)

    # elements (svg_load.cpp:1210-1520)
    def children(self, node):
        for c in node.children:
            f = self.readers.get(c.name)
            if f is not None:
                f(c)

    def finish_path(self, vbegin, all_fields=True):
        vg, a = self.vg, self.stack[-1]
        pi = vg.pathNumber() - 1
        if pi < 0:
            raise SVGError("a path element without path data comes first (the reference writes path -1)")
        vg._ftype[pi] = a["fill"]
        vg._rule[pi] = a["rule"]
        r, g, b, al = a["color"]
        al = _u8(F(al) * a["fopacity"])
        if self.a128:
            al = _u8((F(al) / F(255)) * F(127))
            al = _u8((F(al) / F(127)) * F(255))
        vg._color[pi] = rgba(r, g, b, al)
        vg._findex[pi] = a["findex"]
        vg._stype[pi] = a["stroke"]
        vg._scolor[pi] = rgba(*a["scolor"])
        vg._swidth[pi] = float(a["swidth"])
        vg._sindex[pi] = a["sindex"]
        m = a["mat"]
        self.pmat[pi] = (m, a["fopacity"], a["sopacity"])
        for k in range(vbegin, vg.vertexNumber()):
            x, y = _apply(m, *vg._v[k])
            vg._v[k] = (float(x), float(y))

    def read_path(self, node):
        vbegin = self.vg.vertexNumber()
        self.stack.append(_copy_attr(self.stack[-1]))
        self.read_attrs(node, path_d=True)
        self.finish_path(vbegin)
        self.stack.pop()

    def read_polygon(self, node):
        vbegin = self.vg.vertexNumber()
        self.stack.append(_copy_attr(self.stack[-1]))
        self.read_attrs(node, path_d=True)
        vg = self.vg
        vg.addPath(0, rgba(0, 0, 0, 255), FT_NONE, 0xFFFFFFFF)
        vg.addContour()
        for name, v in node.attrs:
            if name != "points":
                continue
            # istringstream >> float over the values with commas made spaces
            nums, pos, txt = [], 0, v.replace(",", " ")
            while True:
                m = _FNUM.match(txt, pos)
                if not m:
                    break
                nums.append(_f32(m.group(1)))
                pos = m.end()
            first = (nums[0] if nums else F(0), nums[1] if len(nums) > 1 else F(0))
            v0 = first
            for k in range(2, len(nums) - 1, 2):
                v1 = (nums[k], nums[k + 1])
                vg.addLinear(v0, v1)
                v0 = v1
            vg.addLinear(v0, first)
        self.finish_path(vbegin)
        self.stack.pop()

    def read_line(self, node):
        self.stack.append(_copy_attr(self.stack[-1]))
        raise _Stop("rect: the reference's reader table maps rect to its line reader, which throws")

    def read_rect(self, node):
        # circle and ellipse come here (svg_load.cpp:1203-1204)
        vbegin = self.vg.vertexNumber()
        self.stack.append(_copy_attr(self.stack[-1]))
        self.read_attrs(node, path_d=True)
        got = {}
        for name, v in node.attrs:
            if name in ("x", "y", "width", "height"):
                got[name] = _to_float(v, name)
        if len(got) != 4:
            raise SVGError("%s without x, y, width and height: the reference reads them uninitialised" % node.name)
        x, y, w, h = got["x"], got["y"], got["width"], got["height"]
        vg = self.vg
        vg.addPath(0, rgba(0, 0, 0, 255), FT_NONE, 0xFFFFFFFF)
        vg.addContour()
        v = [(x, y), (x, y + h), (x + w, y + h), (x + w, y)]
        for k in range(4):
            vg.addLinear(v[k], v[(k + 1) % 4])
        self.finish_path(vbegin)
        self.stack.pop()

    def read_group(self, node):
        self.stack.append(_copy_attr(self.stack[-1]))
        self.read_attrs(node)
        self.children(node)
        self.stack.pop()

    def _gradient(self, node, kind, keys):
        self.stack.append(_copy_attr(self.stack[-1]))
        g = self.gradient = dict(type=kind, href="", stops=[], transform=_identity())

        def setter(key):
            def f(v):
                g[key] = v if key == "href" else _to_float(v, key)
            return f

        extra = {k: setter(v) for k, v in keys.items()}

        def units(v):
            if v != "userSpaceOnUse":
                raise _Stop("gradientUnits %r" % v)
        extra["gradientUnits"] = units
        extra["gradientTransform"] = lambda v: g.__setitem__("transform", _parse_transform(v))
        extra["spreadMethod"] = lambda v: None
        self.read_attrs(node, extra)
        self.children(node)
        self.vg.svg_gradients.append(g)
        self.gmap[g["href"]] = len(self.vg.svg_gradients) - 1
        self.stack.pop()

    def read_linear(self, node):
        self._gradient(node, "linear", {"id": "href", "x1": "x1", "y1": "y1", "x2": "x2", "y2": "y2"})

    def read_radial(self, node):
        self._gradient(node, "radial", {"id": "href", "cx": "cx", "cy": "cy", "r": "r", "fx": "fx", "fy": "fy"})

    def read_stop(self, node):
        self.stack.append(_copy_attr(self.stack[-1]))
        st = self.stop
        extra = {"offset": lambda v: st.__setitem__("offset", _to_float(v, "offset")),
                 "stop-color": lambda v: st.__setitem__("color", _color(v)),
                 "stop-opacity": lambda v: st.__setitem__("opacity", _to_float(v, "stop-opacity"))}
        self.read_attrs(node, extra)
        r, g, b, a = st["color"]
        st["color"] = (r, g, b, _u8(F(a) * st["opacity"]))
        if self.gradient is not None:
            self.gradient["stops"].append(dict(st))
        self.stack.pop()

    def read_svg(self, node):
        self.stack.append(_new_attr() if not self.stack else _copy_attr(self.stack[-1]))
        for name, v in node.attrs:
            if name in ("width", "height"):
                if v.strip() and not _scan_floats(v, 1):
                    raise _Stop(name)
            elif name == "viewBox":
                vals = _scan_floats(v, 4)
                # sscanf stores the values it converted before a failure
                self.viewbox = tuple(vals) + self.viewbox[len(vals):]
                if len(vals) != 4:
                    raise _Stop("viewBox")
        self.children(node)
        self.stack.pop()

    def run(self, doc):
        self.readers = {"defs": self.children, "g": self.read_group, "path": self.read_path,
                        "rect": self.read_line, "polygon": self.read_polygon, "circle": self.read_rect,
                        "ellipse": self.read_rect, "linearGradient": self.read_linear,
                        "radialGradient": self.read_radial, "stop": self.read_stop}
        stopped = None
        try:
            for node in doc.children:
                if node.name == "svg":
                    self.read_svg(node)
        except _Stop as e:
            stopped = str(e)
        self.post()
        return stopped

    def post(self):
        """svg_load.cpp:1535-1549: viewBox size, origin shift and y flip."""
        vg = self.vg
        vx, vy, vw, vh = self.viewbox
        vg.width = int(np.trunc(vw)) if np.isfinite(vw) else 0
        vg.height = int(np.trunc(vh)) if np.isfinite(vh) else 0
        vg.viewbox = tuple(float(c) for c in self.viewbox)
        hf = F(vg.height & 0xFFFFFFFF)
        vg._v = [(float(F(x) - vx), float(hf - (F(y) - vy))) for x, y in vg._v]


def _affine(m):
    """Rows (a, b, c), (d, e, f) of a 3x3 column-major affine matrix, in float64."""
    return np.array([[m[0][0], m[1][0], m[2][0]], [m[0][1], m[1][1], m[2][1]], [0.0, 0.0, 1.0]], dtype=np.float64)


def _gradient_textures(L, vg, srgb):
    """Gradient table and ramp for the paths of `vg`, the loaded container or its
    stroke-to-fill conversion (see the module notes)."""
    g = GradientTextures(srgb)
    vx, vy = float(L.viewbox[0]), float(L.viewbox[1])
    hf = float(F(vg.height & 0xFFFFFFFF))
    final = np.array([[1.0, 0.0, -vx], [0.0, -1.0, hf + vy], [0.0, 0.0, 1.0]])
    source = getattr(vg, "source", None) or [(pi, "fill") for pi in range(vg.pathNumber())]
    for pi in range(vg.pathNumber()):
        orig, kind = source[pi]
        if vg._ftype[pi] != FT_GRADIENT or vg._stype[pi] == ST_COLOR or orig not in L.pmat:
            g.add_solid()
            continue
        gd = vg.svg_gradients[vg._findex[pi]]
        m, fop, sop = L.pmat[orig]
        opacity = sop if kind == "stroke" else fop
        tr = final @ _affine(m) @ _affine(gd["transform"])
        stops = [(float(st["offset"]),) + tuple(c / 255.0 for c in st["color"]) for st in gd["stops"]]
        words = [tr[0, 0], tr[0, 1], tr[0, 2], tr[1, 0], tr[1, 1], tr[1, 2]]
        val = lambda k, d=0.0: float(gd.get(k, d))  # noqa: E731
        if gd["type"] == "linear":
            g.add("linear", val("x1"), val("y1"), val("x2"), val("y2"), 0.0, stops, float(opacity), words)
        else:
            cx, cy = val("cx"), val("cy")
            g.add("radial", cx, cy, val("fx", cx), val("fy", cy), val("r"), stops, float(opacity), words)
    ramp, table = g.textures()
    vg.gradient_ramp = ramp.astype(np.int64).tolist()
    vg.gradient_table = table.astype(np.int64).tolist()


def loads(text, a128=False, gradients="render", on_error="raise", srgb=True, stroke_to_fill=False):
    """Parse SVG text into a VGContainer (see the module notes).  stroke_to_fill: turn every
    stroked path into fill geometry (VGContainer.strokeToFill) instead of filling it with
    the stroke colour."""
    if gradients not in ("render", "raise", "skip") or on_error not in ("raise", "partial"):
        raise ValueError("gradients: 'render', 'raise' or 'skip'; on_error: 'raise' or 'partial'")
    L = _Loader(a128, gradients)
    stopped = L.run(_parse_xml(text))
    vg = L.vg
    if stopped is not None and on_error == "raise":
        raise SVGError("the reference's SVG::load stops here: %s" % stopped, partial=vg)
    # a colour stroke replaces the fill word, so only gradient fills without one need a gradient
    if stroke_to_fill:
        vg = vg.strokeToFill()
    used = any(t == FT_GRADIENT and st != ST_COLOR for t, st in zip(vg._ftype, vg._stype))
    if gradients == "raise" and used:
        raise SVGError("gradient paints refused (gradients='raise')")
    if gradients == "render" and used:
        _gradient_textures(L, vg, srgb)
    return vg


def load(path, **kw):
    """Parse an SVG file (see `loads`)."""
    with open(path, "rb") as fh:
        return loads(fh.read(), **kw)
