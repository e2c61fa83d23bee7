"""RVG scene files (the reference's text scene format) into a VGContainer.

Follows `Mochimazui::RVG::load` (gpu-scanline/src/rvg.cpp:62-758):

* the file opens with viewport and window corners (rvg.cpp:154-162), then `scene
  dyn_identity`;
* each element is `1 element <rule> dyn_concrete <3 points> <ops...> <transform>
  dyn_paint <opacity> solid <colour>` (rvg.cpp:177-500).  Rules `nzfill` / `ofill` map
  to nonzero / even-odd; other rules and `depth` elements are skipped to the end of
  their line;
* ops: `M p`, `L p`, `Q p p`, `C p p p`, `A mx,my,w p` (rational quadratic; w = 0 is a
  half circle split in two arcs of weight sqrt(1/2), rvg.cpp:290-341), `Z`, `fL p`;
* a new path starts only when the previous element added vertices (rvg.cpp:199-203);
* the element's last contour is closed with a line if it is open and its ends differ
  (rvg.cpp:366-382); earlier contours are left as written;
* `dyn_affine[[a,b,c],[d,e,f]]` maps vertices by (a x + b y + c, d x + e y + f) / 1 in
  float (rvg.cpp:384-419);
* colours: `rgba(r,g,b,a)` / `rgb(r,g,b)` in [0, 1] truncated to 8 bits, alpha scaled
  by the opacity and truncated (rvg.cpp:437-481).  For `rgb(...)` the reference keeps
  only the first len - 6 characters after `rgb(`, dropping the last digit of blue; that
  is kept here too, so the same file gives the same colours.  `a128` requantises alpha
  to 7 bits as the reference's option does;
* comment tokens `//` skip their line;
* gradient paints, `linear` / `radial` followed by the rest of their line: every
  character that cannot belong to a number becomes a blank, the numbers before `pad` are
  x0 y0 x1 y1 [r] and the stops (offset r g b a), the six after it the paint transform
  (rvg.cpp:490-520, 547-580).  The path's fill becomes the gradient (fill word path + 1,
  colour (255, 0, 255, 0)), and the file's gradient table and ramp textures are built as
  the reference builds them (`gradient.GradientTextures`, with the ramp in sRGB when
  `srgb`, the reference's `--srgb` option, is on).  Where the reference would read a
  number that is not there, this loader raises.
"""

import numpy as np

import re

from .gradient import GradientTextures
from .scene import FT_COLOR, FT_GRADIENT, FT_NONE, VGContainer, rgba
from .vgfile import decimal_to_f32

_F32 = np.float32


class RVGError(ValueError):
    pass


def _tokens(text):
    """(token, line number) pairs."""
    out = []
    for ln, line in enumerate(text.splitlines()):
        for tok in line.split():
            out.append((tok, ln))
    return out


def _point(tok):
    s = tok.replace(",", " ").replace(":", " ").split()
    if len(s) < 2:
        raise RVGError("bad point %r" % tok)
    return (_F32(float(s[0])), _F32(float(s[1])))


def _u8(f):
    # (uint8_t)(f * 255) of a float: truncation toward zero, modulo 256 like the cast
    v = _F32(f) * _F32(255)
    return int(np.trunc(v)) & 0xFF


def _colour(tok):
    if tok.startswith("rgba"):
        body = tok[5:len(tok) - 1]
        vals = [float(x) for x in body.replace(",", " ").split()]
        if len(vals) != 4:
            raise RVGError("bad colour %r" % tok)
        return vals
    if tok.startswith("rgb"):
        body = tok[4:4 + max(len(tok) - 6, 0)]
        vals = []
        for x in body.replace(",", " ").split()[:3]:
            try:
                vals.append(float(x))
            except ValueError:
                vals.append(0.0)
        while len(vals) < 3:
            vals.append(0.0)
        return vals + [1.0]
    raise RVGError("unknown colour type %r" % tok)


def _affine(tok):
    t = tok[12:]
    for ch in "][]":
        i = t.find(ch)
        if i >= 0:
            t = t[:i] + t[i + 1:]
    t = t[:-1]
    vals = [float(x) for x in t.replace(",", " ").split()]
    if len(vals) != 6:
        raise RVGError("bad affine %r" % tok)
    return [_F32(v) for v in vals]


_NUM = re.compile(r"[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?")


def _gradient_numbers(rest):
    """The two number streams of a gradient paint's line: (before `pad`, after it), with
    every character that cannot belong to a number blanked first (rvg.cpp:498-509)."""
    ppad = rest.find("pad")
    if ppad < 0:
        raise RVGError("gradient paint without 'pad'")
    chars = list(rest)
    n = len(rest)
    for i, ch in enumerate(rest):
        keep = ("0" <= ch <= "9") or ch in "-.+" or (ch == "e" and i < n - 1 and rest[i + 1] in "+-")
        if not keep:
            chars[i] = " "
    s = "".join(chars)

    def numbers(t):
        out, i = [], 0
        while True:
            while i < len(t) and t[i] == " ":
                i += 1
            if i >= len(t):
                return out
            m = _NUM.match(t, i)
            if not m:
                return out  # the stream fails here: later reads get nothing
            out.append(m.group(0))
            i = m.end()

    return numbers(s[:ppad]), numbers(s[ppad:])


def loads(text, a128=False, gradients="render", srgb=True):
    """Parse RVG text; returns (VGContainer, viewport corners, window corners).

    gradients: "render" builds the gradient textures; "skip" leaves gradient-painted paths
    unpainted; "raise" refuses them.  srgb: the ramps are built for sRGB correction (the
    reference's --srgb), which must match the renderer's setting."""
    toks = _tokens(text)
    n = len(toks)
    pos = 0

    def nxt():
        nonlocal pos
        if pos >= n:
            raise RVGError("unexpected end of file")
        t = toks[pos][0]
        pos += 1
        return t

    def skip_line():
        nonlocal pos
        if pos == 0:
            return
        ln = toks[pos - 1][1]
        while pos < n and toks[pos][1] == ln:
            pos += 1

    vg = VGContainer()
    nxt()
    viewport = (_point(nxt()), _point(nxt()))
    nxt()
    window = (_point(nxt()), _point(nxt()))
    nxt()  # scene
    nxt()  # dyn_identity
    last_path_vertices = -1
    path = -1
    verts = vg._v  # vertex list of the container, appended in place below
    gtex = GradientTextures(srgb)
    n_grad = 0
    while pos < n:
        tok = nxt()
        if tok == "//":
            skip_line()
            continue
        if nxt() != "element":
            raise RVGError("expected 'element'")
        if last_path_vertices != len(verts):
            path = vg.addPath(0, rgba(0, 0, 0), FT_NONE)
            last_path_vertices = len(verts)
        rule = nxt()
        if rule == "nzfill":
            vg._rule[path] = 0
        elif rule == "ofill":
            vg._rule[path] = 1
        else:
            skip_line()
            continue
        tok = nxt()
        if tok == "depth":
            skip_line()
            continue
        if tok != "dyn_concrete":
            raise RVGError("expected 'dyn_concrete'")
        p0 = _point(nxt())
        _point(nxt())
        _point(nxt())
        path_begin = len(verts)
        contour_begin = path_begin
        closed = False
        tok = None
        while pos < n:
            tok = nxt()
            if tok != "fL" and len(tok) != 1:
                break
            op = tok[0]
            if tok != "fL" and op not in "MmZzLlHhVvCcSsQqTtAa":
                break
            if tok == "fL":
                p0 = _point(nxt())
            elif op == "M":
                p0 = _point(nxt())
                contour_begin = len(verts)
                closed = False
            elif op == "L":
                p1 = _point(nxt())
                vg.addLinear(p0, p1)
                p0 = p1
            elif op == "Q":
                p1, p2 = _point(nxt()), _point(nxt())
                vg.addQuadratic(p0, p1, p2)
                p0 = p2
            elif op == "C":
                p1, p2, p3 = _point(nxt()), _point(nxt()), _point(nxt())
                vg.addCubic(p0, p1, p2, p3)
                p0 = p3
            elif op == "A":
                s = nxt().replace(",", " ").replace(":", " ").split()
                mx, my = _F32(float(s[0])), _F32(float(s[1]))
                w1 = _F32(float(s[2])) if len(s) > 2 else _F32(0)
                if not (w1 >= 0 or w1 <= 0):
                    w1 = _F32(0)
                p1 = _point(nxt())
                if w1 == 0:
                    dx = _F32(0.5) * (p1[0] - p0[0])
                    dy = _F32(0.5) * (p1[1] - p0[1])
                    midx = _F32(0.5) * (p0[0] + p1[0]) + mx
                    midy = _F32(0.5) * (p0[1] + p1[1]) + my
                    w = _F32(0.7071067811865475244)
                    vg.addRational(p0, (midx - dx, midy - dy), (midx, midy), w)
                    vg.addRational((midx, midy), (midx + dx, midy + dy), p1, w)
                else:
                    vg.addRational(p0, (mx / w1, my / w1), p1, w1)
                p0 = p1
            elif op == "Z":
                closed = True
            else:
                raise RVGError("unsupported command %r" % tok)
        if len(verts) == path_begin:
            skip_line()
            continue
        if not closed and len(verts) > contour_begin:
            first, last = verts[contour_begin], verts[-1]
            if first != last:
                vg.addLinear(last, first)
        # transform of the element's vertices
        if tok is not None and tok.startswith("dyn_affine"):
            a, b, c, d, e, f = _affine(tok)
            for i in range(path_begin, len(verts)):
                x, y = _F32(verts[i][0]), _F32(verts[i][1])
                nx_ = (a * x + b * y) + c * _F32(1)
                ny_ = (d * x + e * y) + f * _F32(1)
                z = (_F32(0) * x + _F32(0) * y) + _F32(1) * _F32(1)
                verts[i] = (float(nx_ / z), float(ny_ / z))
        elif tok != "dyn_identity":
            raise RVGError("expected a transform, got %r" % tok)
        if nxt() != "dyn_paint":
            raise RVGError("expected 'dyn_paint'")
        opacity = _F32(float(nxt()))
        kind = nxt()
        if kind == "solid":
            gtex.add_solid()
            r, g, b, al = _colour(nxt())
            cr, cg, cb, ca = _u8(r), _u8(g), _u8(b), _u8(al)
            ca = int(np.trunc(_F32(ca) * opacity)) & 0xFF
            if a128:
                ca = int(np.trunc(_F32(ca) / _F32(255) * _F32(127))) & 0xFF
                ca = int(np.trunc(_F32(ca) / _F32(127) * _F32(255))) & 0xFF
            vg._color[path] = rgba(cr, cg, cb, ca)
            vg._ftype[path] = FT_COLOR
        elif kind in ("linear", "radial"):
            if gradients == "raise":
                raise RVGError("gradient paints are not supported")
            ln = toks[pos - 1][1]
            rest = []
            while pos < n and toks[pos][1] == ln:
                rest.append(toks[pos][0])
                pos += 1
            head, tail = _gradient_numbers(" " + " ".join(rest))
            ngeo = 5 if kind == "radial" else 4
            if len(head) < ngeo or len(tail) < 6 or (len(head) - ngeo) % 5:
                raise RVGError("gradient paint with missing numbers: %r" % " ".join(rest))
            h = decimal_to_f32(head) if head else []
            geo = [float(v) for v in h[:ngeo]]
            x0, y0, x1, y1 = geo[:4]
            rad = geo[4] if kind == "radial" else 0.0
            stops = [tuple(float(v) for v in h[k:k + 5]) for k in range(ngeo, len(h), 5)]
            tr = [float(v) for v in decimal_to_f32(tail[:6])]
            gtex.add(kind, x0, y0, x1, y1, rad, stops, opacity, tr)
            vg._ftype[path] = FT_GRADIENT
            vg._color[path] = rgba(255, 0, 255, 0)
            vg._findex[path] = n_grad
            n_grad += 1
            if gradients == "skip":
                vg._ftype[path] = FT_NONE
        else:
            raise RVGError("unknown paint %r" % kind)
    ramp, table = gtex.textures()
    vg.gradient_ramp = ramp.astype(np.int64).tolist()
    vg.gradient_table = table.astype(np.int64).tolist()
    return vg, viewport, window


def load(path, **kw):
    """Parse an RVG file (see `loads`)."""
    with open(path) as fh:
        return loads(fh.read(), **kw)
