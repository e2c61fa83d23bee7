"""The reference's own scene files: `.vgt` (text) and `.vgb` (binary) VGContainer dumps.

Both formats serialise the container's arrays in one fixed order
(VGContainer::saveTxt / loadTxt, gpu-scanline/src/vg_container.cpp:1692-1793, and
saveBinary / loadBinary, :1796-1906):

1. viewport x, y, width, height;
2. counts: paths, contours, curves, vertices, gradient ramp size, gradient table size;
3. per path: first contour, contour count, fill type (u8), fill rule (u8), fill colour
   (u8rgba, i.e. r | g << 8 | b << 16 | a << 24), fill index;
4. per contour: first curve, curve count;
5. per curve: first vertex, curve type (u8), arc weight (the float's bits as an int;
   0x7f7fffff for non-rational curves, vg_container.cpp:94);
6. vertices as float x, y;
7. the gradient ramp and table textures as ints.

Text (`.vgt`): one record a line, fields separated by commas, printed with
`%d` / `%u`, the u8 fields with `%c` (the raw byte, so fill rule 0 is a NUL byte in the
file) and vertices with `%e` (six digits after the point: text files round vertices to
seven significant digits).  Reading follows fscanf: whitespace before a number is
skipped, a comma must follow at once, `%c` takes the next byte whatever it is, and
decimal vertices are rounded once, to the nearest float.

Binary (`.vgb`): the same fields as little-endian u32 / u8 / i32 / f32 arrays, each
field of all paths (curves, ...) stored together, with no padding.

Loading gathers curves the way the rasterizer's input does (ras_base.cpp:179-244): paths
in order, each path's contours in order, each contour's curves in order; the vertex
array is kept whole and curves index into it.  Unlike the reference, which reads
whatever the file holds, a short file or an index out of range raises `VGFileError`;
an unknown suffix raises too (the reference prints a message and does nothing).
"""
import os
import re
import struct
from fractions import Fraction

import numpy as np

from . import _abi
from .scene import FT_GRADIENT, ST_NONE, VGContainer, rgba

NONRATIONAL_W = 0x7F7FFFFF  # arc weight word of lines, quadratics and cubics
_NVERT = {_abi.CURVE_LINEAR: 2, _abi.CURVE_QUADRATIC: 3, _abi.CURVE_CUBIC: 4, _abi.CURVE_RATIONAL: 3}


class VGFileError(ValueError):
    pass


# ---------------------------------------------------------------------------------
# container <-> arrays
# ---------------------------------------------------------------------------------
def _ranges(starts, counts):
    """Concatenation of arange(s, s + c) over the pairs (s, c)."""
    starts = np.asarray(starts, dtype=np.int64)
    counts = np.asarray(counts, dtype=np.int64)
    total = int(counts.sum())
    if total == 0:
        return np.zeros(0, dtype=np.int64)
    ends = np.cumsum(counts)
    first = ends - counts
    idx = np.arange(total, dtype=np.int64)
    return np.repeat(starts - first, counts) + idx


def _fields(vg):
    """The container's file fields as arrays (contours rebuilt from the curve order)."""
    n_path = vg.pathNumber()
    cfirst = np.asarray(vg._cfirst, dtype=np.int64)
    cowner = np.asarray(vg._cowner, dtype=np.int64)
    n_contour = len(cfirst)
    cnext = np.append(cfirst[1:], vg.curveNumber()).astype(np.int64)[:n_contour]
    ccount = cnext - cfirst
    pcount = np.bincount(cowner, minlength=n_path) if n_contour else np.zeros(n_path, dtype=np.int64)
    pfirst = np.cumsum(pcount) - pcount
    ctype = np.asarray(vg._ctype, dtype=np.uint8)
    w = np.asarray(vg._cw, dtype=np.float32)
    w1s = np.where(ctype == _abi.CURVE_RATIONAL, w.view(np.int32), np.int32(NONRATIONAL_W)).astype(np.int32)
    return dict(
        view=(int(vg.x), int(vg.y), int(vg.width), int(vg.height)),
        p_cindex=pfirst.astype(np.uint32), p_cnumber=pcount.astype(np.uint32),
        p_ftype=np.asarray(vg._ftype, dtype=np.uint8), p_frule=np.asarray(vg._rule, dtype=np.uint8),
        p_color=np.asarray(vg._color, dtype=np.uint64).astype(np.uint32),
        p_findex=np.asarray(vg._findex, dtype=np.uint64).astype(np.uint32),
        c_index=cfirst.astype(np.uint32), c_number=ccount.astype(np.uint32),
        v_index=np.asarray(vg._cpos, dtype=np.uint32), c_type=ctype, c_w1s=w1s,
        vertex=np.asarray(vg._v, dtype=np.float32).reshape(-1, 2),
        ramp=np.asarray(vg.gradient_ramp, dtype=np.int64).astype(np.int32),
        table=np.asarray(vg.gradient_table, dtype=np.int64).astype(np.int32),
    )


def _container(f, gradients):
    """VGContainer of file fields, curves gathered path by path and contour by contour."""
    n_path, n_contour = len(f["p_frule"]), len(f["c_index"])
    n_curve, n_vertex = len(f["c_type"]), len(f["vertex"])
    pci = f["p_cindex"].astype(np.int64)
    pcn = f["p_cnumber"].astype(np.int64)
    if n_path and (pci + pcn > n_contour).any():
        raise VGFileError("a path's contours lie beyond the contour array")
    cons = _ranges(pci, pcn)
    con_owner = np.repeat(np.arange(n_path, dtype=np.int64), pcn)
    cci = f["c_index"].astype(np.int64)[cons]
    ccn = f["c_number"].astype(np.int64)[cons]
    if len(cons) and (cci + ccn > n_curve).any():
        raise VGFileError("a contour's curves lie beyond the curve array")
    curves = _ranges(cci, ccn)
    ctype = f["c_type"][curves]
    known = np.isin(ctype, list(_NVERT))
    if not known.all():
        raise VGFileError("unknown curve type 0x%02x" % int(ctype[~known][0]))
    vpos = f["v_index"][curves].astype(np.int64)
    nv = np.vectorize(_NVERT.get, otypes=[np.int64])(ctype) if len(ctype) else np.zeros(0, np.int64)
    if len(vpos) and (vpos + nv > n_vertex).any():
        raise VGFileError("a curve's vertices lie beyond the vertex array")
    ftype = f["p_ftype"]
    if gradients == "raise" and (ftype == FT_GRADIENT).any():
        raise VGFileError("gradient paints are not supported")

    x, y, w, h = f["view"]
    vg = VGContainer(w, h, x, y)
    vg._v = [tuple(p) for p in f["vertex"].astype(np.float64).tolist()]
    vg._rule = f["p_frule"].astype(int).tolist()
    vg._color = f["p_color"].astype(np.int64).tolist()
    vg._ftype = ftype.astype(int).tolist()
    vg._findex = f["p_findex"].astype(np.int64).tolist()
    vg._stype = [ST_NONE] * n_path  # the files hold no stroke fields
    vg._scolor = [rgba(0, 0, 0)] * n_path
    vg._swidth = [1.0] * n_path
    vg._sindex = [0xFFFFFFFF] * n_path
    vg._cpos = vpos.tolist()
    vg._ctype = ctype.astype(int).tolist()
    vg._cpath = np.repeat(con_owner, ccn).tolist()
    wbits = f["c_w1s"][curves].astype(np.int32)
    vg._cw = np.where(ctype == _abi.CURVE_RATIONAL, wbits.view(np.float32), np.float32(-1)).astype(
        np.float64).tolist()
    vg._cfirst = (np.cumsum(ccn) - ccn).tolist()
    vg._cowner = con_owner.tolist()
    vg.gradient_ramp = f["ramp"].astype(int).tolist()
    vg.gradient_table = f["table"].astype(int).tolist()
    return vg


# ---------------------------------------------------------------------------------
# binary
# ---------------------------------------------------------------------------------
def save_binary(vg, filename):
    """VGContainer::saveBinary (vg_container.cpp:1796-1850)."""
    f = _fields(vg)
    counts = (len(f["p_frule"]), len(f["c_index"]), len(f["c_type"]), len(f["vertex"]),
              len(f["ramp"]), len(f["table"]))
    with open(filename, "wb") as out:
        out.write(struct.pack("<4i", *f["view"]))
        out.write(struct.pack("<6I", *counts))
        for key, dt in (("p_cindex", "<u4"), ("p_cnumber", "<u4"), ("p_ftype", "u1"), ("p_frule", "u1"),
                        ("p_color", "<u4"), ("p_findex", "<u4"), ("c_index", "<u4"), ("c_number", "<u4"),
                        ("v_index", "<u4"), ("c_type", "u1"), ("c_w1s", "<i4"), ("vertex", "<f4"),
                        ("ramp", "<i4"), ("table", "<i4")):
            out.write(np.ascontiguousarray(f[key], dtype=dt).tobytes())


def _parse_binary(data):
    pos = 0

    def take(dt, n):
        nonlocal pos
        size = np.dtype(dt).itemsize * n
        if pos + size > len(data):
            raise VGFileError("file ends inside its arrays")
        a = np.frombuffer(data, dtype=dt, count=n, offset=pos)
        pos += size
        return a

    if len(data) < 40:
        raise VGFileError("file shorter than its header")
    view = struct.unpack_from("<4i", data, 0)
    n_path, n_contour, n_curve, n_vertex, n_ramp, n_table = struct.unpack_from("<6I", data, 16)
    pos = 40
    f = dict(view=view)
    f["p_cindex"], f["p_cnumber"] = take("<u4", n_path), take("<u4", n_path)
    f["p_ftype"], f["p_frule"] = take("u1", n_path), take("u1", n_path)
    f["p_color"], f["p_findex"] = take("<u4", n_path), take("<u4", n_path)
    f["c_index"], f["c_number"] = take("<u4", n_contour), take("<u4", n_contour)
    f["v_index"], f["c_type"], f["c_w1s"] = take("<u4", n_curve), take("u1", n_curve), take("<i4", n_curve)
    f["vertex"] = take("<f4", 2 * n_vertex).reshape(-1, 2)
    f["ramp"], f["table"] = take("<i4", n_ramp), take("<i4", n_table)
    return f


def load_binary(filename, gradients="render"):
    """VGContainer::loadBinary (vg_container.cpp:1852-1906)."""
    with open(filename, "rb") as fh:
        return _container(_parse_binary(fh.read()), gradients)


# ---------------------------------------------------------------------------------
# text
# ---------------------------------------------------------------------------------
def _e(v):
    """printf("%e") of a float (promoted to double, as printf does)."""
    return ("%e" % float(v)).encode()


def save_text(vg, filename):
    """VGContainer::saveTxt (vg_container.cpp:1692-1742)."""
    f = _fields(vg)
    n_path, n_contour, n_curve = len(f["p_frule"]), len(f["c_index"]), len(f["c_type"])
    parts = [b"%d,%d,%d,%d\n" % f["view"],
             b"%u,%u,%u,%u,%u,%u\n" % (n_path, n_contour, n_curve, len(f["vertex"]), len(f["ramp"]),
                                        len(f["table"]))]
    for i in range(n_path):
        parts.append(b"%u,%u,%c,%c,%u,%u\n" % (int(f["p_cindex"][i]), int(f["p_cnumber"][i]), int(f["p_ftype"][i]),
                                               int(f["p_frule"][i]), int(f["p_color"][i]), int(f["p_findex"][i])))
    for i in range(n_contour):
        parts.append(b"%u,%u\n" % (int(f["c_index"][i]), int(f["c_number"][i])))
    for i in range(n_curve):
        parts.append(b"%u,%c,%d\n" % (int(f["v_index"][i]), int(f["c_type"][i]), int(f["c_w1s"][i])))
    for x, y in f["vertex"]:
        parts.append(_e(x) + b"," + _e(y) + b"\n")
    for v in f["ramp"]:
        parts.append(b"%d\n" % int(v))
    for v in f["table"]:
        parts.append(b"%d\n" % int(v))
    with open(filename, "wb") as out:
        out.write(b"".join(parts))


_WS = re.compile(rb"[ \t\n\v\f\r]*")
_INT = re.compile(rb"[+-]?[0-9]+")
_FLOAT = re.compile(rb"[+-]?(?:(?:[0-9]+\.?[0-9]*|\.[0-9]+)(?:[eE][+-]?[0-9]+)?|inf(?:inity)?|nan)", re.I)


class _Scanner:
    """The subset of fscanf that the text reader's formats use."""

    def __init__(self, data):
        self.buf, self.p = data, 0

    def ws(self):
        self.p = _WS.match(self.buf, self.p).end()

    def _token(self, rx, what):
        self.ws()
        m = rx.match(self.buf, self.p)
        if not m:
            raise VGFileError("expected %s at byte %d" % (what, self.p))
        self.p = m.end()
        return m.group(0)

    def u(self):
        # %u: optional sign, the value taken modulo 2^32
        return int(self._token(_INT, "an integer")) & 0xFFFFFFFF

    def d(self):
        v = int(self._token(_INT, "an integer")) & 0xFFFFFFFF
        return v - (1 << 32) if v & 0x80000000 else v

    def e(self):
        return self._token(_FLOAT, "a number")

    def lit(self, ch):
        if self.buf[self.p:self.p + 1] != ch:
            raise VGFileError("expected %r at byte %d" % (ch, self.p))
        self.p += 1

    def c(self):
        if self.p >= len(self.buf):
            raise VGFileError("file ends inside a record")
        v = self.buf[self.p]
        self.p += 1
        return v


def decimal_to_f32(tokens):
    """Decimal strings rounded once to the nearest float32 (ties to even), as fscanf %e
    into a float does.  Going through a double first can only go wrong when the double
    lands exactly halfway between two floats; those few are settled exactly."""
    d = np.array([float(t) for t in tokens], dtype=np.float64)
    f = d.astype(np.float32)
    fd = f.astype(np.float64)
    with np.errstate(invalid="ignore", over="ignore"):
        other = np.nextafter(f, np.where(d > fd, np.float32(np.inf), np.float32(-np.inf))).astype(np.float64)
        half = np.isfinite(d) & np.isfinite(other) & (d != fd) & ((fd + other) / 2 == d)
    for i in np.flatnonzero(half):
        t = tokens[i].decode() if isinstance(tokens[i], bytes) else tokens[i]
        exact, mid = Fraction(t), Fraction(float(d[i]))
        if exact != mid:  # not a true tie: the side of the decimal decides
            f[i] = np.float32(other[i]) if (exact > mid) == (other[i] > fd[i]) else f[i]
    return f


def _parse_text(data):
    s = _Scanner(data)
    view = []
    for k in range(4):
        if k:
            s.lit(b",")
        view.append(s.d())
    counts = []
    for k in range(6):
        if k:
            s.lit(b",")
        counts.append(s.u())
    n_path, n_contour, n_curve, n_vertex, n_ramp, n_table = counts
    f = dict(view=tuple(view))
    cols = [[] for _ in range(6)]
    for _ in range(n_path):
        cols[0].append(s.u()); s.lit(b",")
        cols[1].append(s.u()); s.lit(b",")
        cols[2].append(s.c()); s.lit(b",")
        cols[3].append(s.c()); s.lit(b",")
        cols[4].append(s.u()); s.lit(b",")
        cols[5].append(s.u())
    for key, col, dt in zip(("p_cindex", "p_cnumber", "p_ftype", "p_frule", "p_color", "p_findex"), cols,
                            ("u4", "u4", "u1", "u1", "u4", "u4")):
        f[key] = np.array(col, dtype=np.int64).astype(dt)
    ci, cn = [], []
    for _ in range(n_contour):
        ci.append(s.u()); s.lit(b",")
        cn.append(s.u())
    f["c_index"], f["c_number"] = np.array(ci, dtype=np.uint32), np.array(cn, dtype=np.uint32)
    vi, ct, cw = [], [], []
    for _ in range(n_curve):
        vi.append(s.u()); s.lit(b",")
        ct.append(s.c()); s.lit(b",")
        cw.append(s.d())
    f["v_index"] = np.array(vi, dtype=np.uint32)
    f["c_type"] = np.array(ct, dtype=np.uint8)
    f["c_w1s"] = np.array(cw, dtype=np.int64).astype(np.int32)
    toks = []
    for _ in range(n_vertex):
        toks.append(s.e()); s.lit(b",")
        toks.append(s.e())
    f["vertex"] = decimal_to_f32(toks).reshape(-1, 2)
    f["ramp"] = np.array([s.d() for _ in range(n_ramp)], dtype=np.int64).astype(np.int32)
    f["table"] = np.array([s.d() for _ in range(n_table)], dtype=np.int64).astype(np.int32)
    return f


def load_text(filename, gradients="render"):
    """VGContainer::loadTxt (vg_container.cpp:1744-1793)."""
    with open(filename, "rb") as fh:
        return _container(_parse_text(fh.read()), gradients)


# ---------------------------------------------------------------------------------
def _suffix(filename):
    return os.fspath(filename)[-3:]


def save(vg, filename):
    """VGContainer::save: the suffix picks the format."""
    sfx = _suffix(filename)
    if sfx == "vgt":
        save_text(vg, filename)
    elif sfx == "vgb":
        save_binary(vg, filename)
    else:
        raise VGFileError("unknown scene file suffix %r (vgt or vgb)" % sfx)


def load(filename, gradients="render"):
    """VGContainer::load: the suffix picks the format.  The file's gradient ramp and table
    textures travel with the container, so gradient-filled paths render ("render" or
    "skip", which keep them as they are); "raise" refuses them."""
    if gradients not in ("render", "raise", "skip"):
        raise ValueError("gradients must be 'render', 'raise' or 'skip'")
    sfx = _suffix(filename)
    if sfx == "vgt":
        return load_text(filename, gradients)
    if sfx == "vgb":
        return load_binary(filename, gradients)
    raise VGFileError("unknown scene file suffix %r (vgt or vgb)" % sfx)
