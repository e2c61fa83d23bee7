"""Scene containers in the reference's structure-of-arrays layout.

``VGContainer`` mirrors the builder API of the reference's VGContainer
(gpu-scanline/src/vg_container.h:69-140: addPath / addContour / addLinear / addQuadratic /
addCubic / addRational, vertices stored per curve) and produces the arrays the C ABI's
``sl_scene`` expects.  ``generate`` wraps the library's seeded synthetic scenes.
"""
import ctypes

import numpy as np

from . import _abi


class SceneArrays:
    """Owns numpy arrays laid out as include/scanline.h's sl_scene."""

    def __init__(self, vertex, curve_vertex_pos, curve_type, curve_path_id, curve_arc_w,
                 fill_rule, fill_color, gradient_ramp=None, gradient_table=None):
        """fill_color holds the fill word of each path (ras_base.cpp:144-166): an RGBA8
        colour, or path + 1 for a gradient paint found in the gradient textures."""
        self.vertex = np.ascontiguousarray(vertex, dtype=np.float32).reshape(-1, 2)
        self.curve_vertex_pos = np.ascontiguousarray(curve_vertex_pos, dtype=np.uint32)
        self.curve_type = np.ascontiguousarray(curve_type, dtype=np.uint8)
        self.curve_path_id = np.ascontiguousarray(curve_path_id, dtype=np.uint32)
        self.curve_arc_w = np.ascontiguousarray(curve_arc_w, dtype=np.float32)
        self.fill_rule = np.ascontiguousarray(fill_rule, dtype=np.uint8)
        self.fill_color = np.ascontiguousarray(fill_color, dtype=np.uint32)
        # gradient textures (vg_container.h:344-345): RGBA8 ramp texels, 1024 a row, and
        # 12 words of gradient table per path slot
        self.gradient_ramp = np.ascontiguousarray(
            np.zeros(0) if gradient_ramp is None else np.asarray(gradient_ramp, dtype=np.int64) & 0xFFFFFFFF,
            dtype=np.uint32).view(np.int32)
        self.gradient_table = np.ascontiguousarray(
            np.zeros(0) if gradient_table is None else np.asarray(gradient_table, dtype=np.int64) & 0xFFFFFFFF,
            dtype=np.uint32).view(np.int32)

    @property
    def n_paths(self):
        return int(self.fill_rule.shape[0])

    @property
    def n_curves(self):
        return int(self.curve_type.shape[0])

    @property
    def n_vertices(self):
        return int(self.vertex.shape[0])

    def as_c(self):
        """sl_scene struct pointing into this object's arrays (keep self alive)."""
        s = _abi.SlScene()
        s.n_paths, s.n_curves, s.n_vertices = self.n_paths, self.n_curves, self.n_vertices
        s.vertex = self.vertex.ctypes.data
        s.curve_vertex_pos = self.curve_vertex_pos.ctypes.data
        s.curve_type = self.curve_type.ctypes.data
        s.curve_path_id = self.curve_path_id.ctypes.data
        s.curve_arc_w = self.curve_arc_w.ctypes.data
        s.fill_rule = self.fill_rule.ctypes.data
        s.fill_color = self.fill_color.ctypes.data
        s.gradient_ramp = self.gradient_ramp.ctypes.data if self.gradient_ramp.size else None
        s.gradient_ramp_size = self.gradient_ramp.size
        s.gradient_table = self.gradient_table.ctypes.data if self.gradient_table.size else None
        s.gradient_table_size = self.gradient_table.size
        return s

    def oracle_args(self):
        """Positional arrays in the order of the oracle's oracle_run."""
        return (self.n_paths, self.n_curves, self.n_vertices, self.vertex, self.curve_vertex_pos,
                self.curve_type, self.curve_path_id, self.curve_arc_w, self.fill_rule,
                self.fill_color, self.gradient_ramp, self.gradient_table)


# fill and stroke types of a path (VGFillType / VGStrokeType, vg_container.h:19-29)
FT_NONE, FT_COLOR, FT_GRADIENT = 0, 1, 4
ST_NONE, ST_COLOR, ST_GRADIENT = 0, 1, 4


def rgba(r, g, b, a=255):
    """u8rgba packed little-endian (gpu-scanline/src/mochimazui/color.h:29-57)."""
    return (int(r) & 255) | ((int(g) & 255) << 8) | ((int(b) & 255) << 16) | ((int(a) & 255) << 24)


class VGContainer:
    """Host-side scene builder with the reference's VGContainer method names."""

    def __init__(self, width=0, height=0, x=0, y=0):
        self.x, self.y, self.width, self.height = x, y, width, height
        self._v, self._cpos, self._ctype, self._cpath, self._cw = [], [], [], [], []
        self._rule, self._color, self._ftype, self._findex = [], [], [], []
        self._stype, self._scolor = [], []  # stroke type and colour per path
        self._swidth, self._sindex = [], []  # stroke width and gradient index per path
        self._cfirst, self._cowner = [], []  # contours: first curve, owning path
        self.gradient_ramp, self.gradient_table = [], []  # kept for file round trips

    def pathNumber(self):
        return len(self._rule)

    def curveNumber(self):
        return len(self._ctype)

    def vertexNumber(self):
        return len(self._v)

    def contourNumber(self):
        return len(self._cfirst)

    def addPath(self, f_rule=0, f_color=rgba(0, 0, 0), f_mode=FT_COLOR, f_index=0xFFFFFFFF):
        """addPath(f_rule, f_mode, f_color, f_index) (vg_container.cpp:310-320); the paint
        arguments come last here so that colour fills read addPath(rule, colour)."""
        self._rule.append(int(f_rule))
        self._color.append(int(f_color))
        self._ftype.append(int(f_mode))
        self._findex.append(int(f_index) & 0xFFFFFFFF)
        self._stype.append(ST_NONE)
        self._scolor.append(rgba(0, 0, 0))
        self._swidth.append(1.0)
        self._sindex.append(0xFFFFFFFF)
        return len(self._rule) - 1

    def addStrokePath(self, s_color, width):
        """addStrokePath (vg_container.cpp:322-332): a path with no fill and a colour stroke
        of the given width."""
        pid = self.addPath(0, rgba(0, 0, 0), FT_NONE)
        self._stype[pid] = ST_COLOR
        self._scolor[pid] = int(s_color)
        self._swidth[pid] = float(width)
        return pid

    def contours(self):
        """Curve index ranges [first, last) of each contour, with its path."""
        ends = self._cfirst[1:] + [len(self._ctype)]
        return [(f, e, o) for f, e, o in zip(self._cfirst, ends, self._cowner)]

    def strokeToFill(self):
        """strokeToOffsetCurve (vg_container.cpp:555-1353): a container in which every stroked
        path is followed by a nonzero fill path of its stroke outline; see scanline.stroke."""
        from . import stroke
        return stroke.stroke_to_fill(self)

    def addContour(self):
        """addContour (vg_container.cpp:269-280): the next curves form a new contour of the
        last path.  Contours only group curves (the fill does not use them); a curve added
        to a path with no contour opens one."""
        if not self._rule:
            raise ValueError("addPath before addContour")
        self._cfirst.append(len(self._ctype))
        self._cowner.append(len(self._rule) - 1)
        return len(self._cfirst) - 1

    def _curve(self, ctype, pts, w=-1.0):
        if not self._rule:
            raise ValueError("addPath before adding curves")
        if not self._cowner or self._cowner[-1] != len(self._rule) - 1:
            self.addContour()
        self._cpos.append(len(self._v))
        self._ctype.append(ctype)
        self._cpath.append(len(self._rule) - 1)
        self._cw.append(w)
        self._v.extend((float(x), float(y)) for x, y in pts)
        return len(self._ctype) - 1

    def addLinear(self, v0, v1):
        return self._curve(_abi.CURVE_LINEAR, (v0, v1))

    def addQuadratic(self, v0, v1, v2):
        return self._curve(_abi.CURVE_QUADRATIC, (v0, v1, v2))

    def addCubic(self, v0, v1, v2, v3):
        return self._curve(_abi.CURVE_CUBIC, (v0, v1, v2, v3))

    def addRational(self, v0, v1, v2, w):
        return self._curve(_abi.CURVE_RATIONAL, (v0, v1, v2), float(w))

    def addPolygon(self, pts):
        """Closed polygon of lines (a convenience for tests)."""
        for k in range(len(pts)):
            self.addLinear(pts[k], pts[(k + 1) % len(pts)])

    def fillInfo(self):
        """Per-path fill word as the rasterizer's input gathers it (ras_base.cpp:144-166):
        colour fills keep their colour unless its alpha is 0, a gradient fill gives the
        path's index + 1 (the renderer finds its gradient in the gradient table), every
        other fill type gives 0; a path with a colour stroke is filled with the stroke colour
        instead, since the reference's renderer takes its input without stroke-to-fill
        conversion."""
        out = []
        for p, (c, t, sc, st) in enumerate(zip(self._color, self._ftype, self._scolor, self._stype)):
            if t == FT_COLOR:
                w = c if (c & 0xFF000000) else 0
            elif t == FT_GRADIENT:
                w = p + 1
            else:
                w = 0
            if st == ST_COLOR:
                w = sc if sc & 0xFF000000 else 0
            out.append(w)
        return out

    def arrays(self):
        v = np.array(self._v, dtype=np.float32).reshape(-1, 2)
        return SceneArrays(v, self._cpos, self._ctype, self._cpath, self._cw, self._rule,
                           self.fillInfo(), self.gradient_ramp, self.gradient_table)

    @classmethod
    def from_arrays(cls, a, width=0, height=0):
        """A container holding the scene arrays `a` (e.g. a generated scene), one contour
        per path, colour fills; curves must be grouped by path in path order."""
        pid = np.asarray(a.curve_path_id, dtype=np.int64)
        if len(pid) and (np.diff(pid) < 0).any():
            raise ValueError("curves are not grouped by path")
        vg = cls(width, height)
        vg._v = [tuple(p) for p in np.asarray(a.vertex, dtype=np.float32).astype(np.float64).tolist()]
        vg._cpos = np.asarray(a.curve_vertex_pos, dtype=np.int64).tolist()
        vg._ctype = np.asarray(a.curve_type, dtype=np.int64).tolist()
        vg._cpath = pid.tolist()
        vg._cw = np.asarray(a.curve_arc_w, dtype=np.float32).astype(np.float64).tolist()
        vg._rule = np.asarray(a.fill_rule, dtype=np.int64).tolist()
        vg._color = np.asarray(a.fill_color, dtype=np.int64).tolist()
        vg._ftype = [FT_COLOR] * len(vg._rule)
        vg._findex = [0xFFFFFFFF] * len(vg._rule)
        vg._stype = [ST_NONE] * len(vg._rule)
        vg._scolor = [rgba(0, 0, 0)] * len(vg._rule)
        vg._swidth = [1.0] * len(vg._rule)
        vg._sindex = [0xFFFFFFFF] * len(vg._rule)
        owners, first = np.unique(pid, return_index=True)
        vg._cfirst, vg._cowner = first.tolist(), owners.tolist()
        return vg

    def save(self, filename):
        """VGContainer::save (vg_container.cpp:1667-1677): `.vgt` text or `.vgb` binary."""
        from . import vgfile
        vgfile.save(self, filename)

    @staticmethod
    def load(filename, gradients="render"):
        """VGContainer::load (vg_container.cpp:1679-1689) as a constructor: `.vgt` or `.vgb`."""
        from . import vgfile
        return vgfile.load(filename, gradients=gradients)


KIND_CUBIC_CHAINS, KIND_MIXED, KIND_GLYPHS = 1, 2, 3


def generate(kind, width, height, n_paths, seed, min_segs=0, max_segs=0):
    """Seeded synthetic scene from the library generator, copied into numpy arrays."""
    L = _abi.lib()
    obj = ctypes.c_void_p()
    rc = L.sl_scene_generate(kind, width, height, n_paths, ctypes.c_uint64(seed), min_segs,
                             max_segs, ctypes.byref(obj))
    if rc:
        raise _abi.ScanlineError(rc, "sl_scene_generate failed")
    try:
        s = _abi.SlScene()
        L.sl_scene_view(obj, ctypes.byref(s))

        def arr(ptr, n, ct, dt):
            if n == 0:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ct)), (n,)).astype(dt, copy=True)

        V, C, P = s.n_vertices, s.n_curves, s.n_paths
        out = SceneArrays(arr(s.vertex, 2 * V, ctypes.c_float, np.float32),
                          arr(s.curve_vertex_pos, C, ctypes.c_uint32, np.uint32),
                          arr(s.curve_type, C, ctypes.c_uint8, np.uint8),
                          arr(s.curve_path_id, C, ctypes.c_uint32, np.uint32),
                          arr(s.curve_arc_w, C, ctypes.c_float, np.float32),
                          arr(s.fill_rule, P, ctypes.c_uint8, np.uint8),
                          arr(s.fill_color, P, ctypes.c_uint32, np.uint32))
    finally:
        L.sl_scene_free(obj)
    return out


SEED_BASE = 0x5CA1AB1E


def config_scene(config_id, scale=1.0):
    """The benchmark workloads' scenes (kind, canvas, paths, spp) for configs 1-3 and 5.

    ``scale`` < 1 shrinks path counts for quick parity runs at the same density rules.
    """
    seed = SEED_BASE ^ config_id
    if config_id == 1:
        return generate(KIND_CUBIC_CHAINS, 512, 512, max(1, int(200 * scale)), seed), 512, 512, 8
    if config_id == 2:
        return generate(KIND_MIXED, 2048, 2048, max(1, int(20000 * scale)), seed), 2048, 2048, 32
    if config_id == 3:
        return generate(KIND_GLYPHS, 4096, 4096, max(1, int(100000 * scale)), seed), 4096, 4096, 16
    if config_id == 5:
        return generate(KIND_CUBIC_CHAINS, 256, 256, max(1, int(64 * scale)), seed, 6, 6), 256, 256, 16
    raise ValueError("no single-scene workload for config %d" % config_id)
