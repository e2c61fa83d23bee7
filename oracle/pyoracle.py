"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, the smoke check and the benchmark's CPU-baseline leg -- as the checker
or the timed CPU baseline, never as the renderer.  ``build()`` compiles it with the
Makefile beside this file.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_SERIAL = os.path.join(HERE, "build", "liboracle.so")
LIB_OMP = os.path.join(HERE, "build", "liboracle_omp.so")

_DT = {
    "tvertex": (np.float32, 2), "path_flags": (np.uint8, 8), "cuts": (np.float32, 4),
    "ncuts": (np.uint8, 1), "frag_count": (np.uint32, 1), "ts": (np.int32, 2),
    "frag_yx": (np.int32, 1), "frag_pid": (np.int32, 1), "frag_wn": (np.int8, 1),
    "frag_mask": (np.uint32, 4), "sorted_yx": (np.int32, 1), "sorted_idx": (np.int32, 1),
    "wn_scan": (np.int8, 1), "merged_mask": (np.uint32, 4), "draw_list": (np.int32, 4),
    "image": (np.uint8, 4), "counts": (np.int64, 1), "pixel_count": (np.int64, 1),
}


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


_libs = {}


def lib(omp=False):
    # SCANLINE_ORACLE_LIB names another build of the serial oracle (the mutation tests
    # point it at copies with a planted mistake)
    path = LIB_OMP if omp else (os.environ.get("SCANLINE_ORACLE_LIB") or LIB_SERIAL)
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    vp, u32, i = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
    L.oracle_run.restype = vp
    L.oracle_run.argtypes = [u32, u32, u32, vp, vp, vp, vp, vp, vp, vp, i, i, i, i, i, vp, i, vp, u32, vp, u32]
    L.oracle_get.restype = i
    L.oracle_get.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_size_t)]
    L.oracle_render_seconds.restype = ctypes.c_double
    L.oracle_render_seconds.argtypes = [vp]
    L.oracle_free.argtypes = [vp]
    L.oracle_threads.restype = i
    L.oracle_sample_positions.argtypes = [i, vp]
    L.oracle_pixel_table.argtypes = [i, vp]
    L.oracle_fetch_pixel.argtypes = [i, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, vp]
    L.oracle_fetch_pixel_batch.argtypes = [i, ctypes.c_int64, vp, vp]
    L.oracle_vertical_mask.argtypes = [i, ctypes.c_float, vp]
    L.oracle_vertical_mask2.argtypes = [i, ctypes.c_float, ctypes.c_float, vp]
    L.oracle_solve_quad.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                    ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    L.oracle_curve_eval.restype = ctypes.c_float
    L.oracle_curve_eval.argtypes = [i, ctypes.c_float, vp, ctypes.c_float]
    L.oracle_color_tables.argtypes = [i, vp, vp]
    if hasattr(L, "oracle_chord_diagram"):
        f = ctypes.c_float
        L.oracle_chord_sincos.argtypes = [f, ctypes.POINTER(f), ctypes.POINTER(f)]
        L.oracle_chord_ease.restype = f
        L.oracle_chord_ease.argtypes = [i]
        L.oracle_chord_paths.restype = ctypes.c_int64
        L.oracle_chord_paths.argtypes = [i, i, vp, i]
        L.oracle_chord_diagram.restype = ctypes.c_int64
        L.oracle_chord_diagram.argtypes = [i, i, vp, vp, vp, vp, vp, vp, vp, vp, f, f, i, u32, vp, vp, vp, vp, vp]
    if hasattr(L, "oracle_diff_run"):
        L.oracle_diff_run.argtypes = [i, i, i, i, i, i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                      ctypes.POINTER(ctypes.c_int64)]
    _libs[path] = L
    return L


class OracleResult:
    def __init__(self, L, ptr):
        self._L, self._p = L, ptr

    def get(self, name):
        p, n = ctypes.c_void_p(), ctypes.c_size_t()
        if self._L.oracle_get(self._p, name.encode(), ctypes.byref(p), ctypes.byref(n)):
            raise KeyError(name)
        dt, width = _DT[name]
        cnt = n.value // np.dtype(dt).itemsize
        if cnt == 0:
            return np.zeros((0, width) if width > 1 else 0, dtype=dt)
        a = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(np.ctypeslib.as_ctypes_type(dt))), (cnt,)).copy()
        return a.reshape(-1, width) if width > 1 else a

    @property
    def seconds(self):
        return self._L.oracle_render_seconds(self._p)

    def counts(self):
        c = self.get("counts")
        keys = ("path", "curve", "vertex", "curve_fragment", "merged_fragment", "span")
        return {k: int(v) for k, v in zip(keys, c)}

    def __del__(self):
        try:
            self._L.oracle_free(self._p)
        except Exception:
            pass


def run(scene, width, height, samples, srgb=True, ms=False, transform=None, keep=True, omp=False):
    """Render `scene` (scanline.scene.SceneArrays) on the CPU oracle."""
    L = lib(omp)
    P, C, V, vtx, cvpos, ctype, cpath, cw, rule, color, ramp, table = scene.oracle_args()
    m = None
    if transform is not None:
        m = np.ascontiguousarray(np.asarray(transform, dtype=np.float32).reshape(16))
    ptr = L.oracle_run(P, C, V, vtx.ctypes.data, cvpos.ctypes.data, ctype.ctypes.data, cpath.ctypes.data,
                       cw.ctypes.data, rule.ctypes.data, color.ctypes.data, int(width), int(height),
                       int(samples), int(bool(srgb)), int(bool(ms)),
                       m.ctypes.data if m is not None else None, 1 if keep else 0,
                       ramp.ctypes.data if ramp.size else None, ramp.size,
                       table.ctypes.data if table.size else None, table.size)
    if not ptr:
        raise ValueError("oracle_run rejected the arguments")
    return OracleResult(L, ptr)


def diff_run(scene, width, height, points=None, colors=None, grad_image=None, samples=16, double=True):
    """The differentiable model on the CPU: image [H, W, 4] and, with grad_image, the
    gradients of sum(grad_image * image) w.r.t. the vertices [V, 2] and colours [P, 4]
    (float64 arrays for the double model, float32 for the float model).

    points: float32 [V, 2] (default: the scene's vertices); colors: float32 [P, 4] in
    [0, 1] (default: the scene's RGBA8 paint / 255)."""
    L = lib()
    P, C, V = scene.n_paths, scene.n_curves, scene.n_vertices
    pts = np.ascontiguousarray(scene.vertex if points is None else points, dtype=np.float32).reshape(V, 2)
    if colors is None:
        c = scene.fill_color
        colors = np.stack([(c >> (8 * k)) & 0xFF for k in range(4)], axis=1).astype(np.float32) / 255.0
    col = np.ascontiguousarray(colors, dtype=np.float32).reshape(P, 4)
    ft = np.float64 if double else np.float32
    img = np.zeros((height, width, 4), dtype=ft)
    gp = np.zeros((V, 2), dtype=ft)
    gc = np.zeros((P, 4), dtype=ft)
    g = None
    if grad_image is not None:
        g = np.ascontiguousarray(grad_image, dtype=ft).reshape(height, width, 4)
    n = ctypes.c_int64()
    rc = L.oracle_diff_run(1 if double else 0, int(width), int(height), int(samples), P, C, V, pts.ctypes.data,
                           scene.curve_vertex_pos.ctypes.data, scene.curve_type.ctypes.data,
                           scene.curve_path_id.ctypes.data, scene.fill_rule.ctypes.data, col.ctypes.data,
                           g.ctypes.data if g is not None else None, img.ctypes.data, gp.ctypes.data,
                           gc.ctypes.data, ctypes.byref(n))
    if rc:
        raise ValueError("oracle_diff_run: unsupported scene (rational segments?)")
    return img, gp, gc, n.value


def chord_sincos(x):
    """(sin, cos) of a float angle as the chord generator evaluates them."""
    s, c = ctypes.c_float(), ctypes.c_float()
    lib().oracle_chord_sincos(float(x), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def chord_ease(time_ms):
    return lib().oracle_chord_ease(int(time_ms))


def chord_frame(diagrams, time_ms):
    """The scene of `diagrams` (scanline.animation.ChordKeyframes) at time_ms, as dicts of
    the arrays the renderer dumps: vertex [V, 2] f32, vertex_path [V], curve [C, 4] u32
    {vertex pos, path, type | reversed << 8, weight bits}, fill_rule [P] u8, fill [P] u32."""
    L = lib()
    counts = [L.oracle_chord_paths(d.n_keys, d.n_nodes, d.link_offset.ctypes.data, int(time_ms)) for d in diagrams]
    P = int(sum(counts))
    out = {"vertex": np.zeros((24 * P, 2), np.float32), "vertex_path": np.zeros(24 * P, np.uint32),
           "curve": np.zeros((6 * P, 4), np.uint32), "fill_rule": np.zeros(P, np.uint8),
           "fill": np.zeros(P, np.uint32)}
    base = 0
    for d, n in zip(diagrams, counts):
        got = L.oracle_chord_diagram(d.n_keys, d.n_nodes, d.node_t0.ctypes.data, d.node_t1.ctypes.data,
                                     d.node_size.ctypes.data, d.node_color.ctypes.data, d.link_offset.ctypes.data,
                                     d.link_node.ctypes.data, d.link_from.ctypes.data, d.link_to.ctypes.data,
                                     float(d.center[0]), float(d.center[1]), int(time_ms), base,
                                     out["vertex"].ctypes.data, out["vertex_path"].ctypes.data,
                                     out["curve"].ctypes.data, out["fill_rule"].ctypes.data, out["fill"].ctypes.data)
        assert got == n
        base += int(n)
    return out


def chord_scene(frame):
    """A chord_frame result as scanline.scene.SceneArrays, as a host upload of the same
    curves would give it: each curve back in the order it was generated in (the reversed
    flag keeps its direction, which the winding depends on), so that the upload orients
    it again and derives the same flag."""
    from scanline.scene import SceneArrays
    cur = frame["curve"]
    vtx = frame["vertex"].copy()
    for vp, _, info, _ in cur[(cur[:, 2] >> 8) & 1 == 1]:
        nv = int(info & 7)
        vtx[vp:vp + nv] = vtx[vp:vp + nv][::-1].copy()
    return SceneArrays(vertex=vtx.reshape(-1), curve_vertex_pos=cur[:, 0].copy(),
                       curve_type=(cur[:, 2] & 0xFF).astype(np.uint8), curve_path_id=cur[:, 1].copy(),
                       curve_arc_w=np.ones(cur.shape[0], np.float32), fill_rule=frame["fill_rule"].copy(), fill_color=frame["fill"].copy())
