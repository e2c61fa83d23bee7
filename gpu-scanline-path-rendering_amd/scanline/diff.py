"""Differentiable rendering through the C ABI (``sl_diff_*`` in include/scanline.h).

``DiffRenderer`` holds a scene's topology (curve types, vertex slots, paths, fill rules)
on the GPU; ``render(points, colors)`` is a ``torch.autograd.Function`` whose forward and
backward are the library's HIP kernels.  Coverage is a box filter of ``samples`` sample
rows per pixel, each integrated exactly along x; the image is the renderer's src-over
onto white in path order, unquantised, as float RGBA [H, W, 4].  The reference has no
differentiable mode.
"""
import ctypes

import numpy as np
import torch

from . import _abi


class DiffRenderer:
    def __init__(self, scene, width, height, samples=16, device=0):
        self._L = _abi.lib()
        a = scene.arrays() if hasattr(scene, "arrays") else scene
        self.width, self.height, self.samples = int(width), int(height), int(samples)
        self.n_paths, self.n_curves, self.n_vertices = a.n_paths, a.n_curves, a.n_vertices
        h = ctypes.c_void_p()
        rc = self._L.sl_diff_create(device, self.width, self.height, self.samples, a.n_paths, a.n_curves,
                                    a.n_vertices, a.curve_vertex_pos.ctypes.data, a.curve_type.ctypes.data,
                                    a.curve_path_id.ctypes.data, a.fill_rule.ctypes.data, ctypes.byref(h))
        if rc:
            raise _abi.ScanlineError(rc, "sl_diff_create failed (unsupported segment type or no GPU?)")
        self._h = h
        self.device = torch.device("cuda", device)
        self.points0 = torch.tensor(np.asarray(a.vertex, dtype=np.float32).reshape(-1, 2), device=self.device)
        c = a.fill_color.astype(np.uint32)
        rgba = np.stack([(c >> (8 * k)) & 0xFF for k in range(4)], axis=1).astype(np.float32) / 255.0
        self.colors0 = torch.tensor(rgba, device=self.device)
        # the library keeps one forward's state for the adjoint; every forward replaces it
        self.generation = 0

    def _chk(self, rc):
        if rc:
            raise _abi.ScanlineError(rc, self._L.sl_diff_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.sl_diff_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, points, colors):
        if points.shape != (self.n_vertices, 2) or colors.shape != (self.n_paths, 4):
            raise ValueError("points must be [V, 2] and colors [P, 4]")
        if points.dtype != torch.float32 or colors.dtype != torch.float32:
            raise TypeError("float32 tensors expected")
        if not (points.is_cuda and colors.is_cuda):
            raise ValueError("device tensors expected")

    def forward_raw(self, points, colors, keep=True):
        """The image (no autograd); `keep` keeps what backward_raw needs."""
        self._check(points, colors)
        points, colors = points.contiguous(), colors.contiguous()
        img = torch.empty((self.height, self.width, 4), dtype=torch.float32, device=self.device)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        self.generation += 1
        self._chk(self._L.sl_diff_forward(self._h, points.data_ptr(), colors.data_ptr(), img.data_ptr(),
                                          1 if keep else 0, ctypes.c_void_p(stream)))
        self._kept = (points, colors) if keep else None
        return img

    def backward_raw(self, grad_image):
        g = grad_image.contiguous().to(torch.float32)
        gp = torch.empty((self.n_vertices, 2), dtype=torch.float32, device=self.device)
        gc = torch.empty((self.n_paths, 4), dtype=torch.float32, device=self.device)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        self._chk(self._L.sl_diff_backward(self._h, g.data_ptr(), gp.data_ptr(), gc.data_ptr(),
                                           ctypes.c_void_p(stream)))
        return gp, gc

    def crossings(self):
        n = ctypes.c_int64()
        self._chk(self._L.sl_diff_crossings(self._h, ctypes.byref(n)))
        return n.value

    def render(self, points, colors):
        """Autograd-aware render: gradients flow to `points` and `colors`."""
        return _Render.apply(points, colors, self)


class _Render(torch.autograd.Function):
    """The library holds the state of one forward per renderer: a backward is only valid
    for the renderer's latest forward, and raises if a later forward replaced it."""

    @staticmethod
    def forward(ctx, points, colors, renderer):
        ctx.renderer = renderer
        img = renderer.forward_raw(points, colors, keep=True)
        ctx.generation = renderer.generation
        # the adjoint reads the forward's points and colours: keep them, and let autograd
        # notice in-place changes before the backward
        ctx.save_for_backward(points, colors)
        return img

    @staticmethod
    def backward(ctx, grad_image):
        r = ctx.renderer
        if r.generation != ctx.generation:
            raise RuntimeError("DiffRenderer.render: a later forward (render or forward_raw) replaced the state "
                               "this backward needs; one pending backward per renderer")
        ctx.saved_tensors  # raises if the inputs were modified in place since the forward
        gp, gc = r.backward_raw(grad_image)
        return gp, gc, None


def tie_vertices(scene):
    """Index map from the per-curve vertex slots to unique control points: slots of one
    path with identical coordinates (the shared ends of chained segments) become one
    point, so a closed contour stays closed when the points move.
    Returns (unique points [U, 2] float32, index [V] int64)."""
    a = scene.arrays() if hasattr(scene, "arrays") else scene
    V = a.n_vertices
    vpath = np.zeros(V, dtype=np.int64)
    for c in range(a.n_curves):
        n = int(a.curve_type[c]) & 7
        vpath[a.curve_vertex_pos[c]:a.curve_vertex_pos[c] + n] = a.curve_path_id[c]
    keys = {}
    index = np.zeros(V, dtype=np.int64)
    uniq = []
    for v in range(V):
        k = (int(vpath[v]), a.vertex[v, 0].tobytes(), a.vertex[v, 1].tobytes())
        if k not in keys:
            keys[k] = len(uniq)
            uniq.append(a.vertex[v])
        index[v] = keys[k]
    return np.array(uniq, dtype=np.float32).reshape(-1, 2), index
