"""``BatchRasterizer``: many independent scenes rendered by one captured launch sequence.

The reference renders one scene per ``rasterize()`` (ras_base.cpp:1131-1195); a batch gives
every scene the image a separate ``VGRasterizer`` would give it (coverage output), from a
HIP graph that covers all scenes.  ``tune()`` picks the launch shape by timing candidates
on the batch itself.
"""
import ctypes

import numpy as np

from . import _abi


class BatchRasterizer:
    def __init__(self, scenes, sizes, samples, srgb=True, device=0):
        """scenes: SceneArrays / VGContainer list; sizes: (W, H) per scene; samples: 8/16/32."""
        self._L = _abi.lib()
        arrs = [s.arrays() if hasattr(s, "arrays") else s for s in scenes]
        n = len(arrs)
        if n == 0 or len(sizes) != n or len(samples) != n:
            raise ValueError("one size and one sample count per scene")
        self._keep = [a.as_c() for a in arrs]
        cs = (_abi.SlScene * n)(*self._keep)
        self.sizes = [(int(w), int(h)) for w, h in sizes]
        w = np.array([s[0] for s in self.sizes], dtype=np.int32)
        h = np.array([s[1] for s in self.sizes], dtype=np.int32)
        spp = np.array(samples, dtype=np.int32)
        out = ctypes.c_void_p()
        rc = self._L.sl_batch_create(device, ctypes.cast(cs, ctypes.c_void_p), w.ctypes.data, h.ctypes.data,
                                     spp.ctypes.data, n, int(bool(srgb)), ctypes.byref(out))
        if rc:
            raise _abi.ScanlineError(rc, "sl_batch_create failed")
        self._b = out
        self.n = n

    def _chk(self, rc):
        if rc:
            raise _abi.ScanlineError(rc, self._L.sl_batch_last_error(self._b).decode())

    def close(self):
        if getattr(self, "_b", None):
            self._L.sl_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, stream=None):
        """Replay the captured graph on `stream` (an integer hipStream_t or None)."""
        self._chk(self._L.sl_batch_render(self._b, ctypes.c_void_p(stream or 0)))

    def render_eager(self, stream=None):
        """The same kernels issued one by one (no graph)."""
        self._chk(self._L.sl_batch_render_eager(self._b, ctypes.c_void_p(stream or 0)))

    def tune(self, reps=3):
        best = ctypes.c_double()
        self._chk(self._L.sl_batch_tune(self._b, int(reps), ctypes.byref(best)))
        return best.value, self.shape()

    def shape(self):
        s = np.zeros(4, dtype=np.int32)
        self._chk(self._L.sl_batch_get_shape(self._b, s.ctypes.data))
        return tuple(int(v) for v in s)

    def set_shape(self, shape):
        s = np.array(shape, dtype=np.int32)
        self._chk(self._L.sl_batch_set_shape(self._b, s.ctypes.data))

    def elementNumber(self):
        c = _abi.SlCounts()
        self._chk(self._L.sl_batch_element_number(self._b, ctypes.byref(c)))
        return {k: getattr(c, k) for k, _ in _abi.SlCounts._fields_}

    def output(self, i, flip=False):
        W, H = self.sizes[i]
        out = np.empty((H, W, 4), dtype=np.uint8)
        self._chk(self._L.sl_batch_read_output(self._b, int(i), out.ctypes.data, 1 if flip else 0))
        return out


def batch_workload(n_scenes=4096, seed=None, max_side=512, min_side=16, min_paths=5, max_paths=200):
    """The batched workload: canvas sides log-uniform in [16, 512] independently, 5-200
    closed cubic chains per scene (the 512x512 small config's construction in that
    canvas), 8 or 16 samples per scene.  Returns (scenes, sizes, samples)."""
    from . import scene as _scene
    if seed is None:
        seed = _scene.SEED_BASE ^ 4
    rng = np.random.default_rng(seed)
    lo, hi = np.log(min_side), np.log(max_side)
    W = np.rint(np.exp(rng.uniform(lo, hi, n_scenes))).astype(int)
    H = np.rint(np.exp(rng.uniform(lo, hi, n_scenes))).astype(int)
    P = rng.integers(min_paths, max_paths + 1, n_scenes)
    S = rng.choice([8, 16], n_scenes)
    seeds = rng.integers(0, 2**63 - 1, n_scenes, dtype=np.int64)
    scenes = [_scene.generate(_scene.KIND_CUBIC_CHAINS, int(W[i]), int(H[i]), int(P[i]), int(seeds[i]))
              for i in range(n_scenes)]
    return scenes, list(zip(W.tolist(), H.tolist())), S.tolist()
