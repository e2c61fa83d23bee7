"""``VGRasterizer``: Python mirror of the reference's rasterizer interface over the C ABI.

Method names follow RasterizerBase::VGRasterizer (gpu-scanline/src/rasterizer/shared/
ras_base.h:133-378) and its construction through createRasterizer
(ras_factory.h:18-32).  Errors raise ScanlineError with the library's message; there is no
CPU fallback -- without a GPU, construction fails with SL_E_NODEVICE.
"""
import ctypes

import numpy as np

from . import _abi

DUMP_DTYPES = {
    "ts": (np.int32, 2), "frag_yx": (np.int32, 1), "frag_pid": (np.int32, 1),
    "frag_wn": (np.int8, 1), "frag_mask": (np.uint32, 4), "merged_mask": (np.uint32, 4),
    "draw_list": (np.int32, 4), "image": (np.uint8, 4), "pixel_table": (np.uint32, 1),
    "bucket_info": (np.int32, 2), "bucket_items": (np.uint32, 1), "bucket_ext": (np.uint32, 1),
    "bucket_off": (np.uint32, 1), "row_keys": (np.uint32, 1), "row_list": (np.uint32, 1),
    "row_beg": (np.uint32, 1), "row_end": (np.uint32, 1),
    "scene_vertex": (np.float32, 2), "scene_vertex_path": (np.uint32, 1), "scene_curve": (np.uint32, 4),
    "scene_fill_rule": (np.uint8, 1), "scene_fill": (np.uint32, 1),
}


class VGRasterizer:
    def __init__(self, device=0):
        self._L = _abi.lib()
        h = ctypes.c_void_p()
        rc = self._L.sl_create(device, ctypes.byref(h))
        if rc:
            raise _abi.ScanlineError(rc, "sl_create failed (no usable GPU?)")
        self._h = h
        self._w = self._h_px = 0

    # -- helpers --------------------------------------------------------------------
    def _chk(self, rc):
        if rc:
            raise _abi.ScanlineError(rc, self._L.sl_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._L.sl_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- configuration (ras_base.h:155-205) -----------------------------------------
    def setSamples(self, s):
        self._chk(self._L.sl_set_samples(self._h, int(s)))

    def setMultisampleOutput(self, f):
        self._chk(self._L.sl_set_multisample_output(self._h, int(bool(f))))

    def enableSRGBCorrection(self, f=True):
        self._chk(self._L.sl_enable_srgb_correction(self._h, int(bool(f))))

    def setInputTransform(self, mat4):
        """mat4: 4x4 array-like in glm's [column][row] order (16 floats column-major)."""
        m = np.ascontiguousarray(np.asarray(mat4, dtype=np.float32).reshape(16))
        self._chk(self._L.sl_set_input_transform(
            self._h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def resizeOutputBuffer(self, w, h):
        self._chk(self._L.sl_resize_output_buffer(self._h, int(w), int(h)))
        self._w, self._h_px = int(w), int(h)

    def outputBufferSize(self):
        return self._w, self._h_px

    def viewport(self, x, y, width, height):
        """viewport (ras_base.cpp:385-388): where the image lands in the window."""
        self._chk(self._L.sl_viewport(self._h, int(x), int(y), int(width), int(height)))

    def resizeViewport(self, width, height):
        """resizeViewport (ras_base.cpp:391-394)."""
        self._chk(self._L.sl_resize_viewport(self._h, int(width), int(height)))

    def present(self, dev_ptr, pitch, width, height, stream=None):
        """The window blit (ras_base.cpp:1170-1176) into a caller's device RGBA8 buffer."""
        self._chk(self._L.sl_present(self._h, ctypes.c_void_p(dev_ptr), int(pitch), int(width), int(height),
                                     ctypes.c_void_p(stream or 0)))

    def enableStepTiming(self, f=True):
        self._chk(self._L.sl_enable_step_timing(self._h, int(bool(f))))

    def dumpDebugData(self, f=True):
        self._chk(self._L.sl_debug_enable(self._h, int(bool(f))))

    # -- scene ------------------------------------------------------------------------
    def addVg(self, scene):
        """scene: scene.SceneArrays or scene.VGContainer."""
        arrs = scene.arrays() if hasattr(scene, "arrays") else scene
        c = arrs.as_c()
        self._chk(self._L.sl_add_vg(self._h, ctypes.byref(c)))

    def clear(self):
        self._chk(self._L.sl_clear(self._h))

    # -- rendering ----------------------------------------------------------------------
    def rasterize(self, stream=None):
        """stream: None or an integer hipStream_t (e.g. torch.cuda.current_stream().cuda_stream)."""
        self._chk(self._L.sl_rasterize(self._h, ctypes.c_void_p(stream or 0)))

    def rasterizeNoAlloc(self, dev_ptr, pitch, stream=None):
        """The allocation-free frame (no reference counterpart): renders into the caller's
        device image at `dev_ptr` (rows of `pitch` bytes) with every intermediate array
        in the library's static per-GPU scratch; see sl_rasterize_noalloc."""
        self._chk(self._L.sl_rasterize_noalloc(self._h, ctypes.c_void_p(dev_ptr), int(pitch),
                                               ctypes.c_void_p(stream or 0)))

    def saveOutputToFile(self, fileName):
        """saveOutputToFile (ras_base.cpp:1221-1277): the image with rows flipped and alpha
        forced to 255, as a PNG file."""
        from . import png
        png.write(fileName, self.output(flip=True))

    def output(self, flip=False):
        """The RGBA8 image as an (H, W, 4) uint8 array; flip=True gives saveOutputToFile's
        top-down rows with alpha forced to 255 (ras_base.cpp:1249-1262)."""
        out = np.empty((self._h_px, self._w, 4), dtype=np.uint8)
        self._chk(self._L.sl_read_output(self._h, out.ctypes.data, 1 if flip else 0))
        return out

    def outputDevicePtr(self):
        p, pitch = ctypes.c_void_p(), ctypes.c_int()
        self._chk(self._L.sl_output_device_ptr(self._h, ctypes.byref(p), ctypes.byref(pitch)))
        return p.value, pitch.value

    def elementNumber(self):
        c = _abi.SlCounts()
        self._chk(self._L.sl_element_number(self._h, ctypes.byref(c)))
        return {n: getattr(c, n) for n, _ in _abi.SlCounts._fields_}

    def stepTiming(self):
        buf = (ctypes.c_double * 32)()
        n = ctypes.c_int()
        self._chk(self._L.sl_step_timing(self._h, buf, 32, ctypes.byref(n)))
        return {self._L.sl_step_name(i).decode(): buf[i] for i in range(n.value)}

    def pixelCount(self):
        v = ctypes.c_int64()
        self._chk(self._L.sl_pixel_count(self._h, ctypes.byref(v)))
        return v.value

    def requantize(self, values, channel=0):
        """Device 8-bit requantisation of `values` (float32): (codes, decoded values).
        channel 0 = colour (sRGB when enabled), 1 = alpha."""
        v = np.ascontiguousarray(np.asarray(values, dtype=np.float32).reshape(-1))
        codes = np.empty(v.size, dtype=np.uint8)
        dec = np.empty(v.size, dtype=np.float32)
        self._chk(self._L.sl_debug_requantize(self._h, int(channel), v.ctypes.data, v.size, codes.ctypes.data,
                                              dec.ctypes.data))
        return codes, dec

    def requantizeSweep(self, bits_lo, bits_hi, channel=0):
        """Disagreements of the device requantisation with the threshold tables over every
        float bit pattern in [bits_lo, bits_hi)."""
        out = ctypes.c_uint64()
        self._chk(self._L.sl_debug_requantize_sweep(self._h, int(channel), int(bits_lo), int(bits_hi),
                                                    ctypes.byref(out)))
        return out.value

    def valuProbe(self, blocks=8192, iters=64):
        """(ms, wave VALU instructions) of one timed launch of the fma issue-rate probe."""
        ms, n = ctypes.c_double(), ctypes.c_uint64()
        self._chk(self._L.sl_debug_valu_probe(self._h, int(blocks), int(iters), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def dump(self, name):
        """Parity dump of one intermediate array in the reference's layout."""
        n = ctypes.c_size_t()
        self._chk(self._L.sl_debug_dump(self._h, name.encode(), None, 0, ctypes.byref(n)))
        dt, width = DUMP_DTYPES[name]
        out = np.empty(n.value // np.dtype(dt).itemsize, dtype=dt)
        if n.value:
            self._chk(self._L.sl_debug_dump(self._h, name.encode(), out.ctypes.data, n.value,
                                            ctypes.byref(n)))
        return out.reshape(-1, width) if width > 1 else out


def createRasterizer(device=0):
    """createRasterizer(PM_Cut_Mask_Comb_Scanline) (ras_factory.h:18-32)."""
    return VGRasterizer(device)
