"""GPU: the allocation-free frame (sl_rasterize_noalloc) gives sl_rasterize's image.

Two calls in a row, the second on a smaller scene, each rendered into a caller-allocated
device image with a row pitch wider than the image, and compared bit for bit with the
image and counts of sl_rasterize on the same handle.  A third call uses the multisample
output (the merged masks then live in the scratch too).  Device free memory does not move
across the calls.  A scene that needs more than the static scratch is refused with
SL_E_UNSUPPORTED, and the handle keeps working after the refusal.
"""
import ctypes

import numpy as np
import pytest

import scanline

pytestmark = pytest.mark.gpu


def _hip():
    scanline._abi.lib()
    return ctypes.CDLL("libamdhip64.so.7")  # the runtime libscanline.so is bound to


def _free_bytes(hip):
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _configure(r, W, H, spp, ms=False):
    r.setSamples(spp)
    r.enableSRGBCorrection(True)
    r.setMultisampleOutput(ms)
    r.resizeOutputBuffer(W, H)
    r.setInputTransform(np.eye(4, dtype=np.float32))
    r.dumpDebugData(False)
    r.enableStepTiming(False)


def _noalloc_image(r, W, H, pad=64, hip=None):
    """Image of one allocation-free frame in a caller buffer with `pad` bytes after each
    row; with `hip`, also the device free memory just before and just after the call."""
    import torch
    pitch = W * 4 + pad
    buf = torch.full((H * pitch,), 0x5A, dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream()
    st.synchronize()
    free0 = _free_bytes(hip) if hip else None
    r.rasterizeNoAlloc(buf.data_ptr(), pitch, st.cuda_stream)
    st.synchronize()
    free1 = _free_bytes(hip) if hip else None
    return buf.cpu().numpy().reshape(H, pitch)[:, :W * 4].reshape(H, W, 4), buf, (free0, free1)


def test_noalloc_frame_matches_rasterize():
    hip = _hip()
    r = scanline.VGRasterizer(0)
    try:
        cases = [
            ("mixed 384x320, 300 paths, 32 spp", scanline.scene.KIND_MIXED, 384, 320, 300, 32, False),
            ("cubic chains 130x97, 25 paths, 8 spp", scanline.scene.KIND_CUBIC_CHAINS, 130, 97, 25, 8, False),
            ("cubic chains 130x97, 25 paths, 16 spp, ms", scanline.scene.KIND_CUBIC_CHAINS, 130, 97, 25, 16, True),
        ]
        for k, (label, kind, W, H, n, spp, ms) in enumerate(cases):
            print("noalloc case:", label, flush=True)
            scene = scanline.generate(kind, W, H, n, 0x5EED0700 + k)
            _configure(r, W, H, spp, ms)
            r.addVg(scene)
            r.rasterize()
            want = r.output()
            counts = r.elementNumber()
            got, buf, (free0, free1) = _noalloc_image(r, W, H, hip=hip)
            np.testing.assert_array_equal(got, want, err_msg=label)
            assert free1 == free0, (label, free0, free1)  # the call allocated no device memory
            assert r.elementNumber() == counts, label
            # the bytes past each row's end are the caller's: left as they were
            pad = buf.cpu().numpy().reshape(H, -1)[:, W * 4:]
            assert (pad == 0x5A).all(), label
            del buf
    finally:
        r.close()


def test_noalloc_refuses_a_scene_beyond_the_scratch():
    r = scanline.VGRasterizer(0)
    try:
        scene, W, H, spp = scanline.config_scene(2)
        print("noalloc case: 2048x2048 20000 paths (refused)", flush=True)
        _configure(r, W, H, spp)
        r.addVg(scene)
        import torch
        buf = torch.empty((H * W * 4,), dtype=torch.uint8, device="cuda")
        with pytest.raises(scanline.ScanlineError) as e:
            r.rasterizeNoAlloc(buf.data_ptr(), W * 4, torch.cuda.current_stream().cuda_stream)
        assert e.value.code == scanline._abi.SL_E_UNSUPPORTED
        assert "static scratch" in str(e.value)
        # the handle still renders, both ways, after the refusal
        small = scanline.generate(scanline.scene.KIND_MIXED, 200, 150, 60, 0x5EED0710)
        _configure(r, 200, 150, 32)
        r.addVg(small)
        r.rasterize()
        got, _, _ = _noalloc_image(r, 200, 150)
        np.testing.assert_array_equal(got, r.output())
    finally:
        r.close()


def test_noalloc_argument_checks():
    r = scanline.VGRasterizer(0)
    try:
        scene = scanline.generate(scanline.scene.KIND_CUBIC_CHAINS, 64, 48, 5, 3)
        _configure(r, 64, 48, 8)
        r.addVg(scene)
        with pytest.raises(scanline.ScanlineError) as e:
            r.rasterizeNoAlloc(0, 256)
        assert e.value.code == scanline._abi.SL_E_ARG
        import torch
        buf = torch.empty((48 * 256,), dtype=torch.uint8, device="cuda")
        with pytest.raises(scanline.ScanlineError) as e:
            r.rasterizeNoAlloc(buf.data_ptr(), 64 * 4 - 4)
        assert e.value.code == scanline._abi.SL_E_ARG
        r.dumpDebugData(True)
        with pytest.raises(scanline.ScanlineError) as e:
            r.rasterizeNoAlloc(buf.data_ptr(), 256)
        assert e.value.code == scanline._abi.SL_E_STATE
    finally:
        r.close()
