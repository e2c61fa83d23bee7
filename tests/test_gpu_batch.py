"""GPU: batched rendering gives every scene exactly its single-scene image.

A batch renders many independent scenes of uneven size (canvas sides log-uniform, cubic
chain scenes, 8/16/32 samples mixed) from one captured HIP graph.  Every scene's image
must be bit-identical to the same scene rendered alone by a ``VGRasterizer`` (which the
parity tests tie to the oracle), and a few scenes are checked against the oracle
directly.  The graph is replayed several times with eager work, synchronisations and a
poisoned output buffer in between; every replay is checked.
"""
import ctypes

import numpy as np
import pytest

import pyoracle
import scanline

pytestmark = pytest.mark.gpu


def _hip():
    # the HIP runtime libscanline.so is bound to (matched by soname once it is loaded)
    scanline._abi.lib()
    return ctypes.CDLL("libamdhip64.so.7")


@pytest.fixture(scope="module")
def small_batch():
    scenes, sizes, spp = scanline.batch_workload(40, seed=11, max_side=200, max_paths=40)
    spp[0] = 32  # all three sample counts present
    return scenes, sizes, spp


def single_images(scenes, sizes, spp):
    out = []
    r = scanline.VGRasterizer(0)
    try:
        for s, (w, h), n in zip(scenes, sizes, spp):
            r.setSamples(n)
            r.setMultisampleOutput(False)
            r.enableSRGBCorrection(True)
            r.resizeOutputBuffer(w, h)
            r.addVg(s)
            r.rasterize()
            out.append(r.output())
    finally:
        r.close()
    return out


def test_batch_matches_single_scene_renders(small_batch):
    scenes, sizes, spp = small_batch
    print("batch of", len(scenes), "scenes", flush=True)
    ref = single_images(scenes, sizes, spp)
    b = scanline.BatchRasterizer(scenes, sizes, spp)
    try:
        hip = _hip()
        for rep in range(3):
            # poison every image, do unrelated eager work, then replay the graph
            for i in range(len(scenes)):
                p, pitch = ctypes.c_void_p(), ctypes.c_int()
                assert b._L.sl_batch_output_device_ptr(b._b, i, ctypes.byref(p), ctypes.byref(pitch)) == 0
                assert hip.hipMemset(p, ctypes.c_int(rep * 37 + 1), ctypes.c_size_t(pitch.value * sizes[i][1])) == 0
            single_images(scenes[:2], sizes[:2], spp[:2])
            b.render()
            for i in range(len(scenes)):
                np.testing.assert_array_equal(b.output(i), ref[i], err_msg="scene %d replay %d" % (i, rep))
        b.render_eager()
        for i in range(len(scenes)):
            np.testing.assert_array_equal(b.output(i), ref[i], err_msg="scene %d eager" % i)
        cnt = b.elementNumber()
        assert cnt["path"] == sum(s.n_paths for s in scenes)
    finally:
        b.close()


def test_batch_against_oracle(small_batch):
    scenes, sizes, spp = small_batch
    sel = [0, 1, 2, 3]
    b = scanline.BatchRasterizer([scenes[i] for i in sel], [sizes[i] for i in sel], [spp[i] for i in sel])
    try:
        b.render()
        for k, i in enumerate(sel):
            w, h = sizes[i]
            ref = pyoracle.run(scenes[i], w, h, spp[i], srgb=True, ms=False).get("image").reshape(h, w, 4)
            np.testing.assert_array_equal(b.output(k), ref, err_msg="scene %d" % i)
    finally:
        b.close()


def test_every_launch_shape_gives_the_same_images(small_batch):
    scenes, sizes, spp = small_batch
    scenes, sizes, spp = scenes[:12], sizes[:12], spp[:12]
    ref = single_images(scenes, sizes, spp)
    b = scanline.BatchRasterizer(scenes, sizes, spp)
    try:
        for shape in [(4096, 1024, 4, 128), (65536, 4096, 8, 64), (16384, 4096, 8, 256)]:
            b.set_shape(shape)
            assert b.shape() == shape
            b.render()
            for i in range(len(scenes)):
                np.testing.assert_array_equal(b.output(i), ref[i], err_msg="scene %d shape %s" % (i, shape))
        best_ms, shape = b.tune(reps=1)
        assert best_ms > 0
        b.render()
        for i in range(len(scenes)):
            np.testing.assert_array_equal(b.output(i), ref[i], err_msg="scene %d tuned %s" % (i, shape))
    finally:
        b.close()


def test_bad_batch_arguments():
    s = scanline.generate(scanline.scene.KIND_CUBIC_CHAINS, 32, 32, 5, 1)
    with pytest.raises(scanline.ScanlineError):
        scanline.BatchRasterizer([s], [(32, 32)], [12])
    with pytest.raises(scanline.ScanlineError):
        scanline.BatchRasterizer([s], [(0, 32)], [8])


def test_frames_on_two_streams_and_retuning(small_batch):
    """Replays and eager frames alternate between two non-blocking streams without host
    waits, and a shape change and a tuning run follow queued frames: each frame is
    ordered after the last, so every image stays the single-scene image."""
    import torch
    scenes, sizes, spp = small_batch
    scenes, sizes, spp = scenes[:16], sizes[:16], spp[:16]
    ref = single_images(scenes, sizes, spp)
    b = scanline.BatchRasterizer(scenes, sizes, spp)
    try:
        s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
        for k in range(4):
            b.render(s1.cuda_stream)
            b.render_eager(s2.cuda_stream)
            b.render(s2.cuda_stream if k % 2 else s1.cuda_stream)
        for i in range(len(scenes)):
            np.testing.assert_array_equal(b.output(i), ref[i], err_msg="scene %d two streams" % i)
        b.render(s1.cuda_stream)
        b.set_shape((16384, 4096, 8, 128))
        b.render(s2.cuda_stream)
        b.tune(reps=1)
        b.render(s1.cuda_stream)
        for i in range(len(scenes)):
            np.testing.assert_array_equal(b.output(i), ref[i], err_msg="scene %d after retuning" % i)
        torch.cuda.synchronize()
    finally:
        b.close()
