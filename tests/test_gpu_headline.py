"""GPU parity at the full sizes of the timed workloads.

* The plain large scene (2048x2048, 20,000 mixed paths, 32 spp, sRGB): every parity array
  -- intersections, fragment keys, path id | sign, scan winding, raw masks, resolved masks,
  draw list -- the element counts, the reference's count-pixel and the RGBA8 image against
  the OpenMP build of the oracle, bit for bit, in coverage output; then the multisample
  output (the reference's default mode) against the oracle's image; then re-rendered frames
  of the resident scene, which run with no host reads, give the same image again.
* The batched workload (4096 scenes, sides 16-512, 5-200 paths, 8/16 spp): every scene of
  one graph replay equals its own single-scene render, and 64 scenes drawn at random equal
  the oracle.
"""
import ctypes

import numpy as np
import pytest

import pyoracle
import scanline

pytestmark = pytest.mark.gpu

PARITY = ("ts", "frag_yx", "frag_pid", "frag_wn", "frag_mask", "draw_list", "merged_mask")


def _configure(r, W, H, spp, ms, dump):
    r.setSamples(spp)
    r.enableSRGBCorrection(True)
    r.setMultisampleOutput(ms)
    r.resizeOutputBuffer(W, H)
    r.setInputTransform(np.eye(4, dtype=np.float32))
    r.dumpDebugData(dump)
    r.enableStepTiming(False)


@pytest.fixture(scope="module")
def headline():
    scene, W, H, spp = scanline.config_scene(2)
    assert (W, H, spp, scene.n_paths) == (2048, 2048, 32, 20000)
    return scene, W, H, spp


def test_headline_scene_every_stage(rasterizer, headline):
    scene, W, H, spp = headline
    print("parity case: headline 2048x2048 20000 paths 32 spp coverage, oracle (OpenMP)", flush=True)
    ref = pyoracle.run(scene, W, H, spp, srgb=True, ms=False, keep=True, omp=True)
    r = rasterizer
    _configure(r, W, H, spp, False, True)
    r.addVg(scene)
    r.rasterize()
    cnt, cref = r.elementNumber(), ref.counts()
    assert cnt == cref, (cnt, cref)
    assert cref["curve_fragment"] > 5e7
    for name in PARITY:
        got, want = r.dump(name), ref.get(name)
        assert got.shape == want.shape, (name, got.shape, want.shape)
        if not np.array_equal(got, want):
            bad = np.flatnonzero((got != want).reshape(got.shape[0], -1).any(axis=1))
            raise AssertionError("%s: %d rows differ, first at %s" % (name, bad.size, bad[:5]))
        del got, want
    assert r.pixelCount() == int(ref.get("pixel_count")[0])
    np.testing.assert_array_equal(r.output(), ref.get("image").reshape(H, W, 4))


def test_headline_scene_requeued_frames(rasterizer, headline):
    """Frames of the resident scene after the first are queued with the first frame's
    sizes and no host reads; their images and counts equal the oracle's."""
    scene, W, H, spp = headline
    print("parity case: headline re-rendered frames", flush=True)
    ref = pyoracle.run(scene, W, H, spp, srgb=True, ms=False, keep=False, omp=True)
    want = ref.get("image").reshape(H, W, 4)
    r = rasterizer
    _configure(r, W, H, spp, False, False)
    r.addVg(scene)
    r.rasterize()
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime libscanline.so is bound to
    for k in range(3):
        p, pitch = r.outputDevicePtr()
        assert hip.hipMemset(ctypes.c_void_p(p), ctypes.c_int(11 + k), ctypes.c_size_t(pitch * H)) == 0
        r.rasterize()
        np.testing.assert_array_equal(r.output(), want, err_msg="frame %d" % k)
    assert r.elementNumber() == ref.counts()


def test_headline_scene_multisample_output(rasterizer, headline):
    scene, W, H, spp = headline
    print("parity case: headline 2048x2048 32 spp multisample output", flush=True)
    ref = pyoracle.run(scene, W, H, spp, srgb=True, ms=True, keep=False, omp=True)
    r = rasterizer
    _configure(r, W, H, spp, True, False)
    r.addVg(scene)
    r.rasterize()
    assert r.elementNumber() == ref.counts()
    got = r.output().astype(np.int16)
    want = ref.get("image").reshape(H, W, 4).astype(np.int16)
    d = np.abs(got - want)
    assert d.max() <= 1, (d.max(), np.argwhere(d > 1)[:5])
    assert (d == 0).mean() == 1.0


def test_batch_workload_full_size():
    scenes, sizes, spp = scanline.batch_workload(4096)
    print("parity case: batch of 4096 scenes", flush=True)
    b = scanline.BatchRasterizer(scenes, sizes, spp)
    r = scanline.VGRasterizer(0)
    try:
        b.render()
        b.render()
        for i, (s, (w, h), n) in enumerate(zip(scenes, sizes, spp)):
            _configure(r, w, h, n, False, False)
            r.addVg(s)
            r.rasterize()
            np.testing.assert_array_equal(b.output(i), r.output(), err_msg="scene %d" % i)
        sel = np.random.default_rng(4096).choice(len(scenes), 64, replace=False)
        for i in sel:
            w, h = sizes[i]
            ref = pyoracle.run(scenes[i], w, h, spp[i], srgb=True, ms=False, keep=False)
            np.testing.assert_array_equal(b.output(int(i)), ref.get("image").reshape(h, w, 4),
                                          err_msg="scene %d against the oracle" % i)
        cnt = b.elementNumber()
        assert cnt["path"] == sum(s.n_paths for s in scenes)
    finally:
        r.close()
        b.close()
