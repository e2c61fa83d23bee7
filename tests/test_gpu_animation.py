"""GPU parity of the scenes generated on the device (sl_chord_load / sl_chord_frame).

The scene arrays a frame writes are compared bit for bit with the oracle's restatement of
the reference's generator (oracle/chord_oracle.cpp), and the rendered frame with the
oracle's render of the oracle's scene, at sizes from a dozen nodes to the full window of
the reference's demo (two diagrams of 150 nodes and about 1,800 links an interval).
"""
import ctypes

import numpy as np
import pytest

import pyoracle
import scanline
from scanline import animation as an

pytestmark = pytest.mark.gpu

W, H = an.CANVAS


def _diagrams(n_nodes, links, years=3, seed=1):
    return an.refugee_diagrams(n_nodes=n_nodes, n_years=years, links_per_year=links, seed=seed)


def _setup(r, spp=16, ms=False):
    r.setSamples(spp)
    r.enableSRGBCorrection(True)
    r.setMultisampleOutput(ms)
    r.resizeOutputBuffer(W, H)
    r.setInputTransform(np.eye(4, dtype=np.float32))
    r.dumpDebugData(False)
    r.enableStepTiming(False)


def _check_scene(r, ref):
    np.testing.assert_array_equal(r.dump("scene_vertex").view(np.uint32), ref["vertex"].view(np.uint32),
                                  err_msg="vertices")
    np.testing.assert_array_equal(r.dump("scene_vertex_path"), ref["vertex_path"], err_msg="vertex paths")
    np.testing.assert_array_equal(r.dump("scene_curve"), ref["curve"], err_msg="curve words")
    np.testing.assert_array_equal(r.dump("scene_fill_rule"), ref["fill_rule"], err_msg="fill rules")
    np.testing.assert_array_equal(r.dump("scene_fill"), ref["fill"], err_msg="fill words")


@pytest.mark.parametrize("n_nodes,links", [(12, 20), (150, 1500)])
def test_chord_scene_matches_oracle(rasterizer, n_nodes, links):
    ds = _diagrams(n_nodes, links)
    r = rasterizer
    _setup(r)
    anim = an.ChordAnimation(r, ds)
    for t in (0, 3000, 6000, 7000, 8000, 9500, 11999, 12000, 13500, 30000):
        print("chord scene:", n_nodes, links, "t =", t, flush=True)
        anim.frame(t)
        _check_scene(r, pyoracle.chord_frame(ds, t))


@pytest.mark.parametrize("spp,ms", [(16, False), (32, False), (8, True)])
def test_chord_frame_renders_as_oracle(rasterizer, spp, ms):
    ds = _diagrams(40, 120)
    r = rasterizer
    _setup(r, spp, ms)
    anim = an.ChordAnimation(r, ds)
    for t in (0, 8000):
        print("chord render:", spp, "ms" if ms else "cov", "t =", t, flush=True)
        anim.frame(t)
        r.rasterize()
        img = r.output()
        ref = pyoracle.run(pyoracle.chord_scene(pyoracle.chord_frame(ds, t)), W, H, spp, srgb=True, ms=ms,
                           keep=False)
        np.testing.assert_array_equal(img, ref.get("image").reshape(H, W, 4))
        c = r.elementNumber()
        assert c["path"] == anim.paths_at(t) and c["curve"] == 6 * c["path"] and c["vertex"] == 24 * c["path"]
        assert c["curve_fragment"] == ref.counts()["curve_fragment"]


def test_chord_frame_multisample_32spp(rasterizer):
    """The reference's default output (per-sample blending at 32 spp) of a generated frame,
    against the OpenMP build of the oracle."""
    ds = _diagrams(40, 120, seed=3)
    r = rasterizer
    _setup(r, 32, True)
    anim = an.ChordAnimation(r, ds)
    t = 8500
    print("chord render: 32 spp multisample t =", t, flush=True)
    anim.frame(t)
    r.rasterize()
    ref = pyoracle.run(pyoracle.chord_scene(pyoracle.chord_frame(ds, t)), W, H, 32, srgb=True, ms=True, keep=False,
                       omp=True)
    np.testing.assert_array_equal(r.output(), ref.get("image").reshape(H, W, 4))


def test_full_chord_frame_renders_as_oracle(rasterizer):
    ds = _diagrams(150, 1500)
    r = rasterizer
    _setup(r, 16)
    anim = an.ChordAnimation(r, ds)
    t = 8000
    print("chord render: full window t =", t, flush=True)
    anim.frame(t)
    r.rasterize()
    ref = pyoracle.run(pyoracle.chord_scene(pyoracle.chord_frame(ds, t)), W, H, 16, srgb=True, keep=False)
    np.testing.assert_array_equal(r.output(), ref.get("image").reshape(H, W, 4))


def test_generated_scene_equals_uploaded_scene(rasterizer):
    """A frame generated on the device renders exactly as the same scene uploaded with
    sl_add_vg; an upload after a generated frame replaces it."""
    ds = _diagrams(30, 80, seed=4)
    r = rasterizer
    _setup(r, 16)
    anim = an.ChordAnimation(r, ds)
    anim.frame(7300)
    r.rasterize()
    gen = r.output()
    scene = pyoracle.chord_scene(pyoracle.chord_frame(ds, 7300))
    r.addVg(scene)
    r.rasterize()
    np.testing.assert_array_equal(r.output(), gen)
    anim.frame(0)  # and back: the keyframes stay loaded across uploads
    r.rasterize()
    np.testing.assert_array_equal(
        r.output(), pyoracle.run(pyoracle.chord_scene(pyoracle.chord_frame(ds, 0)), W, H, 16, keep=False)
        .get("image").reshape(H, W, 4))


def test_chord_frames_queued_on_a_torch_stream(rasterizer):
    """Generated frames and renders queued on a caller's stream, then a frame on the null
    stream: each frame's scene is written after the previous frame on any stream."""
    import torch
    ds = _diagrams(40, 120, seed=2)
    r = rasterizer
    _setup(r, 16)
    anim = an.ChordAnimation(r, ds)
    st = torch.cuda.Stream()
    times = (6000, 6500, 7000, 7500, 8000)
    with torch.cuda.stream(st):
        for t in times:
            anim.frame(t, st.cuda_stream)
            r.rasterize(st.cuda_stream)
        img = r.output()
    ref = pyoracle.run(pyoracle.chord_scene(pyoracle.chord_frame(ds, times[-1])), W, H, 16, keep=False)
    np.testing.assert_array_equal(img, ref.get("image").reshape(H, W, 4))
    with torch.cuda.stream(st):
        anim.frame(9000, st.cuda_stream)
        r.rasterize(st.cuda_stream)
    anim.frame(10000)  # null stream, after the frame on st
    r.rasterize()
    ref = pyoracle.run(pyoracle.chord_scene(pyoracle.chord_frame(ds, 10000)), W, H, 16, keep=False)
    np.testing.assert_array_equal(r.output(), ref.get("image").reshape(H, W, 4))


def test_chord_errors(rasterizer):
    r = rasterizer
    L, h = r._L, r._h
    fresh = scanline.VGRasterizer(0)
    try:
        assert fresh._L.sl_chord_frame(fresh._h, 0, None) == scanline._abi.SL_E_STATE
    finally:
        fresh.close()
    ds = _diagrams(12, 20)
    an.ChordAnimation(r, ds)
    assert L.sl_chord_frame(h, -1, None) == scanline._abi.SL_E_ARG
    bad = an.refugee_diagrams(n_nodes=12, n_years=3, links_per_year=20, seed=1)[0]
    bad.link_node = bad.link_node.copy()
    bad.link_node[0, 1] = 12
    arr = (scanline._abi.SlChord * 1)(bad.as_c())
    assert L.sl_chord_load(h, arr, 1) == scanline._abi.SL_E_ARG
    assert b"names node 12" in L.sl_last_error(h)
    one = an.refugee_diagrams(n_nodes=12, n_years=3, links_per_year=20, seed=1)[0]
    one.n_keys = 1
    arr = (scanline._abi.SlChord * 1)(one.as_c())
    assert L.sl_chord_load(h, arr, 1) == scanline._abi.SL_E_ARG
    assert L.sl_chord_load(h, arr, 0) == scanline._abi.SL_E_ARG
    assert L.sl_chord_load(h, ctypes.cast(arr, ctypes.POINTER(scanline._abi.SlChord)), 9) == scanline._abi.SL_E_ARG
