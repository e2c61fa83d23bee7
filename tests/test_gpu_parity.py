"""GPU parity: the HIP renderer against the CPU oracle on the same scenes.

Bar (per stage): intersections, raw fragment keys / path ids / winding / masks, resolved
per-sample masks and the draw list are bit-exact; the final RGBA8 image is within 1/255
per channel (it is in fact expected to be identical, as both sides quantise identically).
"""
import os

import numpy as np
import pytest

import pyoracle
import scanline
from scanline import VGContainer, rgba

pytestmark = pytest.mark.gpu

NW = {8: 1, 16: 2, 32: 4}


def render_gpu(r, scene, W, H, spp, srgb=True, ms=False, transform=None):
    r.setSamples(spp)
    r.enableSRGBCorrection(srgb)
    r.setMultisampleOutput(ms)
    r.resizeOutputBuffer(W, H)
    if transform is not None:
        r.setInputTransform(transform)
    else:
        r.setInputTransform(np.eye(4, dtype=np.float32))
    r.dumpDebugData(True)
    r.addVg(scene)
    r.rasterize()
    return r


def compare(r, scene, W, H, spp, srgb=True, ms=False, transform=None, label=""):
    arrs = scene.arrays() if hasattr(scene, "arrays") else scene
    print("parity case:", label, W, H, spp, "srgb" if srgb else "lrgb", "ms" if ms else "cov", flush=True)
    ref = pyoracle.run(arrs, W, H, spp, srgb=srgb, ms=ms, transform=transform)
    render_gpu(r, arrs, W, H, spp, srgb, ms, transform)
    cnt_ref = ref.counts()
    cnt_gpu = r.elementNumber()
    assert cnt_gpu["curve_fragment"] == cnt_ref["curve_fragment"], (cnt_gpu, cnt_ref)
    F = cnt_ref["curve_fragment"]
    if F:
        np.testing.assert_array_equal(r.dump("ts"), ref.get("ts"), err_msg="intersections")
        np.testing.assert_array_equal(r.dump("frag_yx"), ref.get("frag_yx"), err_msg="fragment keys")
        np.testing.assert_array_equal(r.dump("frag_pid"), ref.get("frag_pid"), err_msg="path id | sign")
        np.testing.assert_array_equal(r.dump("frag_wn"), ref.get("frag_wn"), err_msg="scan winding")
        nw = NW[spp]
        np.testing.assert_array_equal(r.dump("frag_mask")[:, :nw], ref.get("frag_mask")[:, :nw],
                                      err_msg="raw sample masks")
        assert cnt_gpu["merged_fragment"] == cnt_ref["merged_fragment"], (cnt_gpu, cnt_ref)
        assert cnt_gpu["span"] == cnt_ref["span"], (cnt_gpu, cnt_ref)
        np.testing.assert_array_equal(r.dump("draw_list"), ref.get("draw_list"), err_msg="draw list")
        np.testing.assert_array_equal(r.dump("merged_mask")[:, :nw], ref.get("merged_mask")[:, :nw],
                                      err_msg="resolved per-sample masks")
        assert r.pixelCount() == int(ref.get("pixel_count")[0])
    img = r.output()
    ref_img = ref.get("image").reshape(H, W, 4)
    diff = np.abs(img.astype(np.int32) - ref_img.astype(np.int32))
    assert diff.max() <= 1, "colour differs by %d/255 at %s" % (diff.max(), np.argwhere(diff > 1)[:5])
    return cnt_ref, float((diff == 0).mean())


def test_generated_cubic_chains_config1(rasterizer):
    """The small plain config: 512x512, 200 closed cubic chains, nonzero, 8 spp."""
    scene, W, H, spp = scanline.config_scene(1)
    cnt, exact = compare(rasterizer, scene, W, H, spp, label="config1")
    assert cnt["curve_fragment"] > 100000
    assert exact == 1.0


@pytest.mark.parametrize("spp", [8, 16, 32])
@pytest.mark.parametrize("srgb", [True, False])
def test_generated_mixed(rasterizer, spp, srgb):
    scene = scanline.generate(scanline.scene.KIND_MIXED, 384, 320, 300, 0x5CA1AB1E ^ 2)
    compare(rasterizer, scene, 384, 320, spp, srgb=srgb, label="mixed")


@pytest.mark.parametrize("spp", [8, 16, 32])
def test_generated_glyphs(rasterizer, spp):
    scene = scanline.generate(scanline.scene.KIND_GLYPHS, 512, 512, 800, 0x5CA1AB1E ^ 3)
    compare(rasterizer, scene, 512, 512, spp, label="glyphs")


@pytest.mark.parametrize("spp", [8, 16, 32])
def test_multisample_output(rasterizer, spp):
    scene = scanline.generate(scanline.scene.KIND_CUBIC_CHAINS, 256, 256, 40, 77)
    compare(rasterizer, scene, 256, 256, spp, ms=True, label="ms")


@pytest.mark.parametrize("spp", [8, 16, 32])
@pytest.mark.parametrize("srgb", [True, False])
def test_multisample_mixed(rasterizer, spp, srgb):
    """Per-sample output of translucent mixed paths (both fill rules, partly covered
    samples at every edge) in both colour encodings: each sample count's register layout
    (1, 2 or 4 columns a state register) and block size of the compositor."""
    scene = scanline.generate(scanline.scene.KIND_MIXED, 320, 200, 200, 0x5CA1AB1E ^ 11)
    compare(rasterizer, scene, 320, 200, spp, srgb=srgb, ms=True, label="ms-mixed")


def test_odd_canvas_and_transform(rasterizer):
    """Odd width/height (clipped last column/row) and a non-identity projective transform."""
    scene = scanline.generate(scanline.scene.KIND_MIXED, 301, 257, 150, 99)
    m = np.eye(4, dtype=np.float32)
    m[0, 0], m[1, 1] = 0.9, 1.1          # column-major: m[col][row]
    m[3, 0], m[3, 1] = 7.25, -3.5        # translation
    m[1, 0] = 0.05                        # shear
    compare(rasterizer, scene, 301, 257, 32, transform=m, label="transform")


@pytest.mark.parametrize("spp", [8, 16, 32])
def test_odd_canvas_multisample(rasterizer, spp):
    """Per-sample output on a canvas whose last 64-pixel tile and last cell row are partial
    (301 x 257), under the same projective transform."""
    scene = scanline.generate(scanline.scene.KIND_MIXED, 301, 257, 150, 99)
    m = np.eye(4, dtype=np.float32)
    m[0, 0], m[1, 1] = 0.9, 1.1
    m[3, 0], m[3, 1] = 7.25, -3.5
    m[1, 0] = 0.05
    compare(rasterizer, scene, 301, 257, spp, ms=True, transform=m, label="odd-ms")


def test_hand_shapes(rasterizer):
    """Rectangles on odd / integral coordinates, nested squares under both rules, a star,
    a rational arc and curves leaving the canvas."""
    vg = VGContainer()
    vg.addPath(0, rgba(200, 10, 10))
    vg.addPolygon([(3, 3), (21, 3), (21, 17), (3, 17)])
    vg.addPath(1, rgba(10, 200, 10, 180))
    vg.addPolygon([(30, 5), (60, 5), (60, 35), (30, 35)])
    vg.addPolygon([(35, 10), (55, 10), (55, 30), (35, 30)])
    vg.addPath(0, rgba(10, 10, 200, 128))
    vg.addPolygon([(30, 45), (60, 45), (60, 75), (30, 75)])
    vg.addPolygon([(35, 50), (55, 50), (55, 70), (35, 70)])
    vg.addPath(0, rgba(90, 90, 20))
    cx, cy, rr = 90.0, 40.0, 25.0
    pts = [(cx + rr * np.cos(2 * np.pi * k * 2 / 5 - np.pi / 2), cy + rr * np.sin(2 * np.pi * k * 2 / 5 - np.pi / 2))
           for k in range(5)]
    vg.addPolygon(pts)
    vg.addPath(1, rgba(20, 90, 90))
    cx, cy = 90.0, 95.0
    pts = [(cx + rr * np.cos(2 * np.pi * k * 2 / 5 - np.pi / 2), cy + rr * np.sin(2 * np.pi * k * 2 / 5 - np.pi / 2))
           for k in range(5)]
    vg.addPolygon(pts)
    vg.addPath(0, rgba(150, 60, 160, 220))
    vg.addRational((10, 90), (10, 120), (40, 120), 0.70710678)
    vg.addRational((40, 120), (40, 90), (10, 90), 2.0)
    vg.addPath(0, rgba(30, 30, 30, 100))
    vg.addQuadratic((-20, 60), (64, -40), (140, 60))
    vg.addCubic((140, 60), (100, 200), (0, -50), (-20, 60))
    for spp in (8, 16, 32):
        compare(rasterizer, vg, 128, 128, spp, label="hand")
        compare(rasterizer, vg, 128, 128, spp, ms=True, label="hand-ms")


def test_empty_and_invisible(rasterizer):
    """Paths entirely outside the canvas produce no fragments: white output, F == 0."""
    vg = VGContainer()
    vg.addPath(0, rgba(0, 0, 0))
    vg.addPolygon([(-50, -50), (-10, -50), (-10, -10)])
    cnt, exact = compare(rasterizer, vg, 64, 64, 32, label="invisible")
    assert cnt["curve_fragment"] == 0
    assert (rasterizer.output() == 255).all()


@pytest.mark.parametrize("length", [120, 230, 460, 1900])
def test_bucket_size_classes(rasterizer, length):
    """Thin slivers lying inside one cell row put ~length fragments into a single
    (path, row) bucket, covering every resolve size class (<=64, <=128, <=256, <=512 and
    the global-scratch path beyond)."""
    vg = VGContainer()
    vg.addPath(0, rgba(200, 30, 30, 200))
    vg.addPolygon([(4.3, 10.2), (4.3 + length, 10.7), (4.3 + length, 11.6), (4.3, 11.1)])
    vg.addPath(1, rgba(30, 30, 200, 150))
    vg.addPolygon([(6.1, 12.1), (6.1 + length, 12.2), (6.1 + length / 2, 13.9)])
    vg.addPolygon([(9.1, 12.3), (9.1 + length / 3, 12.5), (9.1 + length / 4, 13.7)])
    for spp in (8, 32):
        compare(rasterizer, vg, 2048, 24, spp, label="sliver%d" % length)


def test_text_like_scene_reduced(rasterizer):
    """The text-like workload's construction (4096x4096 canvas, glyph contours of quadratics
    with reversed holes, nonzero, 16 spp) with 2% of its glyphs: every stage bit-exact."""
    scene, W, H, spp = scanline.config_scene(3, scale=0.02)
    cnt, exact = compare(rasterizer, scene, W, H, spp, label="config3-2%")
    assert cnt["curve_fragment"] > 100000


def test_text_like_scene_full_image(rasterizer):
    """The full text-like workload (100,000 glyphs at 4096x4096, 16 spp): counts and the
    whole image against the OpenMP build of the oracle."""
    scene, W, H, spp = scanline.config_scene(3)
    print("parity case: config3 full", W, H, spp, flush=True)
    ref = pyoracle.run(scene, W, H, spp, srgb=True, ms=False, keep=False, omp=True)
    rasterizer.setSamples(spp)
    rasterizer.enableSRGBCorrection(True)
    rasterizer.setMultisampleOutput(False)
    rasterizer.resizeOutputBuffer(W, H)
    rasterizer.setInputTransform(np.eye(4, dtype=np.float32))
    rasterizer.dumpDebugData(False)
    rasterizer.addVg(scene)
    rasterizer.rasterize()
    cnt = rasterizer.elementNumber()
    for k, v in ref.counts().items():
        assert cnt[k] == v, (k, cnt[k], v)
    np.testing.assert_array_equal(rasterizer.output(), ref.get("image").reshape(H, W, 4))


@pytest.mark.parametrize("spp", [8, 32])
def test_rvg_sample_scene(rasterizer, spp):
    """A scene read by the RVG loader (tests/golden/sample.rvg: lines, quadratic, cubic,
    half-circle and rational arcs, an affine transform, both fill rules)."""
    from scanline import rvg
    vg, _, _ = rvg.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sample.rvg"))
    cnt, exact = compare(rasterizer, vg, 64, 48, spp, label="rvg sample")
    assert cnt["curve_fragment"] > 0
    assert exact == 1.0


def test_save_output_to_file(rasterizer, tmp_path):
    """saveOutputToFile: the PNG holds the image with rows flipped and alpha 255."""
    from scanline import png
    vg = VGContainer()
    vg.addPath(0, rgba(20, 200, 90, 128))
    vg.addPolygon([(3, 2), (40, 5), (30, 28)])
    render_gpu(rasterizer, vg.arrays(), 48, 32, 8, True, False, None)
    path = str(tmp_path / "out.png")
    rasterizer.saveOutputToFile(path)
    img = png.decode(open(path, "rb").read())
    np.testing.assert_array_equal(img, rasterizer.output(flip=True))
    assert (img[..., 3] == 255).all()
    np.testing.assert_array_equal(img[::-1, :, :3], rasterizer.output()[..., :3])


@pytest.mark.parametrize("suffix", ["vgb", "vgt"])
def test_scene_file_round_trip(rasterizer, tmp_path, suffix):
    """A generated scene saved as a .vgb / .vgt file and loaded back renders as the oracle
    renders the loaded scene; the binary file gives the original image exactly."""
    a = scanline.generate(scanline.scene.KIND_MIXED, 320, 240, 150, 0x51CE)
    fn = str(tmp_path / ("scene." + suffix))
    VGContainer.from_arrays(a, 320, 240).save(fn)
    vg = VGContainer.load(fn)
    cnt, exact = compare(rasterizer, vg, 320, 240, 16, label="scene file " + suffix)
    assert cnt["curve_fragment"] > 0
    assert exact == 1.0
    if suffix == "vgb":
        img = rasterizer.output().copy()
        render_gpu(rasterizer, a, 320, 240, 16)
        np.testing.assert_array_equal(rasterizer.output(), img)


@pytest.mark.parametrize("spp", [8, 32])
def test_svg_sample_scene(rasterizer, spp):
    """A scene read by the SVG loader (tests/golden/sample.svg: every path command, group
    transforms, colour forms, a polygon, arcs, a stroke colour over a gradient fill)."""
    from scanline import svg
    vg = svg.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sample.svg"))
    cnt, exact = compare(rasterizer, vg, vg.width, vg.height, spp, label="svg sample")
    assert cnt["curve_fragment"] > 0
    assert exact == 1.0


# Seeded sweep over scene kind, canvas shape (odd sizes, wide and tall), sample count,
# colour encoding and output mode; every case is checked stage by stage against the oracle.
SWEEP = [
    # kind, W, H, paths, seed, spp, srgb, ms
    (scanline.scene.KIND_CUBIC_CHAINS, 97, 61, 12, 1, 8, True, False),
    (scanline.scene.KIND_CUBIC_CHAINS, 640, 96, 60, 2, 32, False, False),
    (scanline.scene.KIND_MIXED, 128, 511, 90, 3, 16, True, False),
    (scanline.scene.KIND_MIXED, 333, 129, 120, 4, 32, True, True),
    (scanline.scene.KIND_MIXED, 64, 64, 200, 5, 8, False, True),
    (scanline.scene.KIND_GLYPHS, 257, 255, 300, 6, 16, False, False),
    (scanline.scene.KIND_GLYPHS, 1024, 33, 200, 7, 32, True, False),
    (scanline.scene.KIND_MIXED, 700, 700, 400, 8, 16, True, False),
    (scanline.scene.KIND_CUBIC_CHAINS, 2, 2, 3, 9, 32, True, False),
    (scanline.scene.KIND_MIXED, 1, 300, 40, 10, 8, True, False),
]


@pytest.mark.parametrize("case", SWEEP, ids=["sweep%d" % i for i in range(len(SWEEP))])
def test_seeded_sweep(rasterizer, case):
    kind, W, H, n, seed, spp, srgb, ms = case
    scene = scanline.generate(kind, W, H, n, 0x5EED0000 + seed)
    compare(rasterizer, scene, W, H, spp, srgb=srgb, ms=ms, label="sweep seed %d" % seed)


def test_wide_canvas_and_long_pieces(rasterizer):
    """Canvases wider than 64 tiles (the tile bitmaps then mark every tile) and monotone
    pieces with more than 1024 cuts (the merge path's LDS window, then the single-thread
    walk): a thin triangle across a 4200 x 2100 canvas, a cubic and a quadratic spanning it."""
    W, H = 4200, 2100
    vg = VGContainer()
    vg.addPath(0, rgba(200, 30, 60, 200))
    vg.addPolygon([(3.5, 5.25), (4195.0, 2090.5), (4190.0, 2060.0)])
    vg.addPath(1, rgba(20, 120, 220, 150))
    vg.addCubic((10.0, 2000.0), (1500.0, -900.0), (2800.0, 3000.0), (4150.0, 40.0))
    vg.addQuadratic((4150.0, 40.0), (2000.0, 1200.0), (10.0, 2000.0))
    vg.addPath(0, rgba(0, 0, 0, 255))
    vg.addPolygon([(100.0, 100.0), (4100.0, 130.0), (4100.0, 131.5), (100.0, 101.0)])
    cnt, exact = compare(rasterizer, vg, W, H, 16, label="wide canvas, long pieces")
    assert cnt["curve_fragment"] > 10000
    assert exact == 1.0


def test_long_cubic_chains(rasterizer):
    """Few paths whose cubics span the canvas: cubic solve chains of hundreds of cuts in the
    longest length bins, where a wave's lanes each run a long sequential bisection chain;
    every intersection t must equal the oracle's."""
    W, H = 2048, 2048
    rng = np.random.default_rng(0x5EED0C0B)
    vg = VGContainer()
    for k in range(24):
        vg.addPath(k % 2, rgba(int(rng.integers(0, 256)), int(rng.integers(0, 256)), int(rng.integers(0, 256)),
                               int(rng.integers(90, 256))))
        pts = [tuple(float(v) for v in rng.uniform(-100.0, 2148.0, 2)) for _ in range(7)]
        vg.addCubic(pts[0], pts[1], pts[2], pts[3])
        vg.addCubic(pts[3], pts[4], pts[5], pts[6])
        vg.addLinear(pts[6], pts[0])
    cnt, exact = compare(rasterizer, vg, W, H, 32, label="long cubic chains")
    assert cnt["curve_fragment"] > 20000
    assert exact == 1.0


def test_wide_sweep(rasterizer):
    scene = scanline.generate(scanline.scene.KIND_MIXED, 4500, 40, 60, 0x5EED0100)
    compare(rasterizer, scene, 4500, 40, 8, label="4500x40 mixed")
    scene = scanline.generate(scanline.scene.KIND_GLYPHS, 48, 4400, 80, 0x5EED0101)
    compare(rasterizer, scene, 48, 4400, 32, label="48x4400 glyphs")


def test_add_vg_rejects_invalid_scenes(rasterizer):
    """sl_add_vg validates the scene before anything is uploaded: each malformed scene is
    refused with SL_E_ARG and a message naming the first offending curve (the reference
    would read out of bounds instead), and a valid scene loads after the refusals."""
    from scanline.scene import SceneArrays

    def scene(cpos, ctype, cpath, nv=8, paths=2):
        v = np.arange(2 * nv, dtype=np.float32).reshape(-1, 2)
        return SceneArrays(v, cpos, ctype, cpath, [-1.0] * len(ctype), [0] * paths,
                           [rgba(1, 2, 3)] * paths)

    L, Q = scanline.CURVE_LINEAR, scanline.CURVE_QUADRATIC
    bad = [
        (scene([0, 2], [L, 0x07], [0, 1]), "invalid type"),
        (scene([0, 2], [L, L], [0, 2]), "path id"),
        (scene([0, 2], [L, L], [1, 0]), "non-decreasing"),
        (scene([0, 7], [L, Q], [0, 1]), "exceed n_vertices"),
        (scene([0, 1], [L, L], [0, 1]), "belongs to two curves"),
    ]
    rasterizer.setSamples(8)
    rasterizer.resizeOutputBuffer(16, 16)
    for s, msg in bad:
        with pytest.raises(scanline.ScanlineError) as e:
            rasterizer.addVg(s)
        assert msg in str(e.value), (msg, str(e.value))
        assert "curve 1" in str(e.value) or "vertex 1" in str(e.value), str(e.value)
    rasterizer.addVg(scene([0, 2], [L, Q], [0, 1]))
    rasterizer.rasterize()


def test_add_vg_validation_large_scene(rasterizer):
    """The same refusals on scenes large enough for the threaded validation: the reported
    curve is the first failing one in curve order, whichever chunk finds it."""
    from scanline.scene import SceneArrays
    n = 120000
    L = scanline.CURVE_LINEAR

    def scene(mod=None):
        cpos = np.arange(n, dtype=np.uint32) * 2
        ctype = np.full(n, L, dtype=np.uint8)
        cpath = (np.arange(n, dtype=np.uint32) // 4)
        if mod:
            mod(cpos, ctype, cpath)
        v = np.random.default_rng(1).uniform(0, 60, (2 * n, 2)).astype(np.float32)
        return SceneArrays(v, cpos, ctype, cpath, np.full(n, -1, np.float32), np.zeros(n // 4, np.uint8),
                           np.full(n // 4, rgba(9, 9, 9, 90), np.uint32))

    def shared(cpos, ctype, cpath):
        cpos[100001] = cpos[7] + 1         # shares a vertex with curve 7: reported at 100001
    def type_then_shared(cpos, ctype, cpath):
        ctype[90000] = 0x05
        cpos[100001] = cpos[7] + 1
    def shared_then_type(cpos, ctype, cpath):
        cpos[50001] = cpos[50000] + 1      # reported at 50001, before the bad type
        ctype[110000] = 0x05
    def order(cpos, ctype, cpath):
        cpath[70000] = 0

    cases = [(shared, "vertex %d belongs to two curves" % (2 * 7 + 1)),
             (type_then_shared, "curve 90000 has invalid type 0x5"),
             (shared_then_type, "vertex %d belongs to two curves" % (2 * 50000 + 1)),
             (order, "curve 70000: curve_path_id must be non-decreasing")]
    rasterizer.setSamples(8)
    rasterizer.resizeOutputBuffer(64, 64)
    for mod, msg in cases:
        with pytest.raises(scanline.ScanlineError) as e:
            rasterizer.addVg(scene(mod))
        assert msg in str(e.value), (msg, str(e.value))
    good = scene()
    cnt, exact = compare(rasterizer, good, 64, 64, 8, label="large valid scene")
    assert exact == 1.0


def test_rerender_unchanged_scene(rasterizer):
    """Frames after the first of an unchanged scene are queued without host reads.  Each
    one renders the image again: it is poisoned between frames and must come back equal
    to the oracle's.  A changed setting makes the next frame a full one again."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime libscanline.so is bound to
    W, H, spp = 256, 192, 16
    scene = scanline.generate(scanline.scene.KIND_MIXED, W, H, 80, 0x5EED0200)
    r = rasterizer
    r.setSamples(spp)
    r.enableSRGBCorrection(True)
    r.setMultisampleOutput(False)
    r.resizeOutputBuffer(W, H)
    r.setInputTransform(np.eye(4, dtype=np.float32))
    r.dumpDebugData(False)
    r.enableStepTiming(False)
    r.addVg(scene)

    def poison(v):
        p, pitch = r.outputDevicePtr()
        assert hip.hipMemset(ctypes.c_void_p(p), ctypes.c_int(v), ctypes.c_size_t(pitch * H)) == 0

    def check(img, transform=None):
        ref = pyoracle.run(scene, W, H, spp, srgb=True, transform=transform).get("image").reshape(H, W, 4)
        assert np.abs(img.astype(np.int32) - ref.astype(np.int32)).max() <= 1

    r.rasterize()
    first = r.output()
    check(first)
    counts = r.elementNumber()
    for k in range(3):
        poison(17 + k)
        r.rasterize()
        np.testing.assert_array_equal(r.output(), first)
        assert r.elementNumber() == counts
    m = np.eye(4, dtype=np.float32)
    m[0, 0], m[1, 1], m[3, 0] = 0.8, 1.15, 9.5
    r.setInputTransform(m)
    r.rasterize()
    moved = r.output()
    check(moved, m)
    assert not np.array_equal(moved, first)
    poison(99)
    r.rasterize()
    np.testing.assert_array_equal(r.output(), moved)


def test_rerender_on_a_torch_stream(rasterizer):
    """Frames queued on a caller's own stream: the read-back waits for the frame through
    the event recorded on that stream, and torch work queued after it sees the image."""
    import torch
    W, H, spp = 192, 160, 8
    scene = scanline.generate(scanline.scene.KIND_CUBIC_CHAINS, W, H, 30, 0x5EED0300)
    r = rasterizer
    r.setSamples(spp)
    r.enableSRGBCorrection(True)
    r.setMultisampleOutput(False)
    r.resizeOutputBuffer(W, H)
    r.setInputTransform(np.eye(4, dtype=np.float32))
    r.dumpDebugData(False)
    r.enableStepTiming(False)
    r.addVg(scene)
    ref = pyoracle.run(scene, W, H, spp, srgb=True).get("image").reshape(H, W, 4)
    st = torch.cuda.Stream()
    ptr, pitch = None, None
    with torch.cuda.stream(st):
        for k in range(4):
            r.rasterize(st.cuda_stream)
            if k == 3:
                # a device-side copy of the image, queued after the frame on the same stream
                ptr, pitch = r.outputDevicePtr()
                dev = torch.empty((H, W, 4), dtype=torch.uint8, device="cuda")
                import ctypes
                hip = ctypes.CDLL("libamdhip64.so.7")
                assert hip.hipMemcpyAsync(ctypes.c_void_p(dev.data_ptr()), ctypes.c_void_p(ptr),
                                          ctypes.c_size_t(pitch * H), ctypes.c_int(3),
                                          ctypes.c_void_p(st.cuda_stream)) == 0
        img = r.output()
    st.synchronize()
    assert np.abs(img.astype(np.int32) - ref.astype(np.int32)).max() <= 1
    np.testing.assert_array_equal(dev.cpu().numpy(), img)


def test_queued_frames_then_new_scene_and_second_stream(rasterizer):
    """Frames still queued on a caller's non-blocking stream share the handle's buffers:
    uploading the next scene waits for them, and a frame issued on a second stream is
    ordered after them, so every read-back is the scene last added."""
    import torch
    W, H, spp = 512, 384, 16
    a = scanline.generate(scanline.scene.KIND_MIXED, W, H, 400, 0x5EED0401)
    b = scanline.generate(scanline.scene.KIND_CUBIC_CHAINS, W, H, 60, 0x5EED0402)
    r = rasterizer
    r.setSamples(spp)
    r.enableSRGBCorrection(True)
    r.setMultisampleOutput(False)
    r.resizeOutputBuffer(W, H)
    r.setInputTransform(np.eye(4, dtype=np.float32))
    r.dumpDebugData(False)
    r.enableStepTiming(False)
    ref_b = pyoracle.run(b, W, H, spp, srgb=True).get("image").reshape(H, W, 4)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    r.addVg(a)
    r.rasterize(s1.cuda_stream)            # full frame: sizes the buffers
    for _ in range(6):
        r.rasterize(s1.cuda_stream)        # queued, not waited for
    r.addVg(b)                             # overwrites the scene buffers
    r.rasterize(s1.cuda_stream)
    img = r.output()
    assert np.abs(img.astype(np.int32) - ref_b.astype(np.int32)).max() <= 1
    for _ in range(4):
        r.rasterize(s1.cuda_stream)
    r.rasterize(s2.cuda_stream)            # same buffers, another stream
    img2 = r.output()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(img2, img)


def test_image_dump_after_queued_frames(rasterizer):
    """A parity dump read straight after frames queued on a caller's stream sees the
    finished image: the dump waits for the queued frames."""
    import torch
    W, H, spp = 256, 224, 8
    scene = scanline.generate(scanline.scene.KIND_MIXED, W, H, 150, 0x5EED0501)
    r = rasterizer
    r.setSamples(spp)
    r.enableSRGBCorrection(True)
    r.setMultisampleOutput(False)
    r.resizeOutputBuffer(W, H)
    r.setInputTransform(np.eye(4, dtype=np.float32))
    r.dumpDebugData(False)
    r.enableStepTiming(False)
    r.addVg(scene)
    ref = pyoracle.run(scene, W, H, spp, srgb=True).get("image").reshape(H, W, 4)
    st = torch.cuda.Stream()
    for _ in range(6):
        r.rasterize(st.cuda_stream)
    img = r.dump("image").reshape(H, W, 4)
    assert np.abs(img.astype(np.int32) - ref.astype(np.int32)).max() <= 1
    torch.cuda.synchronize()


def test_output_after_resize_needs_a_new_frame():
    """After resizeOutputBuffer the last image has the old size: reading it (host copy,
    device pointer, image dump) fails with SL_E_STATE until the next frame, which then
    reads back at the new size."""
    r = scanline.VGRasterizer(0)
    try:
        scene = scanline.generate(scanline.scene.KIND_MIXED, 96, 64, 20, 0x5EED0600)
        render_gpu(r, scene, 96, 64, 8)
        r.dumpDebugData(False)
        assert r.output().shape == (64, 96, 4)
        for w, h in ((160, 120), (40, 30)):
            r.resizeOutputBuffer(w, h)
            with pytest.raises(scanline.ScanlineError) as e:
                r.output()
            assert e.value.code == scanline._abi.SL_E_STATE
            with pytest.raises(scanline.ScanlineError):
                r.outputDevicePtr()
            with pytest.raises(scanline.ScanlineError):
                r.dump("image")
            r.rasterize()
            ref = pyoracle.run(scene, w, h, 8, srgb=True).get("image").reshape(h, w, 4)
            np.testing.assert_array_equal(r.output(), ref)
            assert r.outputDevicePtr()[1] == w * 4
    finally:
        r.close()


def test_viewport_present():
    """viewport / resizeViewport place the image in a window buffer one pixel for one
    (the reference's final blit, ras_base.cpp:1170-1176), clipped to both rectangles."""
    import torch
    r = scanline.VGRasterizer(0)
    try:
        scene = scanline.generate(scanline.scene.KIND_MIXED, 80, 60, 30, 0x5EED0650)
        render_gpu(r, scene, 80, 60, 8)
        img = r.output()
        for (vx, vy, vw, vh), (WW, WH) in [((10, 5, 0, 0), (128, 96)), ((-7, -3, 50, 40), (64, 48)),
                                          ((100, 90, 80, 60), (120, 100))]:
            if vw:
                r.viewport(vx, vy, vw, vh)
            else:
                r.viewport(vx, vy, 0, 0)
            pitch = WW * 4 + 16
            buf = torch.full((WH * pitch,), 7, dtype=torch.uint8, device="cuda")
            r.present(buf.data_ptr(), pitch, WW, WH, torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            got = buf.cpu().numpy().reshape(WH, pitch)[:, :WW * 4].reshape(WH, WW, 4)
            want = np.full((WH, WW, 4), 7, np.uint8)
            w, h = (vw or 80), (vh or 60)
            for y in range(h):
                for x in range(w):
                    dx, dy = vx + x, vy + y
                    if 0 <= dx < WW and 0 <= dy < WH and x < 80 and y < 60:
                        want[dy, dx] = img[y, x]
            np.testing.assert_array_equal(got, want, err_msg=str((vx, vy, vw, vh)))
        r.resizeViewport(20, 10)
        r.viewport(0, 0, 20, 10)
    finally:
        r.close()


def test_transform_changes_queue_without_status_reads():
    """A frame under a changed transform reads its three sizes back in one round trip and
    returns without waiting for the frame; its element counts and error flags come back
    when the frame is settled.  Frames under three transforms queued back to back, then
    replays of the last: counts and image are the oracle's for the last transform."""
    import time
    scene = scanline.generate(scanline.scene.KIND_MIXED, 640, 480, 400, 0x7A5F)
    arrs = scene.arrays() if hasattr(scene, "arrays") else scene
    r = scanline.VGRasterizer(0)
    try:
        r.setSamples(16)
        r.enableSRGBCorrection(True)
        r.setMultisampleOutput(False)
        r.resizeOutputBuffer(640, 480)
        r.addVg(arrs)
        mats = []
        for k in range(3):
            m = np.eye(4, dtype=np.float32)
            m[0, 0], m[1, 1] = 0.8 + 0.1 * k, 1.05 - 0.05 * k
            m[3, 0], m[3, 1] = 5.25 * k + 0.5, -3.75 * k + 0.25
            mats.append(m)
        for m in mats:
            print("queued transform frame", float(m[0, 0]), flush=True)
            r.setInputTransform(m)
            t0 = time.perf_counter()
            r.rasterize()
            print("  call returned after %.2f ms" % (1e3 * (time.perf_counter() - t0)), flush=True)
        r.rasterize()
        r.rasterize()
        ref = pyoracle.run(arrs, 640, 480, 16, srgb=True, ms=False, transform=mats[-1], keep=False)
        got = r.elementNumber()
        want = ref.counts()
        for k in ("curve_fragment", "merged_fragment", "span"):
            assert got[k] == want[k], (k, got, want)
        np.testing.assert_array_equal(r.output(), ref.get("image").reshape(480, 640, 4))
    finally:
        r.close()


def test_bucket_beyond_65536_fragments(rasterizer):
    """One path whose single cell row holds about 160k fragments (80 near-horizontal edges
    across a 4096-px row): the largest-bucket resolve sorts 64-bit (cell x, index) keys, so
    any bucket size renders, and every stage matches the oracle."""
    vg = VGContainer()
    vg.addPath(1, rgba(30, 120, 200, 200))
    pts = []
    for k in range(40):
        y = 10.1 + 0.04 * k
        pts.append((0.5, y))
        pts.append((4095.5, y + 0.02))
    vg.addPolygon(pts)
    vg.addPath(0, rgba(200, 60, 20))
    vg.addPolygon([(100.0, 2.0), (300.0, 2.0), (300.0, 14.0), (100.0, 14.0)])
    compare(rasterizer, vg, 4096, 16, 8, label="huge bucket")
