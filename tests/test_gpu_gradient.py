"""GPU: gradient paint against the oracle.

The renderer evaluates gradient paints per pixel in both compositors (coverage and
multisample output); the oracle restates the output shaders' evaluation
(output_32.vert.glsl:55-97, output_32.frag.glsl:45-51) in the same float operations, so the
images must be identical.  Scenes: the RVG gradient fixture (linear and radial paints, a
solid path between them, a paint transform) in every output mode, sample count and colour
encoding; the fixture under a projective input transform; and a generated scene with random
linear and radial paints on a third of its paths.  Batches refuse gradient paints.
"""
import os

import numpy as np
import pytest

import pyoracle
import scanline
from scanline import rvg
from scanline.gradient import GradientTextures
from scanline.scene import SceneArrays

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "gradient.rvg")


def render_both(r, arrays, W, H, spp, srgb, ms, transform=None):
    print("gradient case:", W, H, spp, "srgb" if srgb else "lrgb", "ms" if ms else "cov", flush=True)
    ref = pyoracle.run(arrays, W, H, spp, srgb=srgb, ms=ms, transform=transform, keep=False)
    r.setSamples(spp)
    r.enableSRGBCorrection(srgb)
    r.setMultisampleOutput(ms)
    r.resizeOutputBuffer(W, H)
    r.setInputTransform(np.eye(4, dtype=np.float32) if transform is None else transform)
    r.dumpDebugData(False)
    r.addVg(arrays)
    r.rasterize()
    assert r.elementNumber() == ref.counts()
    return r.output(), ref.get("image").reshape(H, W, 4)


@pytest.mark.parametrize("spp", [8, 16, 32])
@pytest.mark.parametrize("ms", [False, True])
@pytest.mark.parametrize("srgb", [False, True])
def test_gradient_fixture(rasterizer, spp, ms, srgb):
    vg, _, _ = rvg.load(FIXTURE, srgb=srgb)
    got, want = render_both(rasterizer, vg.arrays(), 96, 64, spp, srgb, ms)
    np.testing.assert_array_equal(got, want)
    # the paints are drawn (not a blank or solid-only image)
    assert len(np.unique(got[8:26, 10:38, 0])) > 10


@pytest.mark.parametrize("ms", [False, True])
def test_gradient_under_a_projective_transform(rasterizer, ms):
    vg, _, _ = rvg.load(FIXTURE, srgb=True)
    m = np.eye(4, dtype=np.float32)
    m[0, 0], m[1, 1], m[1, 0] = 1.3, 0.9, 0.1
    m[3, 0], m[3, 1] = -6.5, 3.25
    m[0, 3] = 1.5e-3
    m[3, 2] = 20.0  # depth: the perspective scaling of the transform kernel
    got, want = render_both(rasterizer, vg.arrays(), 128, 96, 32, True, ms, transform=m)
    np.testing.assert_array_equal(got, want)


def gradient_scene(W, H, n, seed, srgb):
    """A generated mixed scene with random gradient paints on about a third of its paths."""
    a = scanline.generate(scanline.scene.KIND_MIXED, W, H, n, seed)
    rng = np.random.default_rng(seed)
    g = GradientTextures(srgb=srgb)
    fill = a.fill_color.copy()
    for p in range(a.n_paths):
        if rng.random() < 0.35:
            k = int(rng.integers(2, 5))
            offs = np.sort(rng.random(k))
            offs[0], offs[-1] = 0.0, 1.0
            stops = [(float(o),) + tuple(float(v) for v in rng.random(4)) for o in offs]
            x0, y0 = rng.uniform(0, W), rng.uniform(0, H)
            tr = [1.0 + rng.normal(0, 0.2), rng.normal(0, 0.2), rng.normal(0, 10),
                  rng.normal(0, 0.2), 1.0 + rng.normal(0, 0.2), rng.normal(0, 10)]
            if rng.random() < 0.5:
                g.add("linear", x0, y0, x0 + rng.uniform(-200, 200), y0 + rng.uniform(-200, 200), 0, stops,
                      float(rng.uniform(0.3, 1)), tr)
            else:
                g.add("radial", x0, y0, x0 + rng.normal(0, 10), y0 + rng.normal(0, 10), rng.uniform(5, 150), stops,
                      float(rng.uniform(0.3, 1)), tr)
            fill[p] = p + 1
        else:
            g.add_solid()
    ramp, table = g.textures()
    return SceneArrays(a.vertex, a.curve_vertex_pos, a.curve_type, a.curve_path_id, a.curve_arc_w, a.fill_rule,
                       fill, ramp, table)


@pytest.mark.parametrize("spp,ms,srgb", [(32, False, True), (16, False, False), (8, True, True), (32, True, False)])
def test_generated_scene_with_gradients(rasterizer, spp, ms, srgb):
    s = gradient_scene(512, 384, 400, 0x6AD1E47 + spp, srgb)
    got, want = render_both(rasterizer, s, 512, 384, spp, srgb, ms)
    np.testing.assert_array_equal(got, want)


def test_batches_refuse_gradient_paints():
    s = gradient_scene(64, 64, 20, 5, True)
    plain = scanline.generate(scanline.scene.KIND_CUBIC_CHAINS, 64, 64, 5, 1)
    with pytest.raises(scanline.ScanlineError) as e:
        scanline.BatchRasterizer([plain, s], [(64, 64), (64, 64)], [8, 8])
    assert e.value.code == scanline._abi.SL_E_UNSUPPORTED


@pytest.mark.parametrize("spp,ms,srgb", [(32, False, True), (8, False, False), (16, True, True), (32, True, False)])
def test_svg_gradient_fixture(rasterizer, spp, ms, srgb):
    from scanline import svg
    vg = svg.load(os.path.join(HERE, "golden", "gradient.svg"), srgb=srgb)
    got, want = render_both(rasterizer, vg.arrays(), 96, 64, spp, srgb, ms)
    np.testing.assert_array_equal(got, want)
    assert len(np.unique(got[40:56, 14:86, 2])) > 10
