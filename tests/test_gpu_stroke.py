"""GPU: scenes converted by stroke-to-fill render as the oracle renders them.

Stroke-to-fill (scanline/stroke.py) runs on the host and hands the renderer fill geometry:
butt-capped, bevel-joined outlines of offset curves.  The converted scenes have the curve
shapes the conversion produces (many short offset pieces, thin bands, self-overlapping
joins), and are checked stage by stage against the oracle: `sample.svg` converted (its
gradient-filled path is then drawn with its gradient, with the stroke outline over it) and
a generated scene with every path stroked, filled or not, in coverage and multisample
output.
"""
import os

import numpy as np
import pytest

import pyoracle
import scanline
from scanline import svg
from scanline.scene import FT_NONE, ST_COLOR

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def check(r, arrays, W, H, spp, ms, srgb=True):
    print("stroke case:", W, H, spp, "ms" if ms else "cov", flush=True)
    ref = pyoracle.run(arrays, W, H, spp, srgb=srgb, ms=ms)
    r.setSamples(spp)
    r.enableSRGBCorrection(srgb)
    r.setMultisampleOutput(ms)
    r.resizeOutputBuffer(W, H)
    r.setInputTransform(np.eye(4, dtype=np.float32))
    r.dumpDebugData(True)
    r.addVg(arrays)
    r.rasterize()
    assert r.elementNumber() == ref.counts()
    for name in ("ts", "frag_yx", "frag_pid", "frag_wn", "frag_mask", "draw_list", "merged_mask"):
        np.testing.assert_array_equal(r.dump(name), ref.get(name), err_msg=name)
    np.testing.assert_array_equal(r.output(), ref.get("image").reshape(H, W, 4))
    r.dumpDebugData(False)
    return r.output()


@pytest.mark.parametrize("spp,ms", [(32, False), (16, True)])
def test_sample_svg_stroke_to_fill(rasterizer, spp, ms):
    vg = svg.load(os.path.join(HERE, "golden", "sample.svg"), stroke_to_fill=True)
    img = check(rasterizer, vg.arrays(), vg.width, vg.height, spp, ms)
    plain = svg.load(os.path.join(HERE, "golden", "sample.svg"))
    ref = pyoracle.run(plain.arrays(), plain.width, plain.height, spp, ms=ms).get("image")
    assert not np.array_equal(img.reshape(-1), ref)   # the stroke now has its own geometry


def stroked_scene(W, H, n, seed):
    a = scanline.generate(scanline.scene.KIND_MIXED, W, H, n, seed)
    vg = scanline.VGContainer.from_arrays(a, W, H)
    rng = np.random.default_rng(seed)
    for p in range(vg.pathNumber()):
        vg._stype[p] = ST_COLOR
        vg._scolor[p] = scanline.rgba(*rng.integers(0, 256, 3), int(rng.integers(100, 256)))
        vg._swidth[p] = float(rng.uniform(0.5, 4.0))
        if rng.random() < 0.5:
            vg._ftype[p] = FT_NONE
    return vg.strokeToFill()


@pytest.mark.parametrize("spp,ms", [(32, False), (8, True)])
def test_generated_stroked_scene(rasterizer, spp, ms):
    vg = stroked_scene(320, 256, 60, 0x57E0 + spp)
    assert vg.pathNumber() > 60
    check(rasterizer, vg.arrays(), 320, 256, spp, ms)
