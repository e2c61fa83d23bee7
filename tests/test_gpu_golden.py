"""GPU: the renderer's parity dumps against the committed golden vectors.

The same hashes tests/test_golden.py checks the oracle against (tests/golden/
oracle_hashes.json): intersections, fragment keys, path id | sign, scan winding, raw and
resolved sample masks, draw list, counts, pixel count and the image, bit for bit.
"""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from golden_cases import cases, digest, stage_arrays  # noqa: E402

pytestmark = pytest.mark.gpu
GOLDEN = json.load(open(os.path.join(HERE, "golden", "oracle_hashes.json")))
CASES = list(cases())


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_gpu_matches_golden_vectors(rasterizer, case):
    name, scene, W, H, spp, srgb = case
    g = GOLDEN[name]
    print("golden case:", name, flush=True)
    r = rasterizer
    r.setSamples(spp)
    r.enableSRGBCorrection(srgb)
    r.setMultisampleOutput(False)
    r.resizeOutputBuffer(W, H)
    r.setInputTransform(np.eye(4, dtype=np.float32))
    r.dumpDebugData(True)
    r.addVg(scene)
    r.rasterize()
    cnt = r.elementNumber()
    assert {k: cnt[k] for k in g["counts"]} == g["counts"]
    assert r.pixelCount() == g["pixel_count"]
    got = {k: digest(v) for k, v in stage_arrays(r.dump, spp).items()}
    got["image"] = digest(r.dump("image"))
    assert got == g["sha256"]
