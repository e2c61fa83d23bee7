"""The oracle against its committed golden vectors (CPU).

tests/golden/oracle_hashes.json holds, for four seeded scenes, the sha256 of every parity
array the oracle produces (made by tests/golden/make_golden.py).  Re-running the oracle
must reproduce them bit for bit; the GPU test of the same file compares the renderer's
dumps with the same hashes without running the oracle.
"""
import json
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import pyoracle  # noqa: E402
from golden_cases import cases, digest, stage_arrays  # noqa: E402

GOLDEN = json.load(open(os.path.join(HERE, "golden", "oracle_hashes.json")))
CASES = list(cases())


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_oracle_reproduces_golden_vectors(case):
    name, scene, W, H, spp, srgb = case
    g = GOLDEN[name]
    assert (g["width"], g["height"], g["spp"], g["srgb"]) == (W, H, spp, srgb)
    ref = pyoracle.run(scene.arrays() if hasattr(scene, "arrays") else scene, W, H, spp, srgb=srgb, omp=True)
    assert ref.counts() == g["counts"]
    assert int(ref.get("pixel_count")[0]) == g["pixel_count"]
    got = {k: digest(v) for k, v in stage_arrays(ref.get, spp).items()}
    got["image"] = digest(ref.get("image"))
    assert got == g["sha256"]
