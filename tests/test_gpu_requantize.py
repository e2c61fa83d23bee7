"""GPU: the compositor's 8-bit framebuffer requantisation is exact.

The coverage compositor re-quantises every blended channel without a table search:
UNORM8 by one exact fma sign test around an estimate, sRGB8 by bracketing an estimate
from the transfer function between two thresholds.  These tests prove it equal to the
oracle's definition (largest code k with value >= thr[k], ras_base.cpp:1044-1060 blending
into an RGBA8 / SRGB8_ALPHA8 attachment): exhaustively over every float in [0, 1] against
the library's threshold tables, and on the thresholds themselves (+-3 ulps) plus random
values against the oracle's tables.
"""
import numpy as np
import pytest

import pyoracle

pytestmark = pytest.mark.gpu

ONE_BITS = int(np.float32(1.0).view(np.uint32))


def oracle_tables(srgb):
    L = pyoracle.lib()
    dec = np.zeros(256, np.float32)
    thr = np.zeros(256, np.float32)
    L.oracle_color_tables(int(srgb), dec.ctypes.data, thr.ctypes.data)
    return dec, thr


@pytest.mark.parametrize("srgb,channel", [(True, 0), (False, 0), (True, 1)])
def test_requantize_exhaustive(rasterizer, srgb, channel):
    rasterizer.enableSRGBCorrection(srgb)
    print("requantize sweep", "srgb" if srgb else "lrgb", "channel", channel, flush=True)
    bad = rasterizer.requantizeSweep(0, ONE_BITS + 4096, channel)
    assert bad == 0


@pytest.mark.parametrize("srgb,channel", [(True, 0), (False, 0), (False, 1)])
def test_requantize_matches_oracle(rasterizer, srgb, channel):
    rasterizer.enableSRGBCorrection(srgb)
    dec, thr = oracle_tables(srgb if channel == 0 else False)
    rng = np.random.default_rng(7 + channel + 2 * srgb)
    edges = thr[1:].view(np.uint32).astype(np.int64)
    near = (edges[:, None] + np.arange(-3, 4)[None, :]).reshape(-1).astype(np.uint32).view(np.float32)
    vals = np.concatenate([near, rng.random(1 << 20, dtype=np.float32), dec,
                           np.array([0.0, 1.0, np.nextafter(np.float32(1), np.float32(2))], np.float32)])
    codes, got = rasterizer.requantize(vals, channel)
    ref = np.searchsorted(thr[1:], vals, side="right")
    np.testing.assert_array_equal(codes, ref.astype(np.uint8))
    np.testing.assert_array_equal(got.view(np.uint32), dec[ref].view(np.uint32))
