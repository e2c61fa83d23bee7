"""The PNG writer behind VGRasterizer.saveOutputToFile (CPU)."""
import zlib

import numpy as np

from scanline import png


def test_roundtrip_and_structure(tmp_path):
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, (17, 23, 4), dtype=np.uint8)
    path = tmp_path / "a.png"
    png.write(str(path), img)
    data = path.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    assert data[12:16] == b"IHDR" and int.from_bytes(data[16:20], "big") == 23
    assert int.from_bytes(data[20:24], "big") == 17
    # every chunk's CRC covers its type and data
    pos = 8
    while pos < len(data):
        n = int.from_bytes(data[pos:pos + 4], "big")
        crc = int.from_bytes(data[pos + 8 + n:pos + 12 + n], "big")
        assert crc == zlib.crc32(data[pos + 4:pos + 8 + n]) & 0xFFFFFFFF
        pos += 12 + n
    np.testing.assert_array_equal(png.decode(data), img)
