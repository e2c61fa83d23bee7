"""Minimal PNG writer for RGBA8 images (the output side of `saveOutputToFile`,
gpu-scanline/src/rasterizer/shared/ras_base.cpp:1221-1277, which writes through stb)."""
import struct
import zlib

import numpy as np


def _chunk(kind, data):
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def encode(rgba):
    """PNG bytes of an (H, W, 4) uint8 array, rows top to bottom, filter 0 on every row."""
    a = np.ascontiguousarray(rgba, dtype=np.uint8)
    if a.ndim != 3 or a.shape[2] != 4:
        raise ValueError("expected an (H, W, 4) uint8 image")
    h, w = a.shape[:2]
    raw = np.zeros((h, 1 + 4 * w), dtype=np.uint8)
    raw[:, 1:] = a.reshape(h, 4 * w)
    ihdr = struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)
    return (b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr) + _chunk(b"IDAT", zlib.compress(raw.tobytes(), 6))
            + _chunk(b"IEND", b""))


def write(path, rgba):
    with open(path, "wb") as fh:
        fh.write(encode(rgba))


def decode(data):
    """Inverse of `encode` for the files it writes (filter 0 rows, one IDAT chunk)."""
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG")
    pos, chunks = 8, {}
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        kind = data[pos + 4:pos + 8]
        chunks.setdefault(kind, b"")
        chunks[kind] += data[pos + 8:pos + 8 + n]
        pos += 12 + n
    w, h, depth, ctype = struct.unpack(">IIBB", chunks[b"IHDR"][:10])
    if (depth, ctype) != (8, 6):
        raise ValueError("only RGBA8 is read")
    raw = np.frombuffer(zlib.decompress(chunks[b"IDAT"]), dtype=np.uint8).reshape(h, 1 + 4 * w)
    if raw[:, 0].any():
        raise ValueError("only filter 0 rows are read")
    return raw[:, 1:].reshape(h, w, 4).copy()
