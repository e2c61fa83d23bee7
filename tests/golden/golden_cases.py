"""Cases and hashing shared by the golden-vector script and the tests.

Each case is a seeded scene rendered by the CPU oracle; the committed fixture holds the
sha256 of every parity array (the reference's dump layouts: intersections, fragment keys,
path id | sign, scan winding, raw masks, draw list, resolved masks, image) and the counts.
"""
import hashlib

import numpy as np

NW = {8: 1, 16: 2, 32: 4}


def cases():
    import scanline
    from scanline import scene as sc
    yield "config1_512_cubic_chains_8spp", scanline.config_scene(1)[0], 512, 512, 8, True
    yield "mixed_384x320_32spp_srgb", scanline.generate(sc.KIND_MIXED, 384, 320, 300, sc.SEED_BASE ^ 2), 384, 320, 32, True
    yield "mixed_384x320_16spp_linear", scanline.generate(sc.KIND_MIXED, 384, 320, 300, sc.SEED_BASE ^ 2), 384, 320, 16, False
    yield "glyphs_512_16spp", scanline.generate(sc.KIND_GLYPHS, 512, 512, 400, sc.SEED_BASE ^ 3), 512, 512, 16, True


def digest(a):
    a = np.ascontiguousarray(a)
    return hashlib.sha256(a.tobytes()).hexdigest()


def stage_arrays(get, spp):
    """The parity arrays from a getter (oracle result or GPU dump) in comparable form."""
    nw = NW[spp]
    return {
        "ts": get("ts"),
        "frag_yx": get("frag_yx"),
        "frag_pid": get("frag_pid"),
        "frag_wn": get("frag_wn"),
        "frag_mask": np.ascontiguousarray(get("frag_mask")[:, :nw]),
        "draw_list": get("draw_list"),
        "merged_mask": np.ascontiguousarray(get("merged_mask")[:, :nw]),
    }
