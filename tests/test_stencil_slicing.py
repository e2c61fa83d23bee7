"""The bit-sliced cell-group evaluation of the resolve kernels equals the byte-lane stencil.

The resolve kernels evaluate a merged fragment's per-sample fill with the reference's
byte-lane counters (k_gen_stencil_cuda_sm, ras_cut_mask_comb_scanline.cu:665-716) only for
large cell groups or large windings; groups of at most 7 fragments with |w| <= 120 use four
two's-complement bit planes of the per-sample count instead.  This CPU test checks that
claim on random masks, signs and windings over the whole admitted range (and that it fails
outside it, so the bound is not vacuous).  The GPU parity tests compare the kernels
themselves with the oracle's byte-lane stencil.
"""
import numpy as np

M32 = np.uint64(0xFFFFFFFF)


def byte_lane(w, masks, negs, rule):
    """Reference arithmetic on one 32-bit mask word: 8 counter words, bits b, b+8, b+16, b+24."""
    w = np.asarray(w, dtype=np.int64)
    init = (w.astype(np.uint64) * np.uint64(0x01010101)) & M32  # w * 0x01010101 mod 2^32
    acc = [init.copy() for _ in range(8)]
    for m, ng in zip(masks, negs):
        for b in range(8):
            a = (m.astype(np.uint64) >> np.uint64(b)) & np.uint64(0x01010101)
            acc[b] = np.where(ng, (acc[b] - a) & M32, (acc[b] + a) & M32)
    out = np.zeros_like(masks[0], dtype=np.uint64)
    for b in range(8):
        v = (np.uint64(0x80808080) + acc[b]) & np.uint64(0x7F7F7F7F)
        if rule == 0:
            out |= ((v + np.uint64(0x7F7F7F7F)) & np.uint64(0x80808080)) >> np.uint64(7 - b)
        else:
            out |= (v & np.uint64(0x01010101)) << np.uint64(b)
    return out & M32


def sliced(w, masks, negs, rule):
    """Four bit planes of d_j = sum of +/-1 at set bits; nonzero: d != -w, even-odd: d ^ w."""
    w = np.asarray(w, dtype=np.int64)
    d = [np.zeros_like(masks[0], dtype=np.uint64) for _ in range(4)]
    for m, ng in zip(masks, negs):
        nm = np.where(ng, M32, np.uint64(0))
        c = m.astype(np.uint64)
        for p in range(4):
            t = (d[p] ^ nm) & c
            d[p] ^= c
            c = t
    if rule == 1:
        return d[0] ^ np.where(w & 1, M32, np.uint64(0))
    t = -w
    hit = (t >= -8) & (t <= 7)
    ne = np.zeros_like(d[0])
    for p in range(4):
        ne |= d[p] ^ np.where((t >> p) & 1, M32, np.uint64(0))
    return np.where(hit, ne, M32)


def _case(rng, n, g, wlo, whi):
    w = rng.integers(wlo, whi + 1, n)
    masks = [rng.integers(0, 1 << 32, n, dtype=np.uint64) & rng.integers(0, 1 << 32, n, dtype=np.uint64)
             for _ in range(g)]
    negs = [rng.integers(0, 2, n).astype(bool) for _ in range(g)]
    return w, masks, negs


def test_sliced_equals_byte_lane_in_range():
    rng = np.random.default_rng(11)
    for g in range(1, 8):
        for rule in (0, 1):
            w, masks, negs = _case(rng, 40000, g, -120, 120)
            np.testing.assert_array_equal(sliced(w, masks, negs, rule), byte_lane(w, masks, negs, rule))


def test_small_windings_exhaustive_single_fragment():
    # every winding in range, one fragment, all 2^8 patterns of one byte lane group
    w = np.repeat(np.arange(-120, 121), 256)
    m = np.tile(np.arange(256, dtype=np.uint64), 241)
    m = m | (m << np.uint64(8)) | (m << np.uint64(16)) | (m << np.uint64(24))
    for neg in (False, True):
        negs = [np.full(m.shape, neg)]
        for rule in (0, 1):
            np.testing.assert_array_equal(sliced(w, [m], negs, rule), byte_lane(w, [m], negs, rule))


def test_bound_is_needed():
    # near the int8 limits the byte lanes borrow into each other and the planes disagree
    rng = np.random.default_rng(5)
    w, masks, negs = _case(rng, 20000, 7, 121, 127)
    negs = [np.zeros_like(n) for n in negs]  # all +1: counts reach w + 7 > 127
    assert (sliced(w, masks, negs, 0) != byte_lane(w, masks, negs, 0)).any()
