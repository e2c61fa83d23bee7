"""The `.vgt` / `.vgb` scene files (VGContainer::saveTxt / loadTxt / saveBinary /
loadBinary, gpu-scanline/src/vg_container.cpp:1692-1906).

The known-answer files below are written out by hand from the reference's printf / fwrite
calls, field by field, not by the module under test: the reader must give the fields
back, and the writer must give the same bytes.  Decimal rounding of `.vgt` vertices is
checked against exact rational arithmetic, including the cases where a conversion through
double rounds twice.
"""
import struct
from fractions import Fraction

import numpy as np
import pytest

import scanline
from scanline import FT_COLOR, FT_GRADIENT, VGContainer, rgba, vgfile

L, Q, C, R = scanline.CURVE_LINEAR, scanline.CURVE_QUADRATIC, scanline.CURVE_CUBIC, scanline.CURVE_RATIONAL


def sample():
    """Two paths: a triangle (nonzero, opaque red) and a rational arc closed by a line
    (even-odd, half-transparent blue), the second path in two contours."""
    vg = VGContainer(32, 16, 3, -2)
    vg.addPath(0, rgba(255, 0, 0, 255))
    vg.addContour()
    vg.addLinear((1, 1), (9, 1))
    vg.addLinear((9, 1), (5, 7))
    vg.addLinear((5, 7), (1, 1))
    vg.addPath(1, rgba(0, 0, 255, 128))
    vg.addContour()
    vg.addRational((10, 2), (20, 2), (20, 12), 0.5)
    vg.addContour()
    vg.addLinear((20, 12), (10, 2))
    return vg


# the sample as saveTxt prints it: "%d,%d,%d,%d\n", "%u,...\n", paths "%u,%u,%c,%c,%u,%u\n",
# contours "%u,%u\n", curves "%u,%c,%d\n", vertices "%e,%e\n"
SAMPLE_VGT = (
    b"3,-2,32,16\n"
    b"2,3,5,11,0,0\n"
    b"0,1,\x01,\x00,4278190335,4294967295\n"
    b"1,2,\x01,\x01,2164195328,4294967295\n"
    b"0,3\n"
    b"3,1\n"
    b"4,1\n"
    b"0,\x02,2139095039\n"
    b"2,\x02,2139095039\n"
    b"4,\x02,2139095039\n"
    b"6,\x13,1056964608\n"
    b"9,\x02,2139095039\n"
    b"1.000000e+00,1.000000e+00\n9.000000e+00,1.000000e+00\n"
    b"9.000000e+00,1.000000e+00\n5.000000e+00,7.000000e+00\n"
    b"5.000000e+00,7.000000e+00\n1.000000e+00,1.000000e+00\n"
    b"1.000000e+01,2.000000e+00\n2.000000e+01,2.000000e+00\n2.000000e+01,1.200000e+01\n"
    b"2.000000e+01,1.200000e+01\n1.000000e+01,2.000000e+00\n"
)


def sample_vgb():
    """The sample as saveBinary writes it: header, then each field of all items in turn."""
    p = struct.pack
    verts = [(1, 1), (9, 1), (9, 1), (5, 7), (5, 7), (1, 1), (10, 2), (20, 2), (20, 12), (20, 12), (10, 2)]
    return b"".join([
        p("<4I", 3, 0xFFFFFFFE, 32, 16),
        p("<6I", 2, 3, 5, 11, 0, 0),
        p("<2I", 0, 1), p("<2I", 1, 2),           # contourIndex, contourNumber
        bytes([1, 1]), bytes([0, 1]),             # fillType, fillRule
        bytes([255, 0, 0, 255, 0, 0, 255, 128]),  # fillColor as r, g, b, a bytes
        p("<2I", 0xFFFFFFFF, 0xFFFFFFFF),         # fillIndex
        p("<3I", 0, 3, 4), p("<3I", 3, 1, 1),     # curveIndex, curveNumber
        p("<5I", 0, 2, 4, 6, 9),                  # vertexIndex
        bytes([L, L, L, R, L]),                   # type
        p("<5i", 0x7F7FFFFF, 0x7F7FFFFF, 0x7F7FFFFF, 0x3F000000, 0x7F7FFFFF),
        b"".join(p("<2f", x, y) for x, y in verts),
    ])


def check_sample(g):
    assert (g.x, g.y, g.width, g.height) == (3, -2, 32, 16)
    assert g.pathNumber() == 2 and g.contourNumber() == 3 and g.curveNumber() == 5 and g.vertexNumber() == 11
    a = g.arrays()
    assert list(a.curve_type) == [L, L, L, R, L]
    assert list(a.curve_path_id) == [0, 0, 0, 1, 1]
    assert list(a.curve_vertex_pos) == [0, 2, 4, 6, 9]
    np.testing.assert_array_equal(a.curve_arc_w, np.float32([-1, -1, -1, 0.5, -1]))
    assert list(a.fill_rule) == [0, 1]
    assert [int(c) for c in a.fill_color] == [rgba(255, 0, 0, 255), rgba(0, 0, 255, 128)]
    assert g._ftype == [FT_COLOR, FT_COLOR]
    assert g._cfirst == [0, 3, 4] and g._cowner == [0, 1, 1]
    assert a.vertex[7].tolist() == [20.0, 2.0] and a.vertex[3].tolist() == [5.0, 7.0]


def test_text_known_answer(tmp_path):
    fn = tmp_path / "s.vgt"
    fn.write_bytes(SAMPLE_VGT)
    check_sample(VGContainer.load(str(fn)))
    out = tmp_path / "o.vgt"
    sample().save(str(out))
    assert out.read_bytes() == SAMPLE_VGT


def test_binary_known_answer(tmp_path):
    fn = tmp_path / "s.vgb"
    fn.write_bytes(sample_vgb())
    check_sample(VGContainer.load(str(fn)))
    out = tmp_path / "o.vgb"
    sample().save(str(out))
    assert out.read_bytes() == sample_vgb()


def test_text_reader_follows_fscanf(tmp_path):
    # fscanf skips whitespace before numbers (newlines between records are optional and
    # may be doubled); %c takes the next byte whatever it is; %u wraps a minus sign
    text = SAMPLE_VGT.replace(b"\n0,3\n", b"\n\n  0,3\n").replace(b"4294967295\n1,2", b"-1 1,2")
    fn = tmp_path / "w.vgt"
    fn.write_bytes(text)
    check_sample(VGContainer.load(str(fn)))
    # a space before a comma is not skipped
    fn.write_bytes(SAMPLE_VGT.replace(b"0,3\n", b"0 ,3\n"))
    with pytest.raises(vgfile.VGFileError):
        VGContainer.load(str(fn))


def test_generated_scene_round_trips(tmp_path):
    a = scanline.generate(scanline.scene.KIND_MIXED, 256, 256, 60, 77)
    vg = VGContainer.from_arrays(a, 256, 256)
    vg.save(str(tmp_path / "g.vgb"))
    b = VGContainer.load(str(tmp_path / "g.vgb")).arrays()
    for k in ("vertex", "curve_vertex_pos", "curve_type", "curve_path_id", "curve_arc_w", "fill_rule", "fill_color"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)
    vg.save(str(tmp_path / "g.vgt"))
    t = VGContainer.load(str(tmp_path / "g.vgt")).arrays()
    for k in ("curve_vertex_pos", "curve_type", "curve_path_id", "curve_arc_w", "fill_rule", "fill_color"):
        np.testing.assert_array_equal(getattr(a, k), getattr(t, k), err_msg=k)
    # "%e" keeps seven significant digits: relative error at most 5e-7 plus the float's own
    v, u = a.vertex.astype(np.float64), t.vertex.astype(np.float64)
    assert (np.abs(u - v) <= 5e-7 * np.abs(v) + np.spacing(np.abs(v).astype(np.float32)).astype(np.float64)).all()
    # and each read value is the float nearest the printed decimal
    text = (tmp_path / "g.vgt").read_bytes().split(b"\n")
    first_vertex_line = 2 + a.n_paths + vg.contourNumber() + a.n_curves
    for i in range(0, a.n_vertices, 97):
        xs = text[first_vertex_line + i].split(b",")[0].decode()
        assert float(t.vertex[i, 0]) == nearest_f32(xs)


def nearest_f32(s):
    """The float32 nearest the decimal s, ties to even, by exact rational comparison."""
    x = Fraction(s)
    f = np.float32(float(x))
    cands = [f, np.nextafter(f, np.float32(np.inf)), np.nextafter(f, np.float32(-np.inf))]
    best = min(cands, key=lambda c: (abs(Fraction(float(c)) - x), int(np.float32(c).view(np.uint32)) & 1))
    return float(best)


def test_decimal_rounding_is_single():
    rng = np.random.default_rng(3)
    toks = ["%.*e" % (int(rng.integers(3, 12)), float(v)) for v in rng.normal(0, 1e3, 3000)]
    got = vgfile.decimal_to_f32(toks)
    assert [float(g) for g in got] == [nearest_f32(t) for t in toks]
    # 1 + 2^-24 is halfway between 1 and 1 + 2^-23; a decimal a hair above it must round
    # up, although its nearest double is the halfway point itself (which rounds to even)
    half = Fraction(1) + Fraction(1, 1 << 24)
    above = str(half.numerator * 10**40 // half.denominator + 1)
    above = above[0] + "." + above[1:]
    assert float(above) == float(half)  # the double lands on the halfway point
    below = str(half.numerator * 10**40 // half.denominator - 1)
    below = below[0] + "." + below[1:]
    exact = "1.000000059604644775390625"  # exactly halfway: ties to even gives 1
    got = vgfile.decimal_to_f32([above, below, exact])
    assert got[0] == np.nextafter(np.float32(1), np.float32(2))
    assert got[1] == np.float32(1) and got[2] == np.float32(1)


def test_errors(tmp_path):
    with pytest.raises(vgfile.VGFileError):
        sample().save(str(tmp_path / "x.svg"))
    fn = tmp_path / "t.vgb"
    fn.write_bytes(sample_vgb()[:-5])
    with pytest.raises(vgfile.VGFileError):
        VGContainer.load(str(fn))
    bad = bytearray(sample_vgb())
    # vertexIndex of the last curve (offset: header 40, path fields 2*(4+4+1+1+4+4),
    # contour fields 3*(4+4), then 5 u32 vertex indices)
    off = 40 + 2 * 18 + 24 + 16
    bad[off:off + 4] = struct.pack("<I", 10)
    fn.write_bytes(bytes(bad))
    with pytest.raises(vgfile.VGFileError):
        VGContainer.load(str(fn))


def test_gradient_paths(tmp_path):
    vg = sample()
    vg._ftype[1] = FT_GRADIENT
    vg._findex[1] = 0
    vg.gradient_ramp = [1, 2, 3]
    vg.gradient_table = [-4, 5]
    fn = str(tmp_path / "g.vgb")
    vg.save(fn)
    with pytest.raises(vgfile.VGFileError):
        VGContainer.load(fn, gradients="raise")
    g = VGContainer.load(fn)
    assert g.gradient_ramp == [1, 2, 3] and g.gradient_table == [-4, 5]
    # the gradient path's fill word names its table slot (path index + 1)
    assert [int(c) for c in g.arrays().fill_color] == [rgba(255, 0, 0, 255), 2]
    fn2 = str(tmp_path / "g2.vgt")
    g.save(fn2)
    assert VGContainer.load(fn2, gradients="skip")._ftype == [FT_COLOR, FT_GRADIENT]
