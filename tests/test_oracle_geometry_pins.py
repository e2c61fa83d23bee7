"""Pinning the oracle's geometry stages (runs without a GPU).

Covers the curve orientation (ras_base.cpp:523-569), the vertex transform with its
integral-x nudge and culling flags (ras_cut_mask_comb_scanline.cu:36-105,
ras_define.h:119-126), the monotone split (ras_cut.h:1314-1409) and the curved-segment
intersections (ras_cut.h:1033-1106).  Every expectation is stated independently of the
oracle's code:
- hand-worked fixtures (unsigned bit-pattern order, the -1/256 nudge);
- float64 closed forms: Bernstein and rational evaluation, derivative roots by numpy
  polynomial arithmetic, square / cube roots and circle crossings;
- brute force: the culling rule read as "every vertex strictly beyond one canvas edge".
Each is chosen so that a plausible slip fails at least one check: a float compare
instead of the unsigned one, a dropped or misplaced nudge, a transposed matrix, a sign
error in a derivative coefficient, a wrong root branch or a wrong bisection bracket.
"""
import math

import numpy as np
import pytest
from numpy.polynomial import polynomial as Pn

import pyoracle
from scanline import VGContainer, rgba

LINE, QUAD, CUBIC, RATIONAL = 2, 3, 4, 0x13


def run(vg, W, H, spp=8, transform=None):
    return pyoracle.run(vg.arrays(), W, H, spp, transform=transform)


def f32(x):
    return float(np.float32(x))


def ulp32(x):
    return float(np.spacing(np.float32(abs(x))))


# ---------------------------------------------------------------------------------
# float64 curve algebra (power-basis coefficients, lowest degree first)
# ---------------------------------------------------------------------------------
def bernstein_poly(ctrl):
    """Power-basis coefficients of a Bezier component with control values ctrl."""
    n = len(ctrl) - 1
    out = np.zeros(n + 1)
    for k, c in enumerate(ctrl):
        # C(n,k) t^k (1-t)^(n-k)
        b = Pn.polymul(Pn.polypow([0.0, 1.0], k), Pn.polypow([1.0, -1.0], n - k)) * math.comb(n, k)
        out[: b.size] += c * b
    return out


def rational_polys(ctrl, w):
    """Numerator and denominator of a rational quadratic with middle weight w."""
    num = bernstein_poly([ctrl[0], w * ctrl[1], ctrl[2]])
    den = bernstein_poly([1.0, w, 1.0])
    return num, den


def component(ctype, v, w):
    """(value(t), derivative-numerator polynomial) of one coordinate in float64."""
    v = [float(x) for x in v]
    if ctype == RATIONAL:
        num, den = rational_polys(v, float(w))
        dnum = Pn.polysub(Pn.polymul(Pn.polyder(num), den), Pn.polymul(num, Pn.polyder(den)))
        return (lambda t: Pn.polyval(t, num) / Pn.polyval(t, den)), dnum
    p = bernstein_poly(v)
    return (lambda t: Pn.polyval(t, p)), Pn.polyder(p)


def unit_roots(poly):
    """Real roots of a power-basis polynomial strictly inside (0, 1)."""
    poly = np.trim_zeros(np.asarray(poly, dtype=np.float64), "b")
    if poly.size <= 1:
        return np.zeros(0)
    r = Pn.polyroots(poly)
    r = r[np.abs(r.imag) <= 1e-9 * max(1.0, np.abs(r).max())].real
    return np.sort(r[(r > 0) & (r < 1)])


def curve_vertices(scene_arrays, res, c):
    """Transformed (unified) vertices of curve c as float64 (n, 2)."""
    tv = res.get("tvertex").astype(np.float64)
    p = int(scene_arrays.curve_vertex_pos[c])
    n = int(scene_arrays.curve_type[c]) & 7
    return tv[p:p + n]


# ---------------------------------------------------------------------------------
# a2: curve orientation by unsigned bit pattern (ras_base.cpp:523-569)
# ---------------------------------------------------------------------------------
# (first vertex, last vertex, reversed?) worked by hand: the curve keeps its order only if
# (y_first, x_first) < (y_last, x_last) comparing the IEEE bit patterns as uint32.
ORIENTATION_FIXTURES = [
    ((-1.5, 5.25), (3.5, 5.25), True),     # equal y; x bits 0xBFC00000 > 0x40600000
    ((1.5, -2.25), (1.5, 3.25), True),     # y bits 0xC0100000 > 0x40500000
    ((1.5, -2.25), (1.5, -3.25), False),   # 0xC0100000 < 0xC0500000 (float order would reverse)
    ((-0.5, -1.5), (-2.5, -1.5), False),   # equal y; x 0xBF000000 < 0xC0200000
    ((2.5, 3.5), (1.5, 3.5), True),        # positive values: the ordinary order
    ((0.5, 1.5), (0.25, 7.5), False),      # y_first < y_last
    ((0.5, -0.0), (0.75, 0.0), True),      # -0.0 is 0x80000000, above +0.0
    ((0.75, 0.0), (0.5, -0.0), False),
    ((1.25, 1.25), (1.25, 1.25), True),    # equal end points take the reversing branch
]


def test_orientation_uses_unsigned_bit_order():
    vg = VGContainer()
    vg.addPath(0, rgba(1, 2, 3))
    for a, b, _ in ORIENTATION_FIXTURES:
        vg.addLinear(a, b)
    # a cubic whose first vertex has a negative y: all four vertices reverse
    cub = [(3.25, -4.5), (9.75, 20.5), (1.25, 7.5), (5.5, 2.75)]
    vg.addCubic(*cub)
    arr = vg.arrays()
    r = pyoracle.run(arr, 16, 16, 8)
    tv = r.get("tvertex")
    for i, (a, b, rev) in enumerate(ORIENTATION_FIXTURES):
        p = int(arr.curve_vertex_pos[i])
        got = [tuple(tv[p]), tuple(tv[p + 1])]
        # identity transform on a 16x16 canvas is exact for these dyadic values (and maps
        # -0.0 to +0.0); no x here is integral, so no nudge applies
        want = [b, a] if rev else [a, b]
        want = [(f32(x), f32(y) + 0.0) for x, y in want]
        assert got == want, (i, a, b, rev, got)
    p = int(arr.curve_vertex_pos[len(ORIENTATION_FIXTURES)])
    np.testing.assert_array_equal(tv[p:p + 4], np.array(cub[::-1], dtype=np.float32))


# ---------------------------------------------------------------------------------
# a3: transform, nudge, culling (ras_cut_mask_comb_scanline.cu:36-105, ras_define.h:119-126)
# ---------------------------------------------------------------------------------
def test_identity_transform_keeps_vertices_and_nudges_integral_x():
    W, H = 64, 48
    starts = [(4.0, 4.0), (0.0, 7.0), (64.0, 0.0), (-2.0, 12.0), (3.99609375, 9.0),
              (10.5, 10.0), (17.25, 48.0), (31.0, 5.5), (100.0, 20.0), (-0.75, 30.0)]
    vg = VGContainer()
    vg.addPath(0, rgba(1, 2, 3))
    for x, y in starts:
        # rising y with positive coordinates: the curve keeps its vertex order
        vg.addLinear((x, y), (x + 0.25, y + 1.0))
    arr = vg.arrays()
    tv = pyoracle.run(arr, W, H, 8).get("tvertex")
    for i, (x, y) in enumerate(starts):
        p = int(arr.curve_vertex_pos[i])
        for (vx, vy), (gx, gy) in zip([(x, y), (x + 0.25, y + 1.0)], tv[p:p + 2]):
            want_x = vx - 1.0 / 256 if float(vx).is_integer() else vx
            assert gx == np.float32(want_x), (vx, gx)
            assert gy == np.float32(vy), (vy, gy)     # y is never nudged
    # the worked example: 4.0 -> 3.99609375; a value already 1/256 below stays put
    assert tv[int(arr.curve_vertex_pos[0])][0] == np.float32(3.99609375)
    assert tv[int(arr.curve_vertex_pos[4])][0] == np.float32(3.99609375)


def _affine_points(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.uniform(1, 200, (n, 2)) + 0.123).astype(np.float32)


def test_affine_and_projective_transform():
    """glm order m[col][row]: x' = m00 x + m10 y + m30, y' = m01 x + m11 y + m31, each
    divided by w = m03 x + m13 y + m33 (ras_device_func.h:39-41 row dots)."""
    W, H = 300, 260
    m = np.eye(4, dtype=np.float32)
    m[0, 0], m[1, 1] = 0.9, 1.1
    m[1, 0], m[0, 1] = 0.05, -0.07
    m[3, 0], m[3, 1] = 7.25, -3.5
    m[0, 3], m[1, 3] = 2.0e-4, -1.0e-4
    pts = _affine_points(200, 5)
    vg = VGContainer()
    vg.addPath(0, rgba(1, 2, 3))
    for k in range(0, len(pts), 2):
        a, b = pts[k], pts[k + 1]
        if (a[1], a[0]) > (b[1], b[0]):
            a, b = b, a
        vg.addLinear(tuple(a), tuple(b))
    arr = vg.arrays()
    tv = pyoracle.run(arr, W, H, 8, transform=m).get("tvertex").astype(np.float64)
    M = m.astype(np.float64)
    v = arr.vertex.astype(np.float64)
    X = M[0, 0] * v[:, 0] + M[1, 0] * v[:, 1] + M[3, 0]
    Y = M[0, 1] * v[:, 0] + M[1, 1] * v[:, 1] + M[3, 1]
    Wt = M[0, 3] * v[:, 0] + M[1, 3] * v[:, 1] + M[3, 3]
    X, Y = X / Wt, Y / Wt
    tol = 1e-5 * np.maximum(1.0, np.abs(X))
    dx = np.abs(tv[:, 0] - X)
    nudged = np.abs(tv[:, 0] + 1.0 / 256 - X) <= tol
    assert np.all((dx <= tol) | nudged), dx.max()
    np.testing.assert_allclose(tv[:, 1], Y, rtol=1e-5, atol=1e-5)
    # the transposed matrix gives a visibly different vertex set
    Xt = (M[0, 0] * v[:, 0] + M[0, 1] * v[:, 1] + M[0, 3]) / (M[3, 0] * v[:, 0] + M[3, 1] * v[:, 1] + M[3, 3])
    assert np.abs(tv[:, 0] - Xt).max() > 1.0


@pytest.mark.parametrize("depth,factor", [(47.0, 0.5), (0.0, 1.0), (-100.0, 2.0)])
def test_perspective_depth_scaling(depth, factor):
    """With a depth d = z-row . v, points move toward the canvas centre by z / (d + z),
    z = min(W, H), and depths below -z/2 are clamped there (a factor of 2).  The centre
    itself stays fixed."""
    W, H = 63, 47
    z = float(min(W, H))
    m = np.eye(4, dtype=np.float32)
    m[3, 2] = depth
    # offsets chosen so that no scaled x is integral (that x would take the -1/256 nudge)
    pts = [(31.5, 23.5), (31.5 + 10.125, 23.5 + 6.125), (31.5 - 12.625, 23.5 + 3.375)]
    vg = VGContainer()
    vg.addPath(0, rgba(1, 2, 3))
    vg.addLinear(pts[0], pts[1])
    vg.addLinear(pts[2], (pts[2][0] + 0.5, pts[2][1] + 0.5))
    arr = vg.arrays()
    tv = pyoracle.run(arr, W, H, 8, transform=m).get("tvertex").astype(np.float64)
    assert tuple(tv[0]) == (31.5, 23.5)
    for idx, (x, y) in ((1, pts[1]), (2, pts[2])):
        np.testing.assert_allclose(tv[idx], [31.5 + (x - 31.5) * factor, 23.5 + (y - 23.5) * factor],
                                   rtol=0, atol=2e-5)
    assert abs(z / (max(depth, -z / 2) + z) - factor) < 1e-12


def test_path_culling_by_brute_force():
    """A path produces no fragment exactly when all its (transformed) vertices lie
    strictly beyond one edge of the canvas: x < 0, x > W, y < 0 or y > H.  Vertices on an
    edge (x == 0, x == W, ...) count as inside."""
    W, H = 40, 30
    xs = [-30.5, -0.001, 0.0, 0.3, 20.5, 39.8, 40.0, 40.001, 70.5]
    ys = [-25.5, -0.001, 0.0, 0.7, 15.5, 29.6, 30.0, 30.002, 55.5]
    rng = np.random.default_rng(17)
    vg = VGContainer()
    expect = []
    for _ in range(600):
        tri = [(f32(rng.choice(xs)), f32(rng.choice(ys))) for _ in range(3)]
        if len({tri[0], tri[1], tri[2]}) < 3:
            continue
        vg.addPath(0, rgba(9, 9, 9))
        vg.addPolygon(tri)
        px = [p[0] for p in tri]
        py = [p[1] for p in tri]
        culled = all(x < 0 for x in px) or all(x > W for x in px) or \
            all(y < 0 for y in py) or all(y > H for y in py)
        expect.append(culled)
    arr = vg.arrays()
    r = pyoracle.run(arr, W, H, 8)
    fc = r.get("frag_count").astype(np.int64)
    per_path = np.bincount(arr.curve_path_id, weights=fc, minlength=arr.n_paths)
    got = per_path == 0
    assert np.array_equal(got, np.array(expect)), np.flatnonzero(got != np.array(expect))[:10]
    assert 20 < got.sum() < len(expect) - 20
    # the region flags: the 3x3 diagram of the transform kernel, 5|6|7 above the canvas
    # (y > H), 3|in|4 beside it, 0|1|2 below (y < 0); an inside vertex marks 0 and 7
    region = {(0, 0): [0], (0, 1): [1], (0, 2): [2], (1, 0): [3], (1, 1): [0, 7], (1, 2): [4],
              (2, 0): [5], (2, 1): [6], (2, 2): [7]}
    flags = r.get("path_flags")
    v = arr.vertex
    cls = lambda a, hi: 0 if a < 0 else (1 if a <= hi else 2)  # noqa: E731
    for p in range(arr.n_paths):
        want = np.zeros(8, np.uint8)
        for c in np.flatnonzero(arr.curve_path_id == p):
            q = int(arr.curve_vertex_pos[c])
            for x, y in v[q:q + 2]:
                for s in region[(cls(y, H), cls(x, W))]:
                    want[s] = 1
        np.testing.assert_array_equal(flags[p], want, err_msg="path %d" % p)


# ---------------------------------------------------------------------------------
# a4: monotone split (ras_cut.h:1314-1409)
# ---------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def curved_scene():
    """Random quadratics, cubics and rational quadratics, all inside a 256x256 canvas."""
    rng = np.random.default_rng(23)
    vg = VGContainer()
    vg.addPath(0, rgba(9, 9, 9))
    kinds = []
    for _ in range(900):
        k = rng.choice([QUAD, CUBIC, RATIONAL])
        n = {QUAD: 3, CUBIC: 4, RATIONAL: 3}[k]
        pts = [tuple(p) for p in (rng.uniform(3, 253, (n, 2)) + 0.0371).astype(np.float32)]
        if k == QUAD:
            vg.addQuadratic(*pts)
        elif k == CUBIC:
            vg.addCubic(*pts)
        else:
            vg.addRational(pts[0], pts[1], pts[2], f32(rng.uniform(0.2, 3.0)))
        kinds.append(k)
    arr = vg.arrays()
    return arr, pyoracle.run(arr, 256, 256, 8)


def test_monotone_cuts_are_derivative_roots(curved_scene):
    """Each cut is a root of x'(t) or y'(t) in (0, 1) (float64 polynomial algebra on the
    unified vertices, for rationals of N'D - ND'), and every well-separated root is a cut."""
    arr, r = curved_scene
    cuts, ncuts = r.get("cuts"), r.get("ncuts")
    worst, missed, checked = 0.0, 0, 0
    for c in range(arr.n_curves):
        ct = int(arr.curve_type[c])
        v = curve_vertices(arr, r, c)
        roots = []
        for axis in (0, 1):
            _, dpoly = component(ct, v[:, axis], arr.curve_arc_w[c])
            roots.extend(unit_roots(dpoly).tolist())
        roots = np.sort(np.array(roots))
        got = np.sort(cuts[c, :ncuts[c]].astype(np.float64))
        for q in got:
            assert roots.size, (c, got)
            d = np.abs(roots - q).min()
            worst = max(worst, d)
            assert d < 2e-3, (c, q, roots)
            checked += 1
        for x in roots:
            # roots of ill-conditioned (near-double) or edge cases may be lost in float32
            if got.size and np.abs(got - x).min() < 2e-3:
                continue
            if min(x, 1 - x) < 1e-4:
                continue
            missed += 1
    assert checked > 800
    assert missed <= 3, missed
    assert worst < 2e-3


def test_monotone_pieces_are_monotone(curved_scene):
    """Between consecutive cuts (with 0 and 1) both coordinates are monotone: dense
    float64 sampling finds no direction change beyond rounding."""
    arr, r = curved_scene
    cuts, ncuts = r.get("cuts"), r.get("ncuts")
    for c in range(arr.n_curves):
        ct = int(arr.curve_type[c])
        v = curve_vertices(arr, r, c)
        knots = np.concatenate([[0.0], np.sort(cuts[c, :ncuts[c]].astype(np.float64)), [1.0]])
        for axis in (0, 1):
            f, _ = component(ct, v[:, axis], arr.curve_arc_w[c])
            for a, b in zip(knots[:-1], knots[1:]):
                if b - a < 1e-3:
                    continue
                t = np.linspace(a + 5e-4 * (b - a), b - 5e-4 * (b - a), 257)
                d = np.diff(f(t))
                eps = 1e-9 * max(1.0, np.abs(f(t)).max())
                assert np.all(d >= -eps) or np.all(d <= eps), (c, axis, a, b)


def test_a_sign_slip_in_the_cubic_derivative_is_caught():
    """The check above fails for the plausible slip named in its docstring: the cubic's
    linear coefficient with its sign flipped gives roots that are not derivative roots."""
    rng = np.random.default_rng(29)
    bad = 0
    for _ in range(200):
        y = rng.uniform(0, 250, 4)
        a = (y[3] - y[0]) + 3 * (y[1] - y[2])
        b = 2 * ((y[0] - y[1]) + (y[2] - y[1]))
        good = unit_roots([y[1] - y[0], b, a])
        slip = unit_roots([y[1] - y[0], -b, a])
        _, dpoly = component(CUBIC, y, 0)
        true = unit_roots(dpoly)
        np.testing.assert_allclose(good, true, atol=1e-9)
        if slip.size != true.size or (slip.size and np.abs(slip - true).max() > 2e-3):
            bad += 1
    assert bad > 100


# ---------------------------------------------------------------------------------
# a6: curved-segment intersections (ras_cut.h:1033-1106)
# ---------------------------------------------------------------------------------
def _entries(r):
    ts = r.get("ts")
    flags = ts[:, 1] & 3
    t = (ts[:, 1].astype(np.uint32) & 0xFFFFFFFC).view(np.float32).astype(np.float64)
    return ts[:, 0], flags, t


def _y_cut_ts(r, cid):
    cc, flags, t = _entries(r)
    sel = (cc == cid) & (flags == 1) & (t > 0) & (t < 1)
    return t[sel]


def test_quadratic_crossings_closed_form():
    """y(t) = 0.5 + 60 t^2 (x constant at 5.3, so no x cuts): the crossing of y = c is
    t = sqrt((c - 0.5) / 60) for c = 2, 4, ..., 60."""
    vg = VGContainer()
    vg.addPath(0, rgba(9, 9, 9))
    vg.addQuadratic((5.3, 0.5), (5.3, 0.5), (5.3, 60.5))
    r = run(vg, 16, 64)
    got = _y_cut_ts(r, 0)
    want = np.sqrt((np.arange(2, 61, 2) - 0.5) / 60.0)
    assert got.size == want.size
    tol = np.array([4 * ulp32(x) for x in want]) + 1e-6
    assert np.all(np.abs(got - want) <= tol), np.abs(got - want).max()


def test_cubic_crossings_closed_form():
    """y(t) = 0.5 + 60 t^3: the 24-step bisection lands within its bracket width
    (2^-24 of the piece) plus the flag bits of cbrt((c - 0.5) / 60)."""
    vg = VGContainer()
    vg.addPath(0, rgba(9, 9, 9))
    vg.addCubic((5.3, 0.5), (5.3, 0.5), (5.3, 0.5), (5.3, 60.5))
    r = run(vg, 16, 64)
    got = _y_cut_ts(r, 0)
    want = np.cbrt((np.arange(2, 61, 2) - 0.5) / 60.0)
    assert got.size == want.size
    tol = 2.0 ** -24 + np.array([4 * ulp32(x) for x in want])
    assert np.all(np.abs(got - want) <= tol), np.abs(got - want).max()


def test_rational_quarter_circle_crossings():
    """A quarter circle as a rational quadratic (w = sqrt(2)/2): every y crossing lies on
    the circle at height c, and every x crossing at abscissa c (float64 circle algebra)."""
    cx, cy, R = 70.3, 3.7, 50.0
    w = f32(np.sqrt(0.5))
    p0, p1, p2 = (cx + R, cy), (cx + R, cy + R), (cx, cy + R)
    vg = VGContainer()
    vg.addPath(0, rgba(9, 9, 9))
    vg.addRational(p0, p1, p2, w)
    r = run(vg, 128, 64)
    cc, flags, t = _entries(r)
    ctrl = np.array([p0, p1, p2], dtype=np.float32).astype(np.float64)
    n_checked = 0
    for side in (0, 1):
        num, den = rational_polys(ctrl[:, side], w)
        for tv in t[(flags == side) & (t > 0) & (t < 1)]:
            val = Pn.polyval(tv, num) / Pn.polyval(tv, den)
            c = 2 * np.round(val / 2)
            # the exact crossing of the even line c, from c * D(t) - N(t) = 0
            roots = unit_roots(Pn.polysub(c * den, num))
            assert roots.size == 1, (side, tv, roots)
            assert abs(tv - roots[0]) <= 8 * ulp32(tv) + 2e-6, (side, tv, roots[0])
            n_checked += 1
        # the point is on the circle (float32 weight: radius within 1e-3 px)
    assert n_checked == 25 + 25


@pytest.mark.parametrize("ctype", [QUAD, CUBIC, RATIONAL])
def test_curved_intersections_lie_on_cut_lines(curved_scene, ctype):
    """Every solved entry (flags 0 = x cut, 1 = y cut) of a random curve evaluates, in
    float64 at its emitted t, onto an even line: within |B'(t)| times the t error bound
    (closed-form solves: the flag bits; the cubic: its last bisection bracket) plus the
    float32 rounding of the curve value the solver tests.  The cubic's fallback (no sign
    change in the bracket: t stays at the bracket start) must stay rare."""
    arr, r = curved_scene
    cc, flags, t = _entries(r)
    sel = np.flatnonzero(np.isin(flags, (0, 1)) & (t > 0) & (t < 1) & (arr.curve_type[cc] == ctype))
    assert sel.size > 2000
    sel = np.random.default_rng(ctype).choice(sel, 2500, replace=False)
    worst, fallback = 0.0, 0
    prev_t = np.concatenate([[0.0], t[:-1]])
    for i in sel:
        c = int(cc[i])
        v = curve_vertices(arr, r, c)
        axis = int(flags[i])
        f, dpoly = component(ctype, v[:, axis], arr.curve_arc_w[c])
        val = f(t[i])
        line = 2 * np.round(val / 2)
        if ctype == RATIONAL:
            _, den = rational_polys(v[:, axis], arr.curve_arc_w[c])
            deriv = abs(Pn.polyval(t[i], dpoly) / Pn.polyval(t[i], den) ** 2)
        else:
            deriv = abs(Pn.polyval(t[i], dpoly))
        dt = 4 * ulp32(t[i]) + (2.0 ** -24 if ctype == CUBIC else 8 * ulp32(t[i]))
        tol = 2 * dt * deriv + 16 * ulp32(np.abs(v).max())
        err = abs(val - line)
        if err > tol and ctype == CUBIC and cc[i - 1] == c and t[i] == prev_t[i]:
            fallback += 1
            continue
        assert err <= tol, (i, c, axis, t[i], val, line, tol)
        worst = max(worst, err / tol)
    assert fallback <= sel.size // 1000, fallback
