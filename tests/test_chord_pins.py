"""Pins of the chord-animation oracle and of the keyframe builder (no GPU).

The oracle (oracle/chord_oracle.cpp) restates the reference's device generator
(animation.cu:1890-2432); the builder (scanline/animation.py) its host keyframes
(animation.cu:1086-1330).  Neither is checked here by retyping its formulas: the checks
are what the geometry fixes -- points on the circles at the keyframe angles, the closed
form of a cubic arc's midpoint, closed outlines, the chord radius 320 / (2 sin(theta/2) + 1)
-- plus hand-worked keyframes, the ease curve's known values, and sine and cosine against
the C library.  test_chord_oracle_mutations.py plants mistakes in the oracle and shows
these checks fail.
"""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle  # noqa: E402

from scanline import animation as an  # noqa: E402

TWO_PI = 2 * math.pi


def _small_diagrams(seed=3, n_nodes=12, years=3, links=20):
    recs = an.synthetic_records(n_nodes, years, links, seed=seed)
    _, src, dst = an.diagrams_from_records(recs)
    colors = an.palette(n_nodes, seed)
    return [an.chord_keyframes(src, n_nodes, colors, an.CENTERS[0]),
            an.chord_keyframes(dst, n_nodes, colors, an.CENTERS[1])]


def _interp(d, time_ms):
    """The frame's node angles and sizes and link records, from the keyframes (float64)."""
    k = min(time_ms // an.KEY_MS + 1, d.n_keys - 1)
    t = float(pyoracle.chord_ease(time_ms)) if time_ms // an.KEY_MS + 1 <= d.n_keys - 1 else 1.0
    lerp = lambda a, b: a.astype(np.float64) * (1 - t) + b.astype(np.float64) * t
    n0, n1, ns = lerp(d.node_t0[k - 1], d.node_t0[k]), lerp(d.node_t1[k - 1], d.node_t1[k]), \
        lerp(d.node_size[k - 1], d.node_size[k])
    lo, hi = int(d.link_offset[k - 1]), int(d.link_offset[k])
    lr = lerp(d.link_from[lo:hi], d.link_to[lo:hi])
    return n0, n1, ns, d.link_node[lo:hi], lr


def _angle_of(p, c):
    """Angle t of point p about c with p - c along (sin t, cos t), in [0, 2 pi)."""
    return math.atan2(p[0] - c[0], p[1] - c[1]) % TWO_PI


def _same_angle(a, b, tol):
    d = (a - b) % TWO_PI
    return min(d, TWO_PI - d) < tol


def _unorient(frame, c):
    """Vertices of curve c in generation order (undo the upload's orientation)."""
    vp, _, info, _ = frame["curve"][c]
    nv = info & 7
    v = frame["vertex"][vp:vp + nv].astype(np.float64)
    return v[::-1] if (info >> 8) & 1 else v


def _midpoint_radius(R, span):
    """|B(1/2) - centre| of one of the two cubics of an arc of angle `span` on radius R with
    the handle length sin(span / 2) * 0.353 * R: each cubic spans 2a = span / 2, and a
    symmetric cubic from angle -a to a has its midpoint at R cos a + 3/4 h sin a."""
    a = span / 4
    h = math.sin(span / 2) * 0.353 * R
    return R * math.cos(a) + 0.75 * h * math.sin(a)


# ---- sine, cosine, ease ----------------------------------------------------------------

def test_sincos_matches_libm():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-20, 20, 4000), np.float32(np.pi) * np.arange(-8, 9) / 4,
                         [0.0, 1e-30, -1e-30, 1e-6, 7.0]]).astype(np.float32)
    exact = 0
    for x in xs:
        s, c = pyoracle.chord_sincos(float(x))
        ws, wc = math.sin(float(x)), math.cos(float(x))
        for got, want in ((s, ws), (c, wc)):
            ulp = max(np.spacing(np.float32(abs(want))), np.float32(1e-45))
            assert abs(got - want) <= ulp, (x, got, want)
        exact += int(np.float32(s) == np.float32(ws)) + int(np.float32(c) == np.float32(wc))
    assert exact >= 2 * len(xs) - 2  # rounded once from double: correctly rounded but for near-ties


def test_ease_known_values():
    e = pyoracle.chord_ease
    assert e(0) == 1.0 and e(5999) == 1.0     # before the first interval ends: keyframe 1
    assert e(6000) == 0.0 and e(12000) == 0.0  # each interval starts at its first keyframe
    u = lambda ms: ms / (6000 * 0.66666)       # the interval's first two thirds map to [0, 1]
    assert abs(e(7000) - 2 * u(1000) ** 2) < 1e-6  # ease in: 2 u^2
    assert abs(e(8000) - 0.5) < 1e-4
    assert abs(e(9000) - (1 - 2 * (1 - u(3000)) ** 2)) < 1e-6  # ease out
    assert e(10000) == 1.0 and e(11999) == 1.0  # held after two thirds of the interval
    prev = -1.0
    for t in range(6000, 10001, 25):
        assert e(t) >= prev
        prev = e(t)


# ---- keyframe builder ----------------------------------------------------------------

def test_positions_worked_example():
    # links (0,1): 3000 / 1000 and (1,2): 2000 / 2000.  Node sizes 3000, 3000, 2000;
    # ring order 0, 1, 2 (ties by id); total 8000; three non-empty nodes, gap 2 degrees.
    links = {(0, 1): [np.float32(3000), np.float32(1000)], (1, 2): [np.float32(2000), np.float32(2000)]}
    pos, size, t0, t1 = an._positions(links, 3)
    assert [float(s) for s in size] == [3000, 3000, 2000]
    assert [float(x) for x in pos[(0, 1)]] == [0, 0]
    assert [float(x) for x in pos[(1, 2)]] == [1000, 0]
    gap = math.radians(2)
    span = TWO_PI - 3 * gap
    want0 = [2 * gap, 3 * gap + span * 3 / 8, 4 * gap + span * 6 / 8]
    want1 = [2 * gap + span * 3 / 8, 3 * gap + span * 6 / 8, 4 * gap + span]
    np.testing.assert_allclose([float(x) for x in t0], want0, atol=2e-6)
    np.testing.assert_allclose([float(x) for x in t1], want1, atol=2e-6)


def test_positions_ring_invariants():
    recs = an.synthetic_records(30, 1, 80, seed=5)
    _, src, _ = an.diagrams_from_records(recs)
    links = an._fold_self_links(src[0])
    pos, size, t0, t1 = an._positions(links, 30)
    size = np.array(size, np.float64)
    t0, t1 = np.array(t0, np.float64), np.array(t1, np.float64)
    live = size > 0
    # each node's extent is proportional to its size; the gaps are 2 degrees; one turn
    k = (t1[live] - t0[live]) / size[live]
    np.testing.assert_allclose(k, k[0], rtol=1e-4)
    order = np.argsort(t0[live])
    a0, a1 = t0[live][order], t1[live][order]
    np.testing.assert_allclose(a0[1:] - a1[:-1], math.radians(2), atol=2e-5)
    # n extents of 2 pi - n gaps in all, n - 1 gaps between them
    assert abs((a1[-1] - a0[0]) - (TWO_PI - math.radians(2))) < 1e-4
    # links fill their nodes end to end
    for n in np.nonzero(live)[0]:
        ends = []
        for (a, b), (p0, p1) in pos.items():
            s0, s1 = links[(a, b)]
            if a == n:
                ends.append((float(p0), float(s0)))
            if b == n and a != b:
                ends.append((float(p1), float(s1)))
        ends.sort()
        at = 0.0
        for p, s in ends:
            assert abs(p - at) <= 1e-3 * max(1.0, at)
            at += s
        assert abs(at - size[n]) <= 1e-3 * size[n]


def test_records_filter_and_directions():
    R = an.Record
    recs = [R(2001, 0, 1, 500, 900),       # both below 1000: dropped
            R(2001, 2, 1, 5000, 0),        # zero becomes 1; stored under (1, 2) reversed
            R(2001, 3, 3, 1200, 800),      # self link: the sum at end 0
            R(2000, 0, 2, 1500, 2500)]
    years, src, dst = an.diagrams_from_records(recs)
    assert years == [2000, 2001]
    assert src[0] == {(0, 2): [1500, 2500]} and dst[0] == {(0, 2): [2500, 1500]}
    assert src[1] == {(1, 2): [1, 5000], (3, 3): [2000, 0]}
    assert dst[1] == {(1, 2): [5000, 1], (3, 3): [0, 2000]}
    assert an._fold_self_links(dst[1])[(3, 3)] == [2000, 0]


def test_interval_links_union():
    f = np.float32
    k0 = {(0, 1): [f(2000), f(1000)], (1, 2): [f(1500), f(1500)]}
    k1 = {(0, 1): [f(2500), f(500)], (0, 2): [f(3000), f(1000)]}
    d = an.chord_keyframes([k0, k1], 3, [0xFF0000FF, 0xFF00FF00, 0xFFFF0000])
    assert d.link_offset.tolist() == [0, 3]
    # links in the order (j, i) for i, then j <= i
    assert d.link_node.tolist() == [[0, 1], [0, 2], [1, 2]]
    assert d.link_from[1].tolist() == [0, 0, 0, 0]    # (0, 2) grows from nothing
    assert d.link_to[2].tolist() == [0, 0, 0, 0]      # (1, 2) shrinks to nothing
    assert d.link_from[0, 2:].tolist() == [2000, 1000] and d.link_to[0, 2:].tolist() == [2500, 500]
    assert d.link_from[2, :2].tolist() == [1000, 0]   # (1, 2) sat after (0, 1) in node 1


# ---- the generated scene ---------------------------------------------------------------

@pytest.mark.parametrize("time_ms", [0, 7000, 8000, 13500, 30000])
def test_scene_geometry(time_ms):
    ds = _small_diagrams()
    fr = pyoracle.chord_frame(ds, time_ms)
    P = sum(d.paths_at(time_ms) for d in ds)
    assert fr["fill"].shape[0] == P and fr["curve"].shape[0] == 6 * P
    base = 0
    for d in ds:
        c = d.center
        n0, n1, ns, lnode, lr = _interp(d, time_ms)
        for i in range(d.n_nodes):
            p = base + i
            cv = [_unorient(fr, 6 * p + j) for j in range(6)]
            assert fr["fill"][p] == d.node_color[i]
            if n0[i] == n1[i] or abs(n1[i] - n0[i]) < 1e-6:
                continue
            # radial lines: outer -> inner at t0, inner -> outer at t1
            for pt, R, t in ((cv[0][0], 350, n0[i]), (cv[0][1], 320, n0[i]), (cv[1][0], 320, n1[i]),
                             (cv[1][1], 350, n1[i])):
                assert abs(math.hypot(pt[0] - c[0], pt[1] - c[1]) - R) < 2e-3
                assert _same_angle(_angle_of(pt, c), t, 2e-5)
            # arcs: inner t0 -> t1, outer t1 -> t0, meeting at the mean angle
            span = n1[i] - n0[i]
            for j, R in ((2, 320), (3, 320), (4, 350), (5, 350)):
                b = cv[j]
                mid = (b[0] + 3 * b[1] + 3 * b[2] + b[3]) / 8
                assert abs(math.hypot(mid[0] - c[0], mid[1] - c[1]) - _midpoint_radius(R, span)) < 3e-3
                assert abs(math.hypot(b[3][0] - c[0], b[3][1] - c[1]) - R) < 2e-3
            assert _same_angle(_angle_of(cv[2][3], c), (n0[i] + n1[i]) / 2, 2e-5)
            # closed outline, end to start
            chain = [cv[0], cv[2], cv[3], cv[1], cv[4], cv[5]]
            for a, b in zip(chain, chain[1:] + chain[:1]):
                assert np.array_equal(a[-1], b[0])
        for li in range(lnode.shape[0]):
            p = base + d.n_nodes + li
            a, b = lnode[li]
            pos0, pos1, s0, s1 = lr[li]
            ang = lambda n, f: n0[n] * (1 - f) + n1[n] * f
            fa0 = pos0 / ns[a] if ns[a] else 0.0
            fa1 = (pos0 + s0) / ns[a] if ns[a] else 0.0
            fb0 = pos1 / ns[b] if ns[b] else 0.0
            fb1 = (pos1 + s1) / ns[b] if ns[b] else 0.0
            la0, la1, lb0, lb1 = ang(a, fa0), ang(a, fa1), ang(b, fb0), ang(b, fb1)
            cv = [_unorient(fr, 6 * p + j) for j in range(6)]
            col = int(d.node_color[a if s0 > s1 else b]) & 0x7FFFFFFF
            assert fr["fill"][p] == (col if col >> 24 else 0)
            # the four ends on the inner circle at their angles
            for pt, t in ((cv[2][0], la0), (cv[3][3], la1), (cv[4][0], lb0), (cv[5][3], lb1)):
                assert abs(math.hypot(pt[0] - c[0], pt[1] - c[1]) - 320) < 2e-3
                assert _same_angle(_angle_of(pt, c), t, 3e-5)
            # the chords' handles point at the centre, at 320 / (2 sin(theta / 2) + 1)
            th_s, th_l = abs(lb0 - la1), abs(lb1 - la0)
            th_s = th_s - math.pi if th_s > math.pi else th_s
            th_l = th_l - math.pi if th_l > math.pi else th_l
            rs, rl = 320 / (2 * math.sin(th_s / 2) + 1), 320 / (2 * math.sin(th_l / 2) + 1)
            for pt, r, t in ((cv[0][1], rs, la1), (cv[0][2], rs, lb0), (cv[1][1], rl, lb1), (cv[1][2], rl, la0)):
                assert abs(math.hypot(pt[0] - c[0], pt[1] - c[1]) - r) < 3e-3
                if r > 1:
                    assert _same_angle(_angle_of(pt, c), t, 3e-5)
            chain = [cv[2], cv[3], cv[0], cv[4], cv[5], cv[1]]
            for x, y in zip(chain, chain[1:] + chain[:1]):
                assert np.array_equal(x[-1], y[0])
        base += d.paths_at(time_ms)


def test_scene_orientation_and_words():
    ds = _small_diagrams(seed=9)
    fr = pyoracle.chord_frame(ds, 7300)
    v = fr["vertex"].view(np.uint32)
    for c, (vp, path, info, wbits) in enumerate(fr["curve"]):
        nv = info & 7
        assert nv in (2, 4) and path == c // 6
        assert vp - 24 * path in (0, 2, 4, 8, 12, 16, 20) and vp + nv <= 24 * (path + 1)
        f, l = v[vp], v[vp + nv - 1]
        assert f[1] < l[1] or (f[1] == l[1] and f[0] < l[0])
        assert wbits == np.float32(-1).view(np.uint32)
    assert np.array_equal(fr["vertex_path"], np.arange(fr["vertex_path"].shape[0]) // 24)
    assert not fr["fill_rule"].any()


def test_upload_form_orients_back_to_the_frame():
    """chord_scene hands the curves over in generation order; orienting them as an upload
    does (first vertex below the last by bit pattern) gives back the frame's vertices and
    reversed flags."""
    ds = _small_diagrams(seed=11)
    fr = pyoracle.chord_frame(ds, 8200)
    sc = pyoracle.chord_scene(fr)
    v = sc.vertex.copy()
    bits = v.view(np.uint32)
    flags = []
    for vp, nv in zip(sc.curve_vertex_pos, sc.curve_type & 7):
        f, l = bits[vp], bits[vp + nv - 1]
        rev = not (f[1] < l[1] or (f[1] == l[1] and f[0] < l[0]))
        if rev:
            v[vp:vp + nv] = v[vp:vp + nv][::-1].copy()
        flags.append(int(rev))
    np.testing.assert_array_equal(v.view(np.uint32), fr["vertex"].view(np.uint32))
    np.testing.assert_array_equal(flags, (fr["curve"][:, 2] >> 8) & 1)
    assert 0 < sum(flags) < len(flags)
