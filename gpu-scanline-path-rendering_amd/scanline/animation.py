"""Chord-diagram animation: keyframes built on the host, each frame's scene made on the GPU.

The reference's animation mode (``-animation``) draws two chord diagrams of refugee flows,
one by country of origin and one by destination, and moves them from year to year
(gpu-scanline/src/rasterizer/kernel/animation.cu):

* ``diagrams_from_records`` -- the record filtering and per-year link maps of
  load_refugee_animation (animation.cu:1480-1706);
* ``chord_keyframes`` -- RefugeeAnimation::addFrame, calculatePosition and genFrame
  (animation.cu:1086-1330): node sizes and angles per keyframe, and per interval between
  keyframes the links present at either end with their sizes and positions;
* ``ChordAnimation`` -- the keyframes loaded once (sl_chord_load) and the scene of a frame
  generated on the device (sl_chord_frame; chord_animation, animation.cu:2435-2760).

The reference reads a CSV of UNHCR records that it does not ship; ``synthetic_records``
makes seeded records of the same shape.  Host arithmetic is float32, in the reference's
order of operations.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _abi

KEY_MS = 6000                                 # TIME_PER_YEAR (animation.cu:1708)
CANVAS = (1920, 1080)                         # the animation's window
CENTERS = ((480.0, 540.0), (1440.0, 540.0))   # source and target diagrams (animation.cu:2733-2734)
_F = np.float32


@dataclass
class Record:
    """One row of the flow table: year, origin and destination ids, and the two values."""
    year: int
    id0: int
    id1: int
    value0: float
    value1: float


@dataclass
class ChordKeyframes:
    """Keyframes of one diagram, laid out as sl_chord (include/scanline.h)."""
    n_keys: int
    n_nodes: int
    node_t0: np.ndarray      # [n_keys, n_nodes] f32
    node_t1: np.ndarray      # [n_keys, n_nodes] f32
    node_size: np.ndarray    # [n_keys, n_nodes] f32
    node_color: np.ndarray   # [n_nodes] u32 RGBA8
    link_offset: np.ndarray  # [n_keys] u32
    link_node: np.ndarray    # [n_links, 2] u32
    link_from: np.ndarray    # [n_links, 4] f32 {pos0, pos1, size0, size1} at keyframe k - 1
    link_to: np.ndarray      # [n_links, 4] f32 ... at keyframe k
    center: tuple = (0.0, 0.0)

    def as_c(self):
        c = _abi.SlChord()
        c.n_keys, c.n_nodes = self.n_keys, self.n_nodes
        c.node_t0, c.node_t1 = self.node_t0.ctypes.data, self.node_t1.ctypes.data
        c.node_size, c.node_color = self.node_size.ctypes.data, self.node_color.ctypes.data
        c.link_offset, c.link_node = self.link_offset.ctypes.data, self.link_node.ctypes.data
        c.link_from, c.link_to = self.link_from.ctypes.data, self.link_to.ctypes.data
        c.center_x, c.center_y = float(self.center[0]), float(self.center[1])
        return c

    def paths_at(self, time_ms):
        """Paths of this diagram's scene at time_ms: its nodes and its interval's links."""
        k = min(int(time_ms) // KEY_MS + 1, self.n_keys - 1)
        return self.n_nodes + int(self.link_offset[k]) - int(self.link_offset[k - 1])


def diagrams_from_records(records):
    """Per-year link maps of the source and target diagrams (load_refugee_animation,
    animation.cu:1521-1700).  Returns (years, source_maps, target_maps); a map takes
    (low id, high id) to [size0, size1] of the low and high end."""
    kept = []
    for r in records:
        v0, v1 = _F(r.value0), _F(r.value1)
        if v0 < 1000 and v1 < 1000:
            continue
        if r.id0 == r.id1:
            v0, v1 = _F(v0 + v1), _F(0)
        else:
            v0 = _F(1) if v0 == 0 else v0
            v1 = _F(1) if v1 == 0 else v1
        kept.append((r.year, r.id0, r.id1, v0, v1))
    kept.sort(key=lambda x: x[0])  # stable: equal keys keep the records' order
    years, src, dst = [], [], []
    for year, a, b, v0, v1 in kept:
        if not years or years[-1] != year:
            years.append(year)
            src.append({})
            dst.append({})
        key = (min(a, b), max(a, b))
        if a <= b:
            src[-1][key] = [v0, v1]
            dst[-1][key] = [v1, v0]
        else:
            src[-1][key] = [v1, v0]
            dst[-1][key] = [v0, v1]
    return years, src, dst


def _fold_self_links(links):
    """addFrame (animation.cu:1086-1122): a link from a node to itself keeps the sum at end 0."""
    out = {}
    for (a, b), (s0, s1) in links.items():
        out[(a, b)] = [_F(s0 + s1), _F(0)] if a == b else [_F(s0), _F(s1)]
    return out


def _positions(links, n_nodes):
    """calculatePosition (animation.cu:1134-1199) of one keyframe: the links' positions in
    their nodes, node sizes, and node angles.  Returns (pos, size, t0, t1); pos maps a link
    to [pos0, pos1].  Equal node sizes are ordered by id (the reference leaves their order
    to std::sort)."""
    size = [_F(0)] * n_nodes
    pos = {}
    for i in range(n_nodes):
        for j in range(i + 1):
            lk = links.get((j, i))
            if lk is None:
                continue
            pos[(j, i)] = [size[j], size[i]]
            if i == j:
                size[j] = _F(size[j] + lk[0])
            else:
                size[j] = _F(size[j] + lk[0])
                size[i] = _F(size[i] + lk[1])
    order = sorted(range(n_nodes), key=lambda n: (-float(size[n]), n))
    total = _F(0)
    npos = [_F(0)] * n_nodes
    for r, n in enumerate(order):
        total = _F(total + size[n])
        if r:
            prev = order[r - 1]
            npos[n] = _F(npos[prev] + size[prev])
    non_empty = sum(1 for n in range(n_nodes) if size[n] != 0)
    pi = _F(np.pi)
    gap = _F(_F(pi / _F(180)) * _F(2))
    total_gap = _F(gap * _F(non_empty))
    span = _F(_F(2) * pi - total_gap)
    t0 = [_F(0)] * n_nodes
    t1 = [_F(0)] * n_nodes
    before = 1
    for n in order:
        if size[n] != 0:
            before += 1
        base = _F(_F(before) * gap)
        # 0 / 0 when every node is empty: such a keyframe puts all nodes at the gap
        f0 = _F(npos[n] / total) if total != 0 else _F(0)
        f1 = _F(_F(npos[n] + size[n]) / total) if total != 0 else _F(0)
        t0[n] = _F(base + _F(f0 * span))
        t1[n] = _F(base + _F(f1 * span))
    return pos, size, t0, t1


def chord_keyframes(link_maps, n_nodes, colors, center=(0.0, 0.0)):
    """RefugeeAnimation::genAnimation (animation.cu:1125-1133, 1202-1330): keyframes of one
    diagram from its per-year link maps (at least two)."""
    if len(link_maps) < 2:
        raise ValueError("a chord animation needs at least two keyframes")
    K = len(link_maps)
    frames = [_fold_self_links(m) for m in link_maps]
    calc = [_positions(f, n_nodes) for f in frames]
    node_t0 = np.array([c[2] for c in calc], np.float32).reshape(K, n_nodes)
    node_t1 = np.array([c[3] for c in calc], np.float32).reshape(K, n_nodes)
    node_size = np.array([c[1] for c in calc], np.float32).reshape(K, n_nodes)
    offs, nodes, frm, to = [0], [], [], []
    for k in range(1, K):
        prev, cur = frames[k - 1], frames[k]
        ppos, cpos = calc[k - 1][0], calc[k][0]
        for i in range(n_nodes):
            for j in range(i + 1):
                key = (j, i)
                pl, cl = prev.get(key), cur.get(key)
                live = lambda l: l is not None and (l[0] != 0 or l[1] != 0)
                if not (live(pl) or live(cl)):
                    continue
                nodes.append(key)
                # a link missing at one end grows from (or shrinks to) position 0
                frm.append([ppos[key][0], ppos[key][1], pl[0], pl[1]] if pl is not None else [0, 0, 0, 0])
                to.append([cpos[key][0], cpos[key][1], cl[0], cl[1]] if cl is not None else [0, 0, 0, 0])
        offs.append(len(nodes))
    L = len(nodes)
    return ChordKeyframes(
        n_keys=K, n_nodes=n_nodes, node_t0=np.ascontiguousarray(node_t0), node_t1=np.ascontiguousarray(node_t1),
        node_size=np.ascontiguousarray(node_size),
        node_color=np.ascontiguousarray(np.asarray(colors, dtype=np.int64).astype(np.uint32).reshape(n_nodes)),
        link_offset=np.asarray(offs, np.uint32),
        link_node=np.asarray(nodes, np.uint32).reshape(L, 2),
        link_from=np.asarray(frm, np.float32).reshape(L, 4), link_to=np.asarray(to, np.float32).reshape(L, 4),
        center=tuple(center))


def palette(n, seed=0):
    """Opaque RGBA8 node colours (r | g << 8 | b << 16 | 0xFF << 24), seeded."""
    rng = np.random.default_rng(seed)
    rgb = rng.integers(64, 224, size=(n, 3))
    return (rgb[:, 0] | (rgb[:, 1] << 8) | (rgb[:, 2] << 16) | (0xFF << 24)).astype(np.uint32)


def synthetic_records(n_nodes, n_years, links_per_year, seed=0, keep=0.8):
    """Seeded flow records: each year `links_per_year` distinct node pairs with values of
    1000 to ~1e6; a share `keep` of a year's pairs carries over to the next with new
    values, so that the animation both moves and adds and drops links."""
    rng = np.random.default_rng(seed)
    recs = []
    pairs = []
    for y in range(n_years):
        kept = [p for p in pairs if rng.random() < keep]
        seen = set(kept)
        while len(kept) < links_per_year:
            a, b = int(rng.integers(0, n_nodes)), int(rng.integers(0, n_nodes))
            if rng.random() < 0.97 and a == b:
                continue
            if (a, b) in seen or (b, a) in seen:
                continue
            seen.add((a, b))
            kept.append((a, b))
        pairs = kept
        for a, b in pairs:
            v = np.exp(rng.uniform(np.log(1000.0), np.log(1e6), size=2)).astype(np.float32)
            if rng.random() < 0.3:
                v[rng.integers(0, 2)] = 0
            recs.append(Record(2000 + y, a, b, float(v[0]), float(v[1])))
    return recs


def refugee_diagrams(n_nodes=150, n_years=4, links_per_year=1500, seed=0):
    """The source and target diagrams of seeded records, centred as the reference places
    them on its 1920 x 1080 window."""
    recs = synthetic_records(n_nodes, n_years, links_per_year, seed)
    _, src, dst = diagrams_from_records(recs)
    colors = palette(n_nodes, seed)
    return [chord_keyframes(src, n_nodes, colors, CENTERS[0]), chord_keyframes(dst, n_nodes, colors, CENTERS[1])]


class ChordAnimation:
    """Keyframes held on the device of a rasterizer; frame(t) replaces its scene with the
    scene at t milliseconds, generated on the GPU (vg_animation, animation.cu:2767-2786)."""

    def __init__(self, rasterizer, diagrams):
        self.r = rasterizer
        self.diagrams = list(diagrams)
        arr = (_abi.SlChord * len(self.diagrams))(*[d.as_c() for d in self.diagrams])
        self.r._chk(self.r._L.sl_chord_load(self.r._h, arr, len(self.diagrams)))

    def frame(self, time_ms, stream=None):
        self.r._chk(self.r._L.sl_chord_frame(self.r._h, int(time_ms), ctypes.c_void_p(stream or 0)))

    def paths_at(self, time_ms):
        return sum(d.paths_at(time_ms) for d in self.diagrams)
