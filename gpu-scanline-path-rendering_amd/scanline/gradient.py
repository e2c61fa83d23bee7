"""Gradient paint textures: the per-path gradient table and the colour ramp.

Restates the texture construction of `Mochimazui::RVG::load` (gpu-scanline/src/rvg.cpp:
490-758) in float32 arithmetic:

* every painted path gets one 12-word table slot, in path order; a solid paint leaves its
  slot empty (rvg.cpp:488-489), so the renderer finds path p's gradient at slot p;
* a gradient slot holds the inverse of the paint transform mapped into a gradient space
  where the ramp parameter is c = |d|^2 / (sqrt(B^2 + |d|^2 f.z) - B) with d = g - f and
  B = d . f (output_32.frag.glsl:45-51): radial gradients are scaled by 1/r about the
  centre with focal point f = (focus - centre) / r; linear ones are projected on the
  gradient vector and given the focal point (-1, 0), which makes c the linear parameter
  (rvg.cpp:514-577);
* the ramp of a gradient has n = a power of two in [2, 1024] texels, n >= 16 / (smallest
  stop gap) (rvg.cpp:616-628), sampled at t_i = i / (n - 1) between its stops, in linear
  light when sRGB correction is on (rvg.cpp:629-700); ramps are packed into 1024-texel rows
  by decreasing n so that no ramp straddles a row (rvg.cpp:718-746);
* table words: tr[0..5] (float), ramp position p (int), (n - 1) / 1024 (float), focal
  point x, y, z (float), 0 (vg_container.h:59-64, rvg.cpp:747-752).
"""
import math

import numpy as np

_F = np.float32


def safe_rcp(a):
    """safeRcp of rvg.cpp:38-40."""
    a = _F(a)
    return _F(_F(1.0) / a) if a != 0 else _F(0.0)


def srgb_to_lrgb(f):
    """color.h:105-110 in float (powf)."""
    f = _F(f)
    if f <= _F(0.04045):
        return _F(f / _F(12.92))
    return _F(np.power(_F((f + _F(0.055)) / _F(_F(1.0) + _F(0.055))), _F(2.4)))


def lrgb_to_srgb(f):
    """color.h:112-117 in float (powf)."""
    f = _F(f)
    if f <= _F(0.0031308):
        return _F(_F(12.92) * f)
    return _F(_F(_F(1.0) + _F(0.055)) * _F(np.power(f, _F(_F(1.0) / _F(2.4)))) - _F(0.055))


def _cvt_i32(v):
    """(int) of a float as x86 converts it: out of range and NaN give INT_MIN."""
    v = float(v)
    if not math.isfinite(v) or v >= 2147483648.0 or v <= -2147483649.0:
        return -2147483648
    return int(math.trunc(v))


def f32tou8(a):
    """f32tou8 of rvg.cpp:26-31: (int)(a * 255) clamped to [0, 255]."""
    r = _cvt_i32(_F(_F(a) * _F(255.0)))
    return 0 if r < 0 else (255 if r > 255 else r)


def _bits_i32(x):
    return int(np.array([_F(x)], dtype=np.float32).view(np.int32)[0])


class GradientTextures:
    """Collects table slots and ramps path by path; `textures()` packs them."""

    def __init__(self, srgb):
        self.srgb = bool(srgb)
        self._items = []   # [tr(6), p, n, focal(3)] per painted path
        self._ramp = []    # ramp texels (int RGBA8) in item order

    def add_solid(self):
        """The empty slot of a solid-paint path (rvg.cpp:488-489)."""
        self._items.append([[_F(0)] * 6, 0, 0, [_F(0)] * 3])

    def add(self, kind, x0, y0, x1, y1, r, stops, opacity, tr):
        """A gradient paint: kind 'linear' or 'radial'; (x0, y0) the start / centre, (x1, y1)
        the end / focus, r the radius; stops [(offset, r, g, b, a)] in [0, 1]; tr the six
        paint-transform words a b c d e f mapping (x, y) to (a x + b y + c, d x + e y + f)."""
        tr = [_F(v) for v in tr]
        x0, y0, x1, y1, r = _F(x0), _F(y0), _F(x1), _F(y1), _F(r)
        det = _F(float(np.float64(tr[0]) * np.float64(tr[4]) - np.float64(tr[1]) * np.float64(tr[3])))
        idet = safe_rcp(det)
        ti = [_F(0)] * 6
        ti[0] = _F(tr[4] * idet)
        ti[3] = _F(-tr[3] * idet)
        ti[1] = _F(-tr[1] * idet)
        ti[4] = _F(tr[0] * idet)
        ti[2] = _F(-float(np.float64(tr[2]) * np.float64(ti[0]) + np.float64(tr[5]) * np.float64(ti[1])))
        ti[5] = _F(-float(np.float64(tr[2]) * np.float64(ti[3]) + np.float64(tr[5]) * np.float64(ti[4])))
        ti[2] = _F(ti[2] - x0)
        ti[5] = _F(ti[5] - y0)
        if kind == "radial":
            ir = safe_rcp(r)
            ti = [_F(v * ir) for v in ti]
            focal = [_F(_F(x1 - x0) * ir), _F(_F(y1 - y0) * ir)]
        else:
            dx, dy = _F(x1 - x0), _F(y1 - y0)
            s2 = _F(_F(2.0) / _F(_F(dx * dx) + _F(dy * dy)))
            dx, dy = _F(dx * s2), _F(dy * s2)
            t0 = _F(_F(ti[0] * dx) + _F(ti[3] * dy))
            t1 = _F(_F(ti[1] * dx) + _F(ti[4] * dy))
            t2 = _F(_F(ti[2] * dx) + _F(ti[5] * dy))
            ti = [t0, t1, _F(t2 - _F(1.0)), _F(0), _F(0), _F(0)]
            focal = [_F(-1.0), _F(0.0)]
        fz = _F(_F(_F(1.0) - _F(focal[0] * focal[0])) - _F(focal[1] * focal[1]))
        if not (fz > 0):
            fz = _F(0.0)
        focal.append(fz)

        knots = []
        for off, cr, cg, cb, ca in stops:
            knots.extend([_F(off), _F(cr), _F(cg), _F(cb), _F(_F(ca) * _F(opacity))])
        if not knots:
            knots = [_F(0)] * 5
        if len(knots) <= 5:
            knots = knots[:5] + knots[:5]
            knots[0], knots[5] = _F(0.0), _F(1.0)
        min_delta = _F(1.0)
        for i in range(5, len(knots), 5):
            d = _F(knots[i] - knots[i - 5])
            if d > 0 and d < min_delta:
                min_delta = d
        res = _cvt_i32(_F(_F(16.0) / min_delta))
        res = min(max(res, 1), 1023)
        for s in (1, 2, 4, 8):
            res |= res >> s
        res += 1
        p0 = len(self._ramp)
        ires = _F(_F(1.0) / _F(res - 1))
        pk = 0
        for i in range(res):
            t_i = _F(_F(i) * ires)
            while pk + 10 < len(knots):
                a, b = knots[pk], knots[pk + 5]
                if t_i > b or a == b:
                    pk += 5
                else:
                    break
            a, b = knots[pk], knots[pk + 5]
            tl = _F(_F(t_i - a) * safe_rcp(_F(b - a)))
            if not (tl > 0):
                tl = _F(0.0)
            if tl > 1:
                tl = _F(1.0)
            c0 = knots[pk + 1:pk + 5]
            c1 = knots[pk + 6:pk + 10]
            if self.srgb:
                # color.h:163-171: lerp in linear light, back to sRGB
                rgb = []
                for ch in range(3):
                    l0, l1 = srgb_to_lrgb(c0[ch]), srgb_to_lrgb(c1[ch])
                    ll = _F(_F(l0 * _F(_F(1.0) - tl)) + _F(l1 * tl))
                    rgb.append(f32tou8(lrgb_to_srgb(ll)))
            else:
                # rvg.cpp:44-46: a + (b - a) t
                rgb = [f32tou8(_F(c0[ch] + _F(_F(c1[ch] - c0[ch]) * tl))) for ch in range(3)]
            al = f32tou8(_F(c0[3] + _F(_F(c1[3] - c0[3]) * tl)))
            self._ramp.append(rgb[0] | (rgb[1] << 8) | (rgb[2] << 16) | (al << 24))
        self._items.append([ti, p0, res, focal])

    def textures(self):
        """(ramp int32[h * 1024], table int32[12 * slots]).  Ramps are packed by decreasing
        texel count (a stable order among equal counts)."""
        n = len(self._items)
        order = sorted(range(n), key=lambda i: -self._items[i][2])
        h = (len(self._ramp) + 1023) // 1024
        ramp = np.zeros(h * 1024, dtype=np.int64)
        pos = [0] * n
        out = 0
        for i in order:
            p0, cnt = self._items[i][1], self._items[i][2]
            pos[i] = out
            if cnt:
                ramp[out:out + cnt] = [v & 0xFFFFFFFF for v in self._ramp[p0:p0 + cnt]]
                out += cnt
        table = []
        for i, (ti, _, cnt, focal) in enumerate(self._items):
            words = [_bits_i32(v) for v in ti] + [pos[i], _bits_i32(_F(_F(cnt - 1) * _F(1.0 / 1024.0)))]
            words += [_bits_i32(v) for v in focal] + [0]
            table.extend(words)
        if not table:
            table = [0] * 12
        return (ramp.astype(np.uint32).view(np.int32),
                np.array([w & 0xFFFFFFFF for w in table], dtype=np.uint32).view(np.int32))
