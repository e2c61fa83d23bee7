// scene_gen.cpp -- seeded synthetic scenes for the benchmark workloads.
//
// The reference has no generator (its benchmark scenes are external files, README.md:7,
// 49-53); these produce scenes in the reference's VGContainer layout (vertices stored
// per curve, vg_container.cpp:97-125; fill_info as u8rgba little-endian, color.h:29-57).
// PRNG: splitmix64, one substream per path, so any path can be regenerated alone.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/scanline.h"

struct sl_scene_obj {
    std::vector<float> vertex;
    std::vector<uint32_t> cvpos, cpath;
    std::vector<uint8_t> ctype;
    std::vector<float> cw;
    std::vector<uint8_t> frule;
    std::vector<uint32_t> fcolor;
};

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double uni(double a, double b) { return a + (b - a) * u01(); }
    int ui(int a, int b) { return a + (int)(next() % (uint64_t)(b - a + 1)); }  // [a,b]
};

inline uint64_t path_seed(uint64_t seed, uint32_t p) {
    Rng r(seed ^ (0xD1B54A32D192ED03ull * ((uint64_t)p + 1)));
    return r.next();
}

struct Builder {
    sl_scene_obj &o;
    uint32_t path;
    float first_x = 0, first_y = 0;
    explicit Builder(sl_scene_obj &obj, uint32_t p) : o(obj), path(p) {}
    void curve(uint8_t type, const double *xy, float w = -1.f) {
        o.cvpos.push_back((uint32_t)(o.vertex.size() / 2));
        o.ctype.push_back(type);
        o.cpath.push_back(path);
        o.cw.push_back(w);
        for (int k = 0; k < (type & 7); ++k) {
            o.vertex.push_back((float)xy[2 * k]);
            o.vertex.push_back((float)xy[2 * k + 1]);
        }
    }
};

uint32_t rand_color(Rng &r, double amin) {
    uint32_t cr = (uint32_t)r.ui(0, 255), cg = (uint32_t)r.ui(0, 255), cb = (uint32_t)r.ui(0, 255);
    uint32_t a = amin >= 1.0 ? 255u : (uint32_t)std::lround(255.0 * r.uni(amin, 1.0));
    return cr | (cg << 8) | (cb << 16) | (a << 24);
}

// kind 1: closed chains of cubics, control points uniform in the canvas, nonzero, opaque.
void gen_cubic_chain(sl_scene_obj &o, uint32_t p, Rng &r, int W, int H, int mn, int mx) {
    Builder b(o, p);
    int n = r.ui(mn, mx);
    double x0 = r.uni(0, W), y0 = r.uni(0, H);
    double px = x0, py = y0;
    for (int k = 0; k < n; ++k) {
        double nx = (k + 1 == n) ? x0 : r.uni(0, W), ny = (k + 1 == n) ? y0 : r.uni(0, H);
        double c1x = r.uni(0, W), c1y = r.uni(0, H), c2x = r.uni(0, W), c2y = r.uni(0, H);
        double xy[8] = {px, py, c1x, c1y, c2x, c2y, nx, ny};
        b.curve(SL_CURVE_CUBIC, xy);
        px = nx; py = ny;
    }
    o.frule.push_back(SL_FILL_NONZERO);
    o.fcolor.push_back(rand_color(r, 1.0));
}

int gcd(int a, int b) { while (b) { int t = a % b; a = b; b = t; } return a; }

// kind 2: mixed segments in log-uniform bboxes; 10% star polygons; alternating rules.
void gen_mixed(sl_scene_obj &o, uint32_t p, Rng &r, int W, int H, int mn, int mx) {
    Builder b(o, p);
    double smax = std::min(1024.0, (double)std::min(W, H));
    double s = std::exp(r.uni(std::log(4.0), std::log(smax)));
    double ox = r.uni(-0.1 * s, W - 0.9 * s), oy = r.uni(-0.1 * s, H - 0.9 * s);
    if (r.u01() < 0.1) {
        int k = r.ui(5, 17);
        int m = (k - 1) / 2;
        int g = gcd(k, m);
        double cx = ox + s / 2, cy = oy + s / 2, rad = s / 2;
        const double PI = 3.14159265358979323846;
        for (int j = 0; j < g; ++j) {
            int cnt = k / g;
            for (int e = 0; e < cnt; ++e) {
                int i0 = (j + e * m) % k, i1 = (j + (e + 1) * m) % k;
                double a0 = 2 * PI * i0 / k - PI / 2, a1 = 2 * PI * i1 / k - PI / 2;
                double xy[4] = {cx + rad * std::cos(a0), cy + rad * std::sin(a0), cx + rad * std::cos(a1),
                                cy + rad * std::sin(a1)};
                b.curve(SL_CURVE_LINEAR, xy);
            }
        }
    } else {
        int n = r.ui(mn, mx);
        auto rx = [&]() { return ox + s * r.u01(); };
        auto ry = [&]() { return oy + s * r.u01(); };
        double x0 = rx(), y0 = ry(), px = x0, py = y0;
        for (int k = 0; k < n; ++k) {
            bool last = (k + 1 == n);
            double u = r.u01();
            if (u < 0.4) {
                double nx = last ? x0 : rx(), ny = last ? y0 : ry();
                double xy[4] = {px, py, nx, ny};
                b.curve(SL_CURVE_LINEAR, xy);
                px = nx; py = ny;
            } else if (u < 0.7) {
                double cxx = rx(), cyy = ry();
                double nx = last ? x0 : rx(), ny = last ? y0 : ry();
                double xy[6] = {px, py, cxx, cyy, nx, ny};
                b.curve(SL_CURVE_QUADRATIC, xy);
                px = nx; py = ny;
            } else {
                double c1x = rx(), c1y = ry(), c2x = rx(), c2y = ry();
                double nx = last ? x0 : rx(), ny = last ? y0 : ry();
                double xy[8] = {px, py, c1x, c1y, c2x, c2y, nx, ny};
                b.curve(SL_CURVE_CUBIC, xy);
                px = nx; py = ny;
            }
        }
    }
    o.frule.push_back((uint8_t)(p & 1));
    o.fcolor.push_back(rand_color(r, 0.2));
}

// kind 3: glyph-like paths: 1-4 concentric jittered elliptical contours of quadratics,
// outer counter-clockwise, inner reversed (holes under nonzero).
void gen_glyph(sl_scene_obj &o, uint32_t p, Rng &r, int W, int H, int n_paths, int mn, int mx) {
    Builder b(o, p);
    int G = (int)std::ceil(std::sqrt((double)n_paths));
    double cw = (double)W / G, ch = (double)H / G;
    int gx = (int)(p % (uint32_t)G), gy = (int)(p / (uint32_t)G);
    double cx = (gx + 0.5 + r.uni(-0.25, 0.25)) * cw;
    double cy = (gy + 0.5 + r.uni(-0.25, 0.25)) * ch;
    double s = r.uni(8.0, 64.0);
    double ax = 0.5 * s * r.uni(0.6, 1.0), ay = 0.5 * s * r.uni(0.6, 1.0);
    int nc = r.ui(1, 4);
    const double PI = 3.14159265358979323846;
    for (int c = 0; c < nc; ++c) {
        double sc = 1.0 - 0.22 * c;
        int n = r.ui(mn, mx);
        std::vector<double> ang(n), rad(n);
        for (int k = 0; k < n; ++k) {
            ang[k] = 2 * PI * (k + r.uni(-0.3, 0.3)) / n;
            rad[k] = sc * r.uni(0.92, 1.08);
        }
        bool reverse = (c > 0);
        for (int e = 0; e < n; ++e) {
            int k0 = reverse ? (n - e) % n : e;
            int k1 = reverse ? (n - e - 1) % n : (e + 1) % n;
            double a0 = ang[k0], a1 = ang[k1];
            if (!reverse && k1 == 0) a1 += 2 * PI;
            if (reverse && k0 == 0) a0 += 2 * PI;
            double am = 0.5 * (a0 + a1), half = 0.5 * std::fabs(a1 - a0);
            double rm = 0.5 * (rad[k0] + rad[k1]) / std::cos(half);
            double xy[6] = {cx + ax * rad[k0] * std::cos(a0), cy + ay * rad[k0] * std::sin(a0),
                            cx + ax * rm * std::cos(am),      cy + ay * rm * std::sin(am),
                            cx + ax * rad[k1] * std::cos(a1), cy + ay * rad[k1] * std::sin(a1)};
            b.curve(SL_CURVE_QUADRATIC, xy);
        }
    }
    o.frule.push_back(SL_FILL_NONZERO);
    uint32_t col = rand_color(r, 1.0);
    o.fcolor.push_back(col);
}

}  // namespace

extern "C" int sl_scene_generate(int kind, int width, int height, int n_paths, uint64_t seed,
                                 int min_segs, int max_segs, sl_scene_obj **out) {
    if (!out || width <= 0 || height <= 0 || n_paths < 0 || kind < 1 || kind > 3) return SL_E_ARG;
    *out = nullptr;
    int mn = min_segs, mx = max_segs;
    if (mn <= 0 || mx <= 0) {
        if (kind == 1) { mn = 4; mx = 12; }
        else if (kind == 2) { mn = 3; mx = 64; }
        else { mn = 10; mx = 40; }
    }
    if (mx < mn) return SL_E_ARG;
    sl_scene_obj *o = new (std::nothrow) sl_scene_obj;
    if (!o) return SL_E_OOM;
    try {
        for (int p = 0; p < n_paths; ++p) {
            Rng r(path_seed(seed, (uint32_t)p));
            if (kind == 1) gen_cubic_chain(*o, (uint32_t)p, r, width, height, mn, mx);
            else if (kind == 2) gen_mixed(*o, (uint32_t)p, r, width, height, mn, mx);
            else gen_glyph(*o, (uint32_t)p, r, width, height, n_paths, mn, mx);
        }
    } catch (...) {
        delete o;
        return SL_E_OOM;
    }
    *out = o;
    return SL_OK;
}

extern "C" int sl_scene_view(const sl_scene_obj *o, sl_scene *s) {
    if (!o || !s) return SL_E_ARG;
    s->n_paths = (uint32_t)o->frule.size();
    s->n_curves = (uint32_t)o->ctype.size();
    s->n_vertices = (uint32_t)(o->vertex.size() / 2);
    s->vertex = o->vertex.data();
    s->curve_vertex_pos = o->cvpos.data();
    s->curve_type = o->ctype.data();
    s->curve_path_id = o->cpath.data();
    s->curve_arc_w = o->cw.data();
    s->fill_rule = o->frule.data();
    s->fill_color = o->fcolor.data();
    s->gradient_ramp = nullptr;
    s->gradient_ramp_size = 0;
    s->gradient_table = nullptr;
    s->gradient_table_size = 0;
    return SL_OK;
}

extern "C" int sl_scene_free(sl_scene_obj *o) {
    delete o;
    return SL_OK;
}
