/*
 * scanline_oracle.cpp -- CPU oracle of the comb-scanline rasterizer.
 *
 * TEST INFRASTRUCTURE ONLY: a literal single-threaded restatement of the reference's
 * per-scanline algorithm, used as the parity checker for the HIP renderer and (built
 * with -fopenmp) as the reported CPU baseline.  Nothing in the product links it.
 *
 * Build: -O2 -ffp-contract=off, no -ffast-math, SSE scalar float (x86-64 default).
 * Arithmetic is written in the reference's operation order; every place where the
 * reference relies on GPU-specific behaviour is made explicit:
 *   - __float2int_rd / (int) casts saturate and map NaN to 0 (CUDA cvt semantics);
 *   - shifts by a count outside [0, width) give 0 (PTX clamps the count);
 *   - round() is half away from zero; (int)(x*k + 0.5) is done in double, as the
 *     reference's double literal promotes it;
 *   - normalize(v) = v * (1/sqrtf(dot(v,v))) (the reference uses rsqrtf);
 *   - a 2D point-sampled, clamped, normalized texture fetch reads texel
 *     clamp(floor(u*256), 0, 255);
 *   - the OpenGL compositor is restated as src-over blending into an 8-bit
 *     framebuffer that is re-quantised after every blend.
 * Pinning (tests/, CPU): orientation, transform, nudge, culling flags, monotone split and
 * curved intersections by test_oracle_geometry_pins.py, whose failure on planted
 * mistakes is checked by test_oracle_mutations.py; sample sets, tables, masks, solvers,
 * scans, stencil and compositing by test_oracle_pinning.py, including the reference's own
 * table tolerance test (ras_qm_mask.h:600-714).
 * Citations are file:line into the reference tree.
 */
#include "scanline_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------------
// scalar helpers
// ---------------------------------------------------------------------------------
inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// __float2int_rd (ras_cut.h:639-657 uses it through d_float2int_rd).
inline int32_t f2i_rd(float x) {
    if (x != x) return 0;
    float f = std::floor(x);
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f < -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}
// C (int) of a float, saturating.
inline int32_t f2i_rz(float x) {
    if (x != x) return 0;
    float f = std::trunc(x);
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f < -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}
// C (int) of a double, saturating.
inline int32_t d2i_rz(double x) {
    if (x != x) return 0;
    double f = std::trunc(x);
    if (f >= 2147483648.0) return INT32_MAX;
    if (f < -2147483648.0) return INT32_MIN;
    return (int32_t)f;
}
inline int32_t addw(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
inline int32_t subw(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
inline uint32_t shr32(uint32_t v, int32_t s) { return (s < 0 || s >= 32) ? 0u : (v >> s); }
inline uint64_t shr64(uint64_t v, int32_t s) { return (s < 0 || s >= 64) ? 0ull : (v >> s); }

// Rasterizer::min / max (ras_device_func.h:20-28): plain ternaries.
template <class T> inline T tmin(T a, T b) { return a < b ? a : b; }
template <class T> inline T tmax(T a, T b) { return a > b ? a : b; }

// safeRcp (cutil_math.h:862-864)
inline float safeRcp(float a) { return (a != 0.f) ? (1.f / a) : 0.f; }
// lerp (cutil_math.h:1019-1029)
inline float lerp1(float a, float b, float t) { return a + t * (b - a); }

enum { CT_LINEAR = 2, CT_QUAD = 3, CT_CUBIC = 4, CT_RATIONAL = 0x13 };

// interpolateGeneralCurve (ras_cut.h:586-632) == d_curve_point per component
// (bezier_curve.h:57-152): repeated lerps; rational divides by the weight polynomial.
inline float curve_eval(int type, float w, const float *v, float t) {
    switch (type) {
    case CT_LINEAR:
        return lerp1(v[0], v[1], t);
    case CT_QUAD: {
        float l0 = lerp1(v[0], v[1], t);
        float l1 = lerp1(v[1], v[2], t);
        return lerp1(l0, l1, t);
    }
    case CT_CUBIC: {
        float q0 = lerp1(v[0], v[1], t);
        float q1 = lerp1(v[1], v[2], t);
        float q2 = lerp1(v[2], v[3], t);
        float l0 = lerp1(q0, q1, t);
        float l1 = lerp1(q1, q2, t);
        return lerp1(l0, l1, t);
    }
    default: {  // CT_RATIONAL
        float iv1 = v[1] * w;
        float l0 = lerp1(v[0], iv1, t);
        float l1 = lerp1(iv1, v[2], t);
        float r = lerp1(l0, l1, t);
        float omt = 1.f - t;
        float ww = omt * omt + 2.f * omt * t * w + t * t;
        return r * safeRcp(ww);
    }
    }
}

// solveQuadEquation (ras_cut.h:33-64)
inline void solve_quad(float a, float b, float c, float *r0, float *r1) {
    if (a == 0.f) {
        float x = -c / b;
        *r0 = x;
        *r1 = x;
        return;
    }
    float A = a;
    float B = b * 0.5f;
    float C = c;
    float tx, ty;
    float R = B * B - A * C;
    if (R > 0.0f) {
        float SR = std::sqrt(R);
        if (B > 0.0f) {
            float TB = B + SR;
            tx = -C / TB;
            ty = -TB / A;
        } else {
            float TB = -B + SR;
            tx = TB / A;
            ty = C / TB;
        }
    } else {
        tx = 0.f;
        ty = 0.f;
    }
    *r0 = tx;
    *r1 = ty;
}

// solveQuadBezierQM (ras_cut.h:373-388)
inline float solve_quad_bezier_qm(float iA, float iB, float iC, float iD, float t0, float t1) {
    float a = iA + (-2.f) * iB + iC;
    float b = -2.f * iA + 2.f * iB;
    float c = iA - iD;
    float r0 = 0.f, r1 = 0.f;
    solve_quad(a, b, c, &r0, &r1);
    if (t0 <= r0 && r0 <= t1) return r0;
    if (t0 <= r1 && r1 <= t1) return r1;
    return t0;
}

// ---------------------------------------------------------------------------------
// sample positions and mask tables (ras_qm_mask.cu:105-175, ras_qm_mask.h:424-598,
// 716-812).  16 spp has no reference layout: its samples follow the 8-spp rule
// (x = bit reversal of the index, y = index) over 16 positions and its masks keep one
// 32-bit word per pixel column with bit = row*16 + sample.
// ---------------------------------------------------------------------------------
const double kVg32X[32] = {
    -0.49479166651144624000, 0.02604168653488159200, -0.24479165673255920000, 0.27604168653488159000,
    -0.36979167163372040000, 0.15104168653488159000, -0.09895834326744079600, 0.38020831346511841000,
    -0.43229166418313980000, 0.06770831346511840800, -0.18229165673255920000, 0.33854168653488159000,
    -0.30729167163372040000, 0.21354168653488159000, -0.05729165673255920400, 0.46354168653488159000,
    -0.46354166790843010000, 0.05729168653488159200, -0.19270834326744080000, 0.30729168653488159000,
    -0.31770832836627960000, 0.16145831346511841000, -0.06770834326744079600, 0.43229168653488159000,
    -0.40104166418313980000, 0.11979168653488159000, -0.13020834326744080000, 0.36979168653488159000,
    -0.27604167163372040000, 0.24479168653488159000, -0.02604165673255920400, 0.47395831346511841000,
};  // vg_32_x (ras_qm_mask.cu:111-113); vg_32_y is (i+1/2)/32 - 1/2 (ras_qm_mask.cu:114-116)

// bit reversal rule of vg_sample_position (ras_qm_mask.cu:131-160) for n = 8 or 16.
void sample_positions(int n, float *xy) {
    if (n == 32) {
        for (int i = 0; i < 32; ++i) {
            xy[2 * i] = (float)(kVg32X[i] + 0.5);
            double y = -0.484375 + 0.03125 * i;  // vg_32_y[i]
            xy[2 * i + 1] = (float)(y + 0.5);
        }
        return;
    }
    double gap = 1.0 / n;
    double hgap = gap / 2;
    for (int i = 0; i < n; ++i) {
        int bc = 0, x = i, y = 0;
        while (x) { y <<= 1; y |= x & 1; x >>= 1; ++bc; }
        int scale = 1 << bc;
        double f = y / (float)scale;
        int yy = (int)(f * n);
        xy[2 * i] = (float)(yy * gap + hgap);
        xy[2 * i + 1] = (float)(i * gap + hgap);
    }
    // already sorted by (y, x) (sort_samples, ras_qm_mask.cu:44-48)
}

inline int words_for(int n) { return n == 8 ? 1 : (n == 16 ? 2 : 4); }

// QM_Mask_2x2::makeTable (ras_qm_mask.h:432-484), generated on the host.
// Cell samples are the pixel samples replicated to the 4 pixels and scaled by 1/2
// (init_qm_mask_table, ras_qm_mask.h:737-759).  Bit layout per texel:
//   8 spp : word0 bit (px*2+py)*8 + s
//   16 spp: word px bit py*16 + s
//   32 spp: word (px*2+py) bit s
void make_pixel_table(int n, uint32_t *out) {
    float s[64];
    sample_positions(n, s);
    const int nw = words_for(n);
    // cell sample coordinates per (pixel q = px*2+py, sample)
    std::vector<float> cx(4 * n), cy(4 * n);
    for (int q = 0; q < 4; ++q) {
        float ox = (q >= 2) ? 1.f : 0.f;  // q = 0:(0,0) 1:(0,1) 2:(1,0) 3:(1,1)
        float oy = (q & 1) ? 1.f : 0.f;
        for (int i = 0; i < n; ++i) {
            cx[q * n + i] = (s[2 * i] + ox) * .5f;
            cy[q * n + i] = (s[2 * i + 1] + oy) * .5f;
        }
    }
    for (int y = 0; y < 256; ++y) {
        for (int x = 0; x < 256; ++x) {
            float xf = ((float)x + 0.5f) * (1.f / 256.f);
            float yf = ((float)y + 0.5f) * (1.f / 256.f);
            float nrx = 2.f * (xf - 0.5f);
            float nry = 2.f * (yf - 0.5f);
            float lg2 = nrx * nrx + nry * nry;
            float lg = std::sqrt(lg2);
            float inv = 1.f / lg;
            float Nx = inv * nrx, Ny = inv * nry;
            float C = tmax(1.f - lg, 0.f) * (1.f / 0.5f);
            if (Nx < 0.f) { Nx = -Nx; Ny = -Ny; C = -C; }
            if (x == 37 && y == 218) {  // "hack for the worst" (ras_qm_mask.h:453-456)
                float vx = 1.f, vy = -1.f;
                float il = 1.f / std::sqrt(vx * vx + vy * vy);
                Nx = vx * il; Ny = vy * il;
                C = 0.f;
            }
            C += 0.5f * (Nx + Ny);
            uint32_t w[4] = {0, 0, 0, 0};
            for (int q = 0; q < 4; ++q) {
                int px = q >> 1, py = q & 1;
                for (int i = 0; i < n; ++i) {
                    float d = (Nx * cx[q * n + i] + Ny * cy[q * n + i]) - C;
                    if (d > 0.f) {
                        if (n == 8) w[0] |= 1u << (q * 8 + i);
                        else if (n == 16) w[px] |= 1u << (py * 16 + i);
                        else w[q] |= 1u << i;
                    }
                }
            }
            for (int k = 0; k < nw; ++k) out[(size_t)(y * 256 + x) * nw + k] = w[k];
        }
    }
}

struct Tables {
    int n = 0;
    std::vector<uint32_t> pixel;  // 65536 * words
};

inline int texel_index(float u) {
    int i = f2i_rd(u * 256.f);
    return tmin(tmax(i, 0), 255);
}

// fetchPixelTable (ras_qm_mask.h:519-542)
inline void fetch_pixel(const Tables &T, float p0x, float p0y, float p1x, float p1y, uint32_t *m) {
    if (p1y < p0y) { std::swap(p0x, p1x); std::swap(p0y, p1y); }
    float Dx = p1x - p0x, Dy = p1y - p0y;
    float vx = Dy, vy = -Dx;
    float il = 1.f / std::sqrt(vx * vx + vy * vy);
    float Nx = vx * il, Ny = vy * il;
    float C = Nx * p0x + Ny * p0y;
    C -= 0.5f * (Nx + Ny);
    float Cs = bitsf((fbits(C) & 0x80000000u) + 0x3f800000u);
    float k = tmax(1.f - C * Cs * 0.5f, 0.f) * Cs;
    float nrx = k * Nx, nry = k * Ny;
    float u = nrx * 0.5f + 0.5f;
    float v = nry * 0.5f + 0.5f;
    int ix = texel_index(u), iy = texel_index(v);
    const int nw = words_for(T.n);
    const uint32_t *t = &T.pixel[(size_t)(iy * 256 + ix) * nw];
    for (int i = 0; i < 4; ++i) m[i] = i < nw ? t[i] : 0u;
}

// fetchVerticalTable (ras_qm_mask.h:545-560)
inline void vertical_mask(int n, float y, uint32_t *m) {
    if (n == 8) {
        int32_t r = f2i_rz(std::round(y * 16.f));
        uint32_t mask = shr32(0xffffu, subw(16, r));
        mask |= mask << 16;
        m[0] = mask; m[1] = m[2] = m[3] = 0;
    } else if (n == 16) {
        int32_t r = f2i_rz(std::round(y * 32.f));
        uint32_t mask = shr32(0xffffffffu, subw(32, r));
        m[0] = m[1] = mask; m[2] = m[3] = 0;
    } else {
        int32_t r = f2i_rz(std::round(y * 64.f));
        uint64_t mask = shr64(0xFFFFFFFFFFFFFFFFull, subw(64, r));
        uint32_t m0 = (uint32_t)mask, m1 = (uint32_t)(mask >> 32);
        m[0] = m0; m[1] = m1; m[2] = m0; m[3] = m1;
    }
}

// fetchVerticalTable2 (ras_qm_mask.h:562-598)
inline void vertical_mask2(int n, float y0, float y1, uint32_t *m) {
    int k = n == 8 ? 16 : (n == 16 ? 32 : 64);
    int32_t s0 = subw(k, d2i_rz((double)(y0 * (float)k) + 0.5));
    int32_t s1 = subw(k, d2i_rz((double)(y1 * (float)k) + 0.5));
    s0 = tmax(s0, 0);
    s1 = tmax(s1, 0);
    if (n == 8) {
        uint32_t mask = shr32(0xffffu, s0) ^ shr32(0xffffu, s1);
        mask |= mask << 16;
        m[0] = mask; m[1] = m[2] = m[3] = 0;
    } else if (n == 16) {
        uint32_t mask = shr32(0xffffffffu, s0) ^ shr32(0xffffffffu, s1);
        m[0] = m[1] = mask; m[2] = m[3] = 0;
    } else {
        uint64_t mask = shr64(~0ull, s0) ^ shr64(~0ull, s1);
        uint32_t m0 = (uint32_t)mask, m1 = (uint32_t)(mask >> 32);
        m[0] = m0; m[1] = m1; m[2] = m0; m[3] = m1;
    }
}

// ---------------------------------------------------------------------------------
// colour tables for the compositor (output_32.vert.glsl:39-51; GL sRGB8 framebuffer)
// dec[k]: the linear value of 8-bit code k; thr[k]: smallest float that encodes to k
// (k >= 1), i.e. code = #{k >= 1 : value >= thr[k]}.
// ---------------------------------------------------------------------------------
double srgb_decode_d(double s) {
    return s <= 0.04045 ? s / 12.92 : std::pow((s + 0.055) / 1.055, 2.4);
}
float ceil_to_float(double d) {
    float f = (float)d;
    if ((double)f < d) f = std::nextafter(f, INFINITY);
    // make it the smallest float >= d
    while (true) {
        float g = std::nextafter(f, -INFINITY);
        if ((double)g >= d) f = g; else break;
    }
    return f;
}
void color_tables(int srgb, float *dec, float *thr) {
    for (int k = 0; k < 256; ++k) {
        dec[k] = srgb ? (float)srgb_decode_d(k / 255.0) : (float)k / 255.f;
        if (k == 0) { thr[k] = -INFINITY; continue; }
        double m = (k - 0.5) / 255.0;
        thr[k] = ceil_to_float(srgb ? srgb_decode_d(m) : m);
    }
}
inline int encode8(const float *thr, float v) {
    int lo = 0, hi = 255;  // largest k with v >= thr[k]; thr[0] = -inf
    if (!(v >= thr[1])) return 0;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (v >= thr[mid]) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------------
// pipeline state
// ---------------------------------------------------------------------------------
struct Scene {
    uint32_t P = 0, C = 0, V = 0;
    std::vector<float> vx, vy;           // vertices (unified order)
    std::vector<uint32_t> cvpos, cpath;  // per curve
    std::vector<uint8_t> ctype, crev;
    std::vector<float> cw;
    std::vector<uint8_t> frule;
    std::vector<uint32_t> finfo;
    std::vector<uint32_t> vpath;  // vertex_path_id
    // gradient textures (ras_base.cpp:339-371): ramp texels in rows of 1024, 12-word slots
    std::vector<int32_t> ramp, table;
    uint32_t gslots = 0, rows = 0;
};

struct Result {
    int W = 0, H = 0, N = 0, srgb = 0, ms = 0;
    double seconds = 0;
    int64_t F = 0, M = 0, S = 0, pixel_count = 0;
    std::vector<float> tvx;  // transformed vertices, x,y interleaved
    std::vector<uint8_t> path_flags;   // 8 per path
    std::vector<float> cuts;           // 4 per curve
    std::vector<uint8_t> ncuts;        // per curve
    std::vector<uint32_t> frag_count;  // per curve
    std::vector<int32_t> ts;           // 2 per fragment
    std::vector<int32_t> yx, pid;      // per fragment (ts order)
    std::vector<int8_t> wn;
    std::vector<uint32_t> mask;        // 4 per fragment
    std::vector<int32_t> sorted_yx, sorted_idx;
    std::vector<int8_t> wn_scan;       // sorted & scanned
    std::vector<uint32_t> merged_mask; // 4 per merged fragment
    std::vector<int32_t> draw;         // 4 per item
    std::vector<uint8_t> image;        // W*H*4
    std::vector<int64_t> counts;       // P C V F M S
};

// k_transformVertex (ras_cut_mask_comb_scanline.cu:36-105)
void transform_vertices(const Scene &s, const float *m, int W, int H, Result &r) {
    float rows[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) rows[i][j] = m[j * 4 + i];
    const float w = (float)W, h = (float)H;
    const float z = (float)std::min(W, H);
    const float cutoff = -z * 0.5f;
    r.tvx.assign(2 * (size_t)s.V, 0.f);
    r.path_flags.assign(8 * (size_t)s.P, 0);
    uint8_t *flags = r.path_flags.data();
#pragma omp parallel for schedule(static)
    for (int64_t ti = 0; ti < (int64_t)s.V; ++ti) {
        float ix = s.vx[ti], iy = s.vy[ti], iz = 0.f, iw = 1.f;
        float o[4];
        for (int k = 0; k < 4; ++k)  // f4dot (ras_device_func.h:39-41)
            o[k] = rows[k][0] * ix + rows[k][1] * iy + rows[k][2] * iz + rows[k][3] * iw;
        o[0] /= o[3];
        o[1] /= o[3];
        o[2] /= o[3];
        if (o[2] < cutoff) o[2] = cutoff;
        float w2 = w * .5f, h2 = h * .5f;
        o[0] = (o[0] - w2) * z / (o[2] + z) + w2;
        o[1] = (o[1] - h2) * z / (o[2] + z) + h2;
        int xf = o[0] < 0 ? 0 : (o[0] <= w ? 1 : 2);
        int yf = o[1] < 0 ? 0 : (o[1] <= h ? 1 : 2);
        uint8_t *pf = flags + (size_t)s.vpath[ti] * 8;
        int slot0 = -1, slot1 = -1;
        switch ((yf << 4) | xf) {
        case 0x00: slot0 = 0; break;
        case 0x01: slot0 = 1; break;
        case 0x02: slot0 = 2; break;
        case 0x10: slot0 = 3; break;
        case 0x11: slot0 = 0; slot1 = 7; break;
        case 0x12: slot0 = 4; break;
        case 0x20: slot0 = 5; break;
        case 0x21: slot0 = 6; break;
        case 0x22: slot0 = 7; break;
        default: break;
        }
        if (slot0 >= 0) {
#pragma omp atomic write
            pf[slot0] = 1;
        }
        if (slot1 >= 0) {
#pragma omp atomic write
            pf[slot1] = 1;
        }
        if (std::floor(o[0]) == o[0]) o[0] = o[0] - 1.f / 256.f;
        r.tvx[2 * ti] = o[0];
        r.tvx[2 * ti + 1] = o[1];
    }
}

// PATH_VISIBLE (ras_define.h:119-126), flags read as a little-endian uint64.
inline bool path_visible(const uint8_t *f) {
    uint64_t m = 0;
    for (int k = 0; k < 8; ++k) m |= (uint64_t)f[k] << (8 * k);
    bool inv = !(m & 0x0101010101000000ull) || !(m & 0x0001010001000101ull) ||
               !(m & 0x0000000101010101ull) || !(m & 0x0101000100010100ull);
    return !inv;
}

struct CurveData {
    int type;
    float w;
    float cx[4], cy[4];
};

inline void load_curve(const Scene &s, const Result &r, uint32_t c, CurveData &cd) {
    cd.type = s.ctype[c];
    cd.w = (cd.type == CT_RATIONAL) ? s.cw[c] : 0.f;
    uint32_t p0 = s.cvpos[c];
    int nv = cd.type & 7;
    for (int j = 0; j < 4; ++j) {
        if (j < nv) {
            cd.cx[j] = r.tvx[2 * (size_t)(p0 + j)];
            cd.cy[j] = r.tvx[2 * (size_t)(p0 + j) + 1];
        } else {
            cd.cx[j] = cd.cy[j] = 0.f;
        }
    }
}

// monotonize + insertion sort (k_make_intersection_0, ras_cut.h:1314-1460)
int monotone_cuts(const CurveData &cd, float q[4]) {
    float tq[4] = {0.f, 0.f, 0.f, 0.f};
    int n = 0;
    for (int c = 0; c < 2; ++c) {
        const float *v = c ? cd.cy : cd.cx;
        switch (cd.type) {
        case CT_LINEAR:
            break;
        case CT_QUAD: {
            float b = (v[0] - v[1]) + (v[2] - v[1]);
            if (std::fabs(b) < 1e-6f) {
            } else {
                float t = (v[0] - v[1]) / b;
                if (t > 0.f && t < 1.f) tq[n++] = t;
            }
            break;
        }
        case CT_CUBIC: {
            float r0 = 0.f, r1 = 0.f;
            solve_quad((v[3] - v[0]) + 3.f * (v[1] - v[2]), 2.f * ((v[0] - v[1]) + (v[2] - v[1])),
                       v[1] - v[0], &r0, &r1);
            if (r0 > 0.f && r0 < 1.f) tq[n++] = r0;
            if (r1 > 0.f && r1 < 1.f && r1 != r0) tq[n++] = r1;
            break;
        }
        case CT_RATIONAL: {
            float w0 = 1.f, w1 = cd.w, w2 = 1.f;
            float p0 = w0 * v[0];
            float p1 = w1 * v[1];
            float p2 = w2 * v[2];
            float pw01 = p0 * w1 - p1 * w0;
            float pw12 = p1 * w2 - p2 * w1;
            float pw20 = p2 * w0 - p0 * w2;
            float r0 = 0.f, r1 = 0.f;
            solve_quad(pw01 + pw12 + pw20, -(2.f * pw01 + pw20), pw01, &r0, &r1);
            if (r0 > 0.f && r0 < 1.f) tq[n++] = r0;
            if (r1 > 0.f && r1 < 1.f && r1 != r0) tq[n++] = r1;
            break;
        }
        default:
            break;
        }
    }
    float q0 = tq[0], q1 = tq[1], q2 = tq[2], q3 = tq[3];
    if (n >= 2) {
        float t1 = q1, t0 = q0;
        if (t1 < t0) { q1 = t0; q0 = t1; }
    }
    if (n >= 3) {
        float t2 = q2, t1 = q1;
        if (t2 < t1) {
            q2 = t1;
            float t0 = q0;
            if (t2 < t0) { q1 = t0; q0 = t2; } else { q1 = t2; }
        }
    }
    if (n >= 4) {
        float t3 = q3, t2 = q2;
        if (t3 < t2) {
            q3 = t2;
            float t1 = q1;
            if (t3 < t1) {
                q2 = t1;
                float t0 = q0;
                if (t3 < t0) { q1 = t0; q0 = t3; } else { q1 = t3; }
            } else {
                q2 = t3;
            }
        }
    }
    q[0] = q0; q[1] = q1; q[2] = q2; q[3] = q3;
    return n;
}

inline int32_t floor2(float x) { return f2i_rd(x) & (int32_t)0xFFFFFFFE; }  // d_float2int_rd<2>

// d_get_xy_begin_end_delta (ras_cut.h:693-717)
inline void xy_begin_end(float v0x, float v0y, float v1x, float v1y, int32_t &xb, int32_t &xe,
                         int32_t &yb, int32_t &ye, float &dx, float &dy) {
    if (v0x <= v1x) { xb = addw(floor2(v0x), 2); xe = floor2(v1x); dx = 2.f; }
    else { xb = addw(floor2(v1x), 2); xe = floor2(v0x); dx = -2.f; }
    if (v0y <= v1y) { yb = addw(floor2(v0y), 2); ye = floor2(v1y); dy = 2.f; }
    else { yb = addw(floor2(v1y), 2); ye = floor2(v0y); dy = -2.f; }
}

// d_udpate_cut_range (ras_cut.h:720-777)
inline void update_cut_range(int W, int H, int32_t &xb, int32_t &xe, int32_t &yb, int32_t &ye,
                             int32_t &nx, int32_t &ny) {
    const int32_t xmax = (W & ~1) + 2, ymax = (H & ~1) + 2;
    if ((xb < 0 && xe < 0) || (xb > xmax && xe > xmax) || (xb > xe)) {
        nx = 0;
    } else {
        xb = tmin(tmax(xb, 0), xmax);
        xe = tmin(tmax(xe, 0), xmax);
        nx = tmax((xe - xb) / 2 + 1, 0);
    }
    if ((yb < 0 && ye < 0) || (yb > ymax && ye > ymax) || (yb > ye)) {
        ny = 0;
    } else {
        yb = tmin(tmax(yb, 0), ymax);
        ye = tmin(tmax(ye, 0), ymax);
        ny = tmax((ye - yb) / 2 + 1, 0);
    }
}

// k_make_intersection_impl_0 (ras_cut.h:780-834): fragments of one curve.
uint32_t count_fragments(const CurveData &cd, const float *tq, int n_cuts, int W, int H) {
    float v0x = cd.cx[0], v0y = cd.cy[0];
    uint32_t p = 0;
    for (int ms = 0; ms < n_cuts; ++ms) {
        float t1 = tq[ms];
        float v1x = curve_eval(cd.type, cd.w, cd.cx, t1);
        float v1y = curve_eval(cd.type, cd.w, cd.cy, t1);
        int32_t xb, xe, yb, ye, nx = 0, ny = 0;
        float dx, dy;
        xy_begin_end(v0x, v0y, v1x, v1y, xb, xe, yb, ye, dx, dy);
        update_cut_range(W, H, xb, xe, yb, ye, nx, ny);
        p += 1 + nx + ny;
        v0x = v1x; v0y = v1y;
    }
    return p;
}

// k_make_intersection_impl_1 (ras_cut.h:837-1241): emit {curve id, t | flags}.
void intersect_curve(const CurveData &cd, int32_t cid, const float *tq, int n_cuts, int W, int H,
                     int32_t *ts /* 2 ints per entry, starting at this curve's offset */) {
    float t0_ms = 0.f;
    float v0x = cd.cx[0], v0y = cd.cy[0];
    int64_t p = 0;
    for (int ms = 0; ms < n_cuts; ++ms) {
        float t1_ms = tq[ms];
        float v1x = curve_eval(cd.type, cd.w, cd.cx, t1_ms);
        float v1y = curve_eval(cd.type, cd.w, cd.cy, t1_ms);
        t1_ms = bitsf(fbits(t1_ms) & 0xFFFFFFFCu);
        if (std::floor(v1x) == v1x) t1_ms = bitsf((fbits(t1_ms) & 0xFFFFFFFCu) | 2u);
        else t1_ms = bitsf(fbits(t1_ms) | 3u);

        int32_t xb, xe, yb, ye, nx = 0, ny = 0;
        float dx, dy;
        xy_begin_end(v0x, v0y, v1x, v1y, xb, xe, yb, ye, dx, dy);
        update_cut_range(W, H, xb, xe, yb, ye, nx, ny);
        int n_loop = nx + ny + 1;
        float x = (float)(dx < 0.f ? xe : xb);
        float y = (float)(dy < 0.f ? ye : yb);
        float slot[2] = {t0_ms, t0_ms};  // POINT_COORDS(8), POINT_COORDS(9)
        uint32_t last = fbits(-1.f);
        for (int i = -1; i < n_loop; ++i) {
            float t_solve = 0.f, c = 0.f;
            int side = 0;
            float next_tx = slot[0], next_ty = slot[1];
            float t_min;
            if (i == -1) {
                t_min = next_tx;
                if (nx == 0) { side = 0; t_solve = 2.f; }
                else if (ny == 0) { side = 1; t_solve = 2.f; }
                else { side = 0; nx--; c = x; x += dx; }
            } else if (next_tx <= next_ty) {
                t_min = next_tx;
                side = 0;
                if (nx > 0) { nx--; c = x; x += dx; } else { t_solve = 2.f; }
            } else {
                t_min = next_ty;
                side = 1;
                if (ny > 0) { ny--; c = y; y += dy; } else { t_solve = 2.f; }
            }
            if (i >= 0) {
                uint32_t out = fbits(t_min);
                if ((out & 0xFFFFFFFCu) == (last & 0xFFFFFFFCu)) {
                    out |= last;
                    ts[2 * (p - 1) + 1] = (int32_t)out;
                }
                ts[2 * p] = cid;
                ts[2 * p + 1] = (int32_t)out;
                last = out;
                p++;
            }
            if (t_solve < 2.f) {
                const float *v = side ? cd.cy : cd.cx;
                if (cd.type == CT_LINEAR) {
                    t_solve = tmin(tmax((c - v[0]) * safeRcp(v[1] - v[0]), t_min), t1_ms);
                } else if (cd.type == CT_QUAD) {
                    t_solve = solve_quad_bezier_qm(v[0], v[1], v[2], c, t_min, t1_ms);
                } else if (cd.type == CT_RATIONAL) {
                    t_solve = solve_quad_bezier_qm(v[0] - c, (v[1] - c) * cd.w, v[2] - c, 0.f, t_min, t1_ms);
                } else {
                    // 24-step bisection (ras_cut.h:1065-1106)
                    float a0 = t_min, a1 = t1_ms;
                    float vt0 = curve_eval(CT_CUBIC, 0.f, v, a0);
                    t_solve = a0;
                    if (vt0 != c) {
                        float raw_t0 = a0;
                        float last_vtm = 0.f;
                        for (int it = 0; it < 24; ++it) {
                            float tm = (a0 + a1) * .5f;
                            float vtm = curve_eval(CT_CUBIC, 0.f, v, tm);
                            t_solve = tm;
                            last_vtm = vtm;
                            if ((int32_t)(fbits(vtm - c) ^ fbits(vt0 - c)) >= 0) { a0 = tm; vt0 = vtm; }
                            else { a1 = tm; }
                        }
                        if (std::fabs(last_vtm - c) > 1.f) t_solve = raw_t0;
                    }
                }
            }
            slot[side] = bitsf((fbits(t_solve) & 0xFFFFFFFCu) | (uint32_t)side);
        }
        t0_ms = t1_ms;
        v0x = v1x; v0y = v1y;
    }
}

// ---------------------------------------------------------------------------------
// fragment + mask generation (k_gen_fragment_and_stencil_mask, .cu:112-441)
// ---------------------------------------------------------------------------------
void gen_fragment(const Scene &s, const Result &r, const Tables &T, int64_t ti, int64_t F, int W,
                  int H, int32_t &o_yx, int32_t &o_pid, int8_t &o_wn, uint32_t *o_mask) {
    int32_t c0 = r.ts[2 * ti], i0 = r.ts[2 * ti + 1];
    int32_t c1, i1;
    if (ti + 1 != F) { c1 = r.ts[2 * ti + 2]; i1 = r.ts[2 * ti + 3]; }
    else { c1 = -1; i1 = 0x3f800000; }
    int f0 = i0 & 1, f1 = i1 & 1;
    float t0 = bitsf((uint32_t)i0 & 0xFFFFFFFCu);
    float t1 = bitsf((uint32_t)i1 & 0xFFFFFFFCu);
    if (t0 < 0) t0 = 0.f;
    if (t1 < 0) t1 = 0.f;
    if (t0 == 0.f || t0 == 1.f) f0 = 1;
    if (t1 == 0.f || t1 == 1.f) f1 = 1;
    const int32_t pathid = (int32_t)s.cpath[c0];
    int32_t yx = (int32_t)0xFFFEFFFE;
    if (c0 != c1) t1 = 1.f;
    int wnc = 0, swn = 0;
    if (t0 < t1) {
        CurveData cd;
        load_curve(s, r, (uint32_t)c0, cd);
        const int rev = s.crev[c0];
        float px[2], py[2];
        // FRAG_MASK_SUBDIV = 1: t = (1-fi)*t0 + fi*t1 for fi = 0, 1
        float ta = (1.f - 0.f) * t0 + 0.f * t1;
        float tb = (1.f - 1.f) * t0 + 1.f * t1;
        px[0] = curve_eval(cd.type, cd.w, cd.cx, ta);
        py[0] = curve_eval(cd.type, cd.w, cd.cy, ta);
        px[1] = curve_eval(cd.type, cd.w, cd.cx, tb);
        py[1] = curve_eval(cd.type, cd.w, cd.cy, tb);
        float vfx, vfy, vlx, vly;
        int sf, sl;
        if (rev) { vfx = px[1]; vfy = py[1]; vlx = px[0]; vly = py[0]; sf = f1; sl = f0; }
        else { vfx = px[0]; vfy = py[0]; vlx = px[1]; vly = py[1]; sf = f0; sl = f1; }

        int32_t rfx = floor2((vfx + vlx) * 0.5f);
        int32_t rfy = floor2((vfy + vly) * 0.5f);
        int32_t pos_x = tmin(tmax(rfx, -2), (int32_t)((W & 0xFFFFFFFE) + 2));
        int32_t pos_y = tmin(tmax(rfy, -2), (int32_t)((H & 0xFFFFFFFE) + 2));
        int32_t y_shift = (int32_t)((uint32_t)(pos_y + 0x7FFF) << 16);
        int32_t x_shift = (pos_x + 0x7FFF) & 0xFFFF;
        if ((uint32_t)rfy < (uint32_t)H) yx = y_shift | x_shift;

        const float wn_y = (float)(pos_y + 1);
        if (vfy == vly) swn = 0;
        else if (vfy <= wn_y && wn_y < vly) swn = -1;
        else if (vly <= wn_y && wn_y < vfy) swn = 1;

        if (vfy < vly) wnc = -1;
        else if (vfy > vly) wnc = 1;
        else wnc = (vfx < vlx) ? 1 : -1;

        float lx = 0, ly = 0, rx = 0, ry = 0;
        int left_flag = 1;
        (void)lx; (void)rx;
        if (vfx < vlx) { lx = vfx; ly = vfy; rx = vlx; ry = vly; left_flag = sf; }
        else if (vfx > vlx) { lx = vlx; ly = vly; rx = vfx; ry = vfy; left_flag = sl; }
        else {
            if (vfx < (float)(pos_x + 1)) {
                if (sf == 0) { lx = vfx; ly = vfy; rx = vlx; ry = vly; left_flag = 0; }
                else if (sl == 0) { lx = vlx; ly = vly; rx = vfx; ry = vfy; left_flag = 0; }
            }
        }
        uint32_t lm[4] = {0, 0, 0, 0};
        if (left_flag == 0) {
            vertical_mask(T.n, (ly - (float)pos_y) * .5f, lm);
            if (ly <= (float)(pos_y + 1)) {
                if (ly >= ry) wnc *= -1;
            } else {
                if (ly < ry) wnc *= -1;
                for (int k = 0; k < 4; ++k) lm[k] = ~lm[k];
            }
        }
        const float bx = (float)pos_x, by = (float)pos_y;
        for (int k = 0; k < 2; ++k) {
            px[k] = (px[k] - bx) * .5f;
            py[k] = (py[k] - by) * .5f;
        }
        uint32_t vm[4], tm[4];
        vertical_mask2(T.n, py[0], py[1], vm);
        fetch_pixel(T, px[0], py[0], px[1], py[1], tm);
        for (int k = 0; k < 4; ++k) o_mask[k] = (tm[k] & vm[k]) ^ lm[k];
    } else {
        for (int k = 0; k < 4; ++k) o_mask[k] = 0;  // never read (sign 0, invalid key)
    }
    o_yx = yx;
    o_pid = (int32_t)((uint32_t)pathid | ((uint32_t)wnc << 30));
    o_wn = (int8_t)swn;
}

inline int16_t key_x(int32_t k) { return (int16_t)(int32_t)((int16_t)(uint16_t)(k & 0xFFFF) - 0x7FFF); }
inline int16_t key_y(int32_t k) {
    return (int16_t)(int32_t)((int16_t)(uint16_t)((uint32_t)k >> 16) - 0x7FFF);
}

// Parallel-friendly inclusive scan of uint32 flags (same result as a serial scan).
void inclusive_scan_u32(const uint32_t *in, uint32_t *out, int64_t n) {
    uint32_t acc = 0;
    for (int64_t i = 0; i < n; ++i) { acc += in[i]; out[i] = acc; }
}

// ---------------------------------------------------------------------------------
// gradient paint (output_32.vert.glsl:55-97, output_32.frag.glsl:45-51, ms_output_*.glsl)
// ---------------------------------------------------------------------------------
// The driver's inverse projection (updateRasterizerTransform, vg_main.cpp:117-143): with R
// the inverse of the input transform, rp = R (W/2, H/2, -min(W, H), 1), rx / ry = R's first
// two columns, rw = R's last column - rp, a = -rp.z.  In double (Gauss-Jordan, partial
// pivoting), rounded to float once; a singular transform gives zeros.
void inverse_projection(const float *m, int W, int H, float ctx[13]) {
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            a[r][c] = (double)m[4 * c + r];
            a[r][4 + c] = (r == c) ? 1.0 : 0.0;
        }
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        for (int r = col + 1; r < 4; ++r)
            if (std::fabs(a[r][col]) > std::fabs(a[piv][col])) piv = r;
        if (a[piv][col] == 0.0) {
            for (int i = 0; i < 13; ++i) ctx[i] = 0.f;
            return;
        }
        if (piv != col)
            for (int c = 0; c < 8; ++c) std::swap(a[col][c], a[piv][c]);
        const double d = a[col][col];
        for (int c = 0; c < 8; ++c) a[col][c] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == col) continue;
            const double f = a[r][col];
            if (f != 0.0)
                for (int c = 0; c < 8; ++c) a[r][c] -= f * a[col][c];
        }
    }
    auto mul = [&](const double v[4], double o[4]) {
        for (int r = 0; r < 4; ++r) o[r] = a[r][4] * v[0] + a[r][5] * v[1] + a[r][6] * v[2] + a[r][7] * v[3];
    };
    const double w = (double)W, h = (double)H;
    const double vp[4] = {w / 2, h / 2, -std::min(w, h), 1.0}, ex[4] = {1, 0, 0, 0}, ey[4] = {0, 1, 0, 0},
                 ew[4] = {0, 0, 0, 1};
    double rp[4], rx[4], ry[4], rw[4];
    mul(vp, rp);
    mul(ex, rx);
    mul(ey, ry);
    mul(ew, rw);
    for (int i = 0; i < 3; ++i) {
        ctx[i] = (float)rx[i];
        ctx[3 + i] = (float)ry[i];
        ctx[6 + i] = (float)(rw[i] - rp[i]);
        ctx[9 + i] = (float)rp[i];
    }
    ctx[12] = (float)(0.0 - rp[2]);
}

// A fill word without alpha names gradient slot c - 1; it paints when the slot's ramp scale
// (word 7) is positive, as the shaders' gradient branch requires.
inline bool gradient_paints(const Scene &s, uint32_t c) {
    const uint32_t gi = c - 1u;
    return c != 0u && gi < s.gslots && bitsf((uint32_t)s.table[12 * (size_t)gi + 7]) > 0.f;
}

// Gradient coordinates of window point (sx, sy): obj = rp + a / rd.z * rd with
// rd = rx sx + ry sy + rw (output_32.vert.glsl:70-76), then the slot's affine map (:79-89).
inline void gradient_coord(const float *k, const float *w, float sx, float sy, float &gx, float &gy) {
    const float rdx = (k[0] * sx + k[3] * sy) + k[6];
    const float rdy = (k[1] * sx + k[4] * sy) + k[7];
    const float rdz = (k[2] * sx + k[5] * sy) + k[8];
    const float q = k[12] / rdz;
    const float ox = k[9] + q * rdx, oy = k[10] + q * rdy;
    gx = (ox * w[0] + oy * w[1]) + w[2];
    gy = (ox * w[3] + oy * w[4]) + w[5];
}

// Paint of gradient word c at pixel (x, y) of the item covering columns [x0, x0 + wd) of y's
// cell row: the line ends at x0 + xoff and x0 + wd + xoff (xoff 1 in coverage output, 0 in
// multisample output, output_32.vert.glsl:68 / ms_output_32.vert.glsl:69), interpolated to
// the pixel centre; ramp parameter c = min(|d|^2 / (sqrt(B^2 + |d|^2 f.z) - B), 1) with
// d = g - f, B = d . f (safeRcpP: 0 unless the divisor exceeds 1e-6); the ramp texture
// (1024 wide, `rows` high) read with linear filtering and clamped edges, texels decoded by
// dc (colour: sRGB or UNORM) and da (alpha); no ramp texture reads (0, 0, 0, 1).
void gradient_paint(const Scene &s, const float *ctx, float xoff, uint32_t c, int x0, int wd, int x, int y,
                    const float *dc, const float *da, float out[4]) {
    const uint32_t gi = c - 1u;
    float w[12];
    for (int k = 0; k < 12; ++k) w[k] = bitsf((uint32_t)s.table[12 * (size_t)gi + k]);
    const float sy = (float)(y & ~1) + ((y & 1) ? 1.5f : 0.5f);
    float ax, ay, bx, by;
    gradient_coord(ctx, w, (float)x0 + xoff, sy, ax, ay);
    gradient_coord(ctx, w, (float)(x0 + wd) + xoff, sy, bx, by);
    const float t = ((float)(x - x0) + 0.5f) / (float)wd;
    const float omt = 1.f - t;
    const float gx = ax * omt + bx * t, gy = ay * omt + by * t;
    const float dx = gx - w[8], dy = gy - w[9];
    const float A = dx * dx + dy * dy, B = dx * w[8] + dy * w[9];
    const float den = std::sqrt(B * B + A * w[10]) - B;
    float cc = A * (den > 1e-6f ? 1.f / den : 0.f);
    cc = cc < 1.f ? cc : 1.f;
    if (s.rows == 0) {
        out[0] = out[1] = out[2] = 0.f;
        out[3] = 1.f;
        return;
    }
    const int32_t ri = s.table[12 * (size_t)gi + 6];
    const float irh = 1.f / (float)s.rows;
    const float u = ((float)(ri & 1023) + 0.5f) * (1.f / 1024.f) + cc * w[7];
    const float v = ((float)(ri >> 10) + 0.5f) * irh;
    const float tu = u * 1024.f - 0.5f, tv = v * (float)s.rows - 0.5f;
    const float fu = std::floor(tu), fv = std::floor(tv);
    const float au = tu - fu, av = tv - fv, bu = 1.f - au, bv = 1.f - av;
    const int iu = f2i_rz(fu), jv = f2i_rz(fv);
    const int rmax = (int)s.rows - 1;
    const int i0 = std::min(std::max(iu, 0), 1023), i1 = std::min(std::max(iu + 1, 0), 1023);
    const int j0 = std::min(std::max(jv, 0), rmax), j1 = std::min(std::max(jv + 1, 0), rmax);
    auto tex = [&](int i, int j, int ch) {
        const uint32_t v8 = ((uint32_t)s.ramp[(size_t)j * 1024 + i] >> (8 * ch)) & 0xFFu;
        return ch < 3 ? dc[v8] : da[v8];
    };
    for (int ch = 0; ch < 4; ++ch)
        out[ch] = (tex(i0, j0, ch) * bu + tex(i1, j0, ch) * au) * bv + (tex(i0, j1, ch) * bu + tex(i1, j1, ch) * au) * av;
}

// ---------------------------------------------------------------------------------
// the whole pipeline
// ---------------------------------------------------------------------------------
void run_pipeline(const Scene &s, const Tables &T, const float *m, int W, int H, int srgb, int ms,
                  bool keep, Result &r) {
    const int N = T.n;
    r.W = W; r.H = H; r.N = N; r.srgb = srgb; r.ms = ms;
    auto t_start = std::chrono::steady_clock::now();

    transform_vertices(s, m, W, H, r);

    // monotone split + fragment counting (k_make_intersection_0, ras_cut.h:1254-1482)
    const int64_t C = s.C;
    r.cuts.assign(4 * (size_t)C, 0.f);
    r.ncuts.assign(C, 0);
    r.frag_count.assign(C, 0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t c = 0; c < C; ++c) {
        CurveData cd;
        load_curve(s, r, (uint32_t)c, cd);
        bool vis = path_visible(&r.path_flags[(size_t)s.cpath[c] * 8]);
        float q[4] = {0, 0, 0, 0};
        int n = 0;
        if (vis) n = monotone_cuts(cd, q);
        for (int k = 0; k < 4; ++k) r.cuts[4 * c + k] = q[k];
        r.ncuts[c] = (uint8_t)n;
        float tq[5] = {q[0], q[1], q[2], q[3], 0.f};
        if (vis) { tq[n] = 1.f; n++; }
        r.frag_count[c] = count_fragments(cd, tq, n, W, H);
    }
    // escan (ras_scan.h:12-36)
    std::vector<uint64_t> off(C + 1, 0);
    for (int64_t c = 0; c < C; ++c) off[c + 1] = off[c] + r.frag_count[c];
    const int64_t F = (int64_t)off[C];
    r.F = F;

    r.image.assign((size_t)W * H * 4, 255);  // white clear (.cu:1396-1404)
    if (F == 0) {
        r.M = r.S = 0;
        r.counts = {(int64_t)s.P, (int64_t)s.C, (int64_t)s.V, 0, 0, 0};
        r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        return;
    }

    // intersections (k_make_intersection_1, ras_cut.h:1484-1562)
    r.ts.assign(2 * (size_t)F, 0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t c = 0; c < C; ++c) {
        if (r.frag_count[c] == 0) continue;
        CurveData cd;
        load_curve(s, r, (uint32_t)c, cd);
        bool vis = path_visible(&r.path_flags[(size_t)s.cpath[c] * 8]);
        int n = r.ncuts[c];
        float tq[5] = {r.cuts[4 * c], r.cuts[4 * c + 1], r.cuts[4 * c + 2], r.cuts[4 * c + 3], 0.f};
        if (vis) { tq[n] = 1.f; n++; }
        intersect_curve(cd, (int32_t)c, tq, n, W, H, &r.ts[2 * off[c]]);
    }

    // fragments (.cu:112-441)
    r.yx.assign(F, 0);
    r.pid.assign(F, 0);
    r.wn.assign(F, 0);
    r.mask.assign(4 * (size_t)F, 0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < F; ++i)
        gen_fragment(s, r, T, i, F, W, H, r.yx[i], r.pid[i], r.wn[i], &r.mask[4 * i]);

    // segmented stable sort by signed key within path segments (cuda_sort.cu:12-14 ->
    // SegSortPairsFromIndices, segments from .cu:424-439)
    std::vector<int64_t> seg(s.P + 1, F);
    {
        // path p's fragments are contiguous; seg[p] = first index with path >= p
        int64_t i = 0;
        for (uint32_t p = 0; p <= s.P; ++p) {
            while (i < F && (uint32_t)(r.pid[i] & 0x3FFFFFFF) < p) ++i;
            seg[p] = i;
        }
    }
    r.sorted_idx.resize(F);
    r.sorted_yx.resize(F);
    for (int64_t i = 0; i < F; ++i) r.sorted_idx[i] = (int32_t)i;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t p = 0; p < (int64_t)s.P; ++p) {
        auto b = r.sorted_idx.begin() + seg[p], e = r.sorted_idx.begin() + seg[p + 1];
        std::stable_sort(b, e, [&](int32_t a, int32_t c) { return r.yx[a] < r.yx[c]; });
    }
    std::vector<int32_t> spid(F);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < F; ++i) {
        r.sorted_yx[i] = r.yx[r.sorted_idx[i]];
        spid[i] = r.pid[r.sorted_idx[i]];
    }
    // k_shuffleFragments (.cu:444-456) + int8 exclusive scan (thrust_impl_scan.cu:16-19)
    r.wn_scan.resize(F);
    {
        int8_t acc = 0;
        for (int64_t i = 0; i < F; ++i) {
            r.wn_scan[i] = acc;
            acc = (int8_t)(acc + r.wn[r.sorted_idx[i]]);
        }
    }
    // k_mark_merged_fragment_and_span (.cu:462-535)
    std::vector<uint32_t> fflag(F), sflag(F);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < F; ++i) {
        int32_t yx1 = r.sorted_yx[i];
        int16_t x1 = key_x(yx1), y1 = key_y(yx1);
        uint32_t ff = 0, sfl = 0;
        bool oob = (x1 < 0 || y1 < 0 || x1 >= W || y1 >= H);
        if (i == 0) {
            ff = oob ? 0 : 1;
        } else {
            uint32_t pid0 = (uint32_t)spid[i - 1] & 0x3FFFFFFF;
            uint32_t pid1 = (uint32_t)spid[i] & 0x3FFFFFFF;
            uint8_t rule = s.frule[pid1];
            int32_t yx0 = r.sorted_yx[i - 1];
            int16_t x0 = key_x(yx0), y0 = key_y(yx0);
            if (oob) ff = 0;
            else if (pid0 != pid1 || yx0 != yx1) ff = 1;
            int8_t w = r.wn_scan[i];
            bool wf = (rule == 0 && w != 0) || (rule == 1 && (w & 1));
            if (y0 == y1 && (x0 + 2) < x1 && pid0 == pid1 && wf) sfl = 1;
        }
        fflag[i] = ff;
        sflag[i] = sfl;
    }
    std::vector<uint32_t> fidx(F), sidx(F);
    inclusive_scan_u32(fflag.data(), fidx.data(), F);
    inclusive_scan_u32(sflag.data(), sidx.data(), F);
    const int64_t M = fidx[F - 1], S = sidx[F - 1];
    r.M = M; r.S = S;

    // stencil resolve (k_gen_stencil_cuda_sm + k_gen_stencil_block_boundary,
    // .cu:554-786): per merged fragment and sample, w = init + sum(sign * bit) in 7-bit
    // byte lanes; the final bit is w != 0 (nonzero) or w & 1 (even-odd).
    const int nwords = words_for(N);
    r.merged_mask.assign(4 * (size_t)M, 0);
    std::vector<int64_t> starts;
    for (int64_t i = 0; i < F; ++i) {
        int32_t k = r.sorted_yx[i];
        int16_t x = key_x(k), y = key_y(k);
        bool oob = (x < 0 || y < 0 || x >= W || y >= H);
        if (k == (int32_t)0xFFFEFFFE || oob) continue;
        int64_t fi = (int64_t)fidx[i] - 1;
        if (i == 0 || fi != (int64_t)fidx[i - 1] - 1) starts.push_back(i);
    }
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t gi = 0; gi < (int64_t)starts.size(); ++gi) {
        int64_t i0 = starts[gi];
        int64_t fi = (int64_t)fidx[i0] - 1;
        int8_t init = r.wn_scan[i0];
        uint32_t init_pos = 0, init_neg = 0;
        if (init < 0) init_neg = 0x01010101u * (uint32_t)(-(int32_t)init);
        else init_pos = 0x01010101u * (uint32_t)init;
        uint32_t pos[4][8], neg[4][8];
        for (int w = 0; w < 4; ++w)
            for (int b = 0; b < 8; ++b) { pos[w][b] = init_pos; neg[w][b] = init_neg; }
        uint32_t rule = 1;
        for (int64_t i = i0; i < F; ++i) {
            int32_t k = r.sorted_yx[i];
            int16_t x = key_x(k), y = key_y(k);
            bool oob = (x < 0 || y < 0 || x >= W || y >= H);
            if (k == (int32_t)0xFFFEFFFE || oob) {
                if ((int64_t)fidx[i] - 1 != fi) break;
                continue;
            }
            if ((int64_t)fidx[i] - 1 != fi) break;
            int32_t raw = r.sorted_idx[i];
            rule = s.frule[(uint32_t)spid[i] & 0x3FFFFFFF];
            int sgn = r.pid[raw] >> 30;
            if (sgn == 0) continue;
            const uint32_t *mk = &r.mask[4 * (size_t)raw];
            for (int w = 0; w < nwords; ++w)
                for (int b = 0; b < 8; ++b) {
                    uint32_t a = (mk[w] >> b) & 0x01010101u;
                    if (sgn > 0) pos[w][b] += a; else neg[w][b] += a;
                }
        }
        for (int w = 0; w < nwords; ++w) {
            uint32_t ret = 0;
            for (int b = 0; b < 8; b += 2) {
                uint32_t b0 = (0x80808080u + pos[w][b] - neg[w][b]) & 0x7f7f7f7fu;
                uint32_t b1 = (0x80808080u + pos[w][b + 1] - neg[w][b + 1]) & 0x7f7f7f7fu;
                if (rule == 0) {
                    b0 |= b0 >> 4; b0 |= b0 >> 2; b0 |= b0 >> 1;
                    b1 |= b1 >> 4; b1 |= b1 >> 2; b1 |= b1 >> 1;
                }
                ret |= (b0 & 0x01010101u) << b;
                ret |= (b1 & 0x01010101u) << (b + 1);
            }
            r.merged_mask[4 * fi + w] = ret;
        }
    }

    // draw list (k_gen_gl_merged_fragment_and_span, .cu:789-868)
    r.draw.assign(4 * (size_t)(M + S), 0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < F; ++i) {
        int64_t ff = fflag[i], fi = fidx[i];
        int64_t sfl = sflag[i], si = sidx[i];
        int64_t nfb = fi - ff, nsb = si - sfl;
        if (ff) {
            int64_t o = nfb + nsb;
            int32_t k = r.sorted_yx[i];
            int16_t x = key_x(k), y = key_y(k);
            uint32_t p = (uint32_t)spid[i] & 0x3FFFFFFF;
            r.draw[4 * o + 0] = (int32_t)(((uint32_t)(int32_t)y << 16) | (uint32_t)(int32_t)x);
            r.draw[4 * o + 1] = 2;
            r.draw[4 * o + 2] = (int32_t)s.finfo[p];
            r.draw[4 * o + 3] = (int32_t)fi;
        }
        if (sfl) {
            int64_t o = nfb + nsb + ff;
            uint32_t p0 = (uint32_t)r.sorted_yx[i - 1], p1 = (uint32_t)r.sorted_yx[i];
            uint32_t p = (uint32_t)spid[i] & 0x3FFFFFFF;
            int16_t y0 = key_y((int32_t)p0);
            int16_t x0 = (int16_t)(int32_t)(key_x((int32_t)p0) + 2);
            int16_t x1 = key_x((int32_t)p1);
            x0 = tmax((int16_t)0, x0);
            r.draw[4 * o + 0] = (int32_t)(((uint32_t)(int32_t)y0 << 16) | (uint32_t)(int32_t)x0);
            r.draw[4 * o + 1] = (int32_t)x1 - (int32_t)x0;
            r.draw[4 * o + 2] = (int32_t)s.finfo[p];
            r.draw[4 * o + 3] = 0;
        }
    }
    int64_t pc = 0;
    for (int64_t k = 0; k < M + S; ++k) pc += 2 * (int64_t)r.draw[4 * k + 1];
    r.pixel_count = pc;

    // compositing (output_{8,32}.{vert,frag}.glsl, ms_output_*.frag.glsl; blend state
    // .cu:1359-1389; resolve ras_base.cpp:1148-1163).  Items are bucketed by 2-row band
    // (stable), then bands are blended independently in draw order.
    float dec_c[256], thr_c[256], dec_a[256], thr_a[256];
    color_tables(srgb, dec_c, thr_c);
    color_tables(0, dec_a, thr_a);
    float gctx[13] = {0};
    if (s.gslots) inverse_projection(m, W, H, gctx);
    const int nbands = (H + 1) / 2;
    std::vector<int64_t> bcount(nbands + 1, 0);
    for (int64_t k = 0; k < M + S; ++k) {
        int y = (int)(int16_t)((uint32_t)r.draw[4 * k] >> 16);
        if (y >= 0 && y < H) bcount[y / 2 + 1]++;
    }
    for (int b = 0; b < nbands; ++b) bcount[b + 1] += bcount[b];
    std::vector<int64_t> border(M + S);
    {
        std::vector<int64_t> cur(bcount.begin(), bcount.end() - 1);
        for (int64_t k = 0; k < M + S; ++k) {
            int y = (int)(int16_t)((uint32_t)r.draw[4 * k] >> 16);
            if (y >= 0 && y < H) border[cur[y / 2]++] = k;
        }
    }
    const int full = ms ? N : 0;
    (void)full;
#pragma omp parallel for schedule(dynamic, 4)
    for (int b = 0; b < nbands; ++b) {
        const int y0 = 2 * b;
        const int rows = (y0 + 1 < H) ? 2 : 1;
        // per pixel state: 4 bytes, or N*4 bytes in multisample mode
        const int spp_state = ms ? N : 1;
        std::vector<uint8_t> st((size_t)rows * W * spp_state * 4, 255);
        for (int64_t q = bcount[b]; q < bcount[b + 1]; ++q) {
            const int32_t *d = &r.draw[4 * border[q]];
            int x = (int)(int16_t)((uint32_t)d[0] & 0xFFFF);
            int wdt = d[1];
            uint32_t col = (uint32_t)d[2];
            const uint32_t *mk = d[3] ? &r.merged_mask[4 * (size_t)(d[3] - 1)] : nullptr;
            const bool grad = !(col & 0xFF000000u) && gradient_paints(s, col);
            if (!(col & 0xFF000000u) && !grad) continue;  // paints nothing
            float sa = dec_a[(col >> 24) & 0xFF];
            float sc[3] = {dec_c[col & 0xFF], dec_c[(col >> 8) & 0xFF], dec_c[(col >> 16) & 0xFF]};
            for (int dy = 0; dy < rows; ++dy) {
                for (int dx = 0; dx < wdt; ++dx) {
                    int px = x + dx;
                    if (px < 0 || px >= W) continue;
                    uint32_t bits;  // coverage bits of this pixel, sample s -> bit s
                    if (!mk || dx > 1) {
                        bits = (N == 32) ? 0xFFFFFFFFu : ((1u << N) - 1u);
                    } else if (N == 8) {
                        bits = (mk[0] >> ((dx * 2 + dy) * 8)) & 0xFFu;
                    } else if (N == 16) {
                        bits = (mk[dx] >> (dy * 16)) & 0xFFFFu;
                    } else {
                        bits = mk[dx * 2 + dy];
                    }
                    uint8_t *px4 = &st[(((size_t)dy * W + px) * spp_state) * 4];
                    if (grad) {
                        float g4[4];
                        gradient_paint(s, gctx, ms ? 0.f : 1.f, col, x, wdt, px, y0 + dy, dec_c, dec_a, g4);
                        sc[0] = g4[0]; sc[1] = g4[1]; sc[2] = g4[2];
                        sa = g4[3];
                    }
                    if (!ms) {
                        int cnt = __builtin_popcount(bits);
                        float a = sa * ((float)cnt / (float)N);
                        float oma = 1.f - a;
                        for (int ch = 0; ch < 3; ++ch) {
                            float o = sc[ch] * a + dec_c[px4[ch]] * oma;
                            px4[ch] = (uint8_t)encode8(thr_c, o);
                        }
                        float oa = a * a + dec_a[px4[3]] * oma;
                        px4[3] = (uint8_t)encode8(thr_a, oa);
                    } else {
                        float a = sa, oma = 1.f - a;
                        for (int smp = 0; smp < N; ++smp) {
                            if (!((bits >> smp) & 1u)) continue;
                            uint8_t *p4 = px4 + 4 * smp;
                            for (int ch = 0; ch < 3; ++ch) {
                                float o = sc[ch] * a + dec_c[p4[ch]] * oma;
                                p4[ch] = (uint8_t)encode8(thr_c, o);
                            }
                            float oa = a * a + dec_a[p4[3]] * oma;
                            p4[3] = (uint8_t)encode8(thr_a, oa);
                        }
                    }
                }
            }
        }
        // resolve into the image
        for (int dy = 0; dy < rows; ++dy)
            for (int px = 0; px < W; ++px) {
                uint8_t *o = &r.image[(((size_t)(y0 + dy)) * W + px) * 4];
                const uint8_t *p4 = &st[(((size_t)dy * W + px) * spp_state) * 4];
                if (!ms) {
                    for (int ch = 0; ch < 4; ++ch) o[ch] = p4[ch];
                } else {
                    for (int ch = 0; ch < 4; ++ch) {
                        const float *dec = ch < 3 ? dec_c : dec_a;
                        const float *thr = ch < 3 ? thr_c : thr_a;
                        float sum = 0.f;
                        for (int smp = 0; smp < N; ++smp) sum += dec[p4[4 * smp + ch]];
                        o[ch] = (uint8_t)encode8(thr, sum / (float)N);
                    }
                }
            }
    }
    r.counts = {(int64_t)s.P, (int64_t)s.C, (int64_t)s.V, F, M, S};
    r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    if (!keep) {
        std::vector<int32_t>().swap(r.ts);
        std::vector<int32_t>().swap(r.yx);
        std::vector<int32_t>().swap(r.pid);
        std::vector<int8_t>().swap(r.wn);
        std::vector<uint32_t>().swap(r.mask);
        std::vector<int32_t>().swap(r.sorted_idx);
        std::vector<int32_t>().swap(r.sorted_yx);
        std::vector<int8_t>().swap(r.wn_scan);
    }
}

// unifyCurveVertexOrder (ras_base.cpp:523-569)
void unify_order(Scene &s) {
    s.crev.assign(s.C, 0);
    for (uint32_t i = 0; i < s.C; ++i) {
        int vn = s.ctype[i] & 7;
        uint32_t p = s.cvpos[i];
        uint32_t ixf = fbits(s.vx[p]), ixl = fbits(s.vx[p + vn - 1]);
        uint32_t iyf = fbits(s.vy[p]), iyl = fbits(s.vy[p + vn - 1]);
        if (iyf < iyl || (iyf == iyl && ixf < ixl)) {
            s.crev[i] = 0;
        } else {
            s.crev[i] = 1;
            for (int k = 0; k < vn / 2; ++k) {
                std::swap(s.vx[p + k], s.vx[p + vn - k - 1]);
                std::swap(s.vy[p + k], s.vy[p + vn - k - 1]);
            }
        }
    }
}

}  // namespace

struct oracle_result {
    Result r;
};

extern "C" {

oracle_result *oracle_run(uint32_t n_paths, uint32_t n_curves, uint32_t n_vertices,
                          const float *vertex, const uint32_t *curve_vertex_pos,
                          const uint8_t *curve_type, const uint32_t *curve_path_id,
                          const float *curve_arc_w, const uint8_t *fill_rule,
                          const uint32_t *fill_color, int width, int height, int samples,
                          int srgb, int ms_output, const float *m16, int flags, const int32_t *ramp,
                          uint32_t n_ramp, const int32_t *table, uint32_t n_table) {
    if (samples != 8 && samples != 16 && samples != 32) return nullptr;
    if ((n_ramp && !ramp) || (n_table && !table)) return nullptr;
    if (width <= 0 || height <= 0 || width > 32766 || height > 32766) return nullptr;
    Scene s;
    s.P = n_paths; s.C = n_curves; s.V = n_vertices;
    s.vx.resize(s.V); s.vy.resize(s.V);
    for (uint32_t i = 0; i < s.V; ++i) { s.vx[i] = vertex[2 * i]; s.vy[i] = vertex[2 * i + 1]; }
    s.cvpos.assign(curve_vertex_pos, curve_vertex_pos + s.C);
    s.ctype.assign(curve_type, curve_type + s.C);
    s.cpath.assign(curve_path_id, curve_path_id + s.C);
    s.cw.assign(s.C, -1.f);
    for (uint32_t i = 0; i < s.C; ++i)
        if (s.ctype[i] == CT_RATIONAL && curve_arc_w) s.cw[i] = curve_arc_w[i];
    s.frule.assign(fill_rule, fill_rule + s.P);
    // fill words (ras_base.cpp:144-166): colours keep their value unless alpha is 0; with a
    // gradient table, a word in [1, 0x01000000] names gradient slot word - 1
    const uint32_t gslots = n_table / 12;
    bool any_grad = false;
    s.finfo.resize(s.P);
    for (uint32_t p = 0; p < s.P; ++p) {
        uint32_t c = fill_color[p];
        const bool g = !(c & 0xFF000000u) && gslots && c != 0u;
        s.finfo[p] = ((c & 0xFF000000u) || g) ? c : 0u;
        any_grad |= g;
    }
    if (any_grad) {
        s.gslots = gslots;
        s.rows = n_ramp >> 10;
        s.table.assign(table, table + (size_t)gslots * 12);
        s.ramp.assign(ramp, ramp + (size_t)s.rows * 1024);
    }
    s.vpath.assign(s.V, 0);
    for (uint32_t c = 0; c < s.C; ++c)
        for (int k = 0; k < (s.ctype[c] & 7); ++k) s.vpath[s.cvpos[c] + k] = s.cpath[c];
    unify_order(s);

    Tables T;
    T.n = samples;
    T.pixel.resize((size_t)65536 * words_for(samples));
    make_pixel_table(samples, T.pixel.data());

    float ident[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    oracle_result *o = new oracle_result;
    run_pipeline(s, T, m16 ? m16 : ident, width, height, srgb, ms_output, (flags & 1) != 0, o->r);
    return o;
}

const char *oracle_names(void) {
    return "tvertex path_flags cuts ncuts frag_count ts frag_yx frag_pid frag_wn frag_mask "
           "sorted_yx sorted_idx wn_scan merged_mask draw_list image counts pixel_count";
}

int oracle_get(const oracle_result *o, const char *name, const void **ptr, size_t *nbytes) {
    if (!o || !name || !ptr || !nbytes) return 1;
    const Result &r = o->r;
    std::string n(name);
#define ORACLE_ARR(key, vec)                         \
    if (n == key) {                                  \
        *ptr = (const void *)(vec).data();           \
        *nbytes = (vec).size() * sizeof((vec)[0]);   \
        return 0;                                    \
    }
    ORACLE_ARR("tvertex", r.tvx)
    ORACLE_ARR("path_flags", r.path_flags)
    ORACLE_ARR("cuts", r.cuts)
    ORACLE_ARR("ncuts", r.ncuts)
    ORACLE_ARR("frag_count", r.frag_count)
    ORACLE_ARR("ts", r.ts)
    ORACLE_ARR("frag_yx", r.yx)
    ORACLE_ARR("frag_pid", r.pid)
    ORACLE_ARR("frag_wn", r.wn)
    ORACLE_ARR("frag_mask", r.mask)
    ORACLE_ARR("sorted_yx", r.sorted_yx)
    ORACLE_ARR("sorted_idx", r.sorted_idx)
    ORACLE_ARR("wn_scan", r.wn_scan)
    ORACLE_ARR("merged_mask", r.merged_mask)
    ORACLE_ARR("draw_list", r.draw)
    ORACLE_ARR("image", r.image)
    ORACLE_ARR("counts", r.counts)
#undef ORACLE_ARR
    if (n == "pixel_count") {
        *ptr = &r.pixel_count;
        *nbytes = sizeof(r.pixel_count);
        return 0;
    }
    return 2;
}

double oracle_render_seconds(const oracle_result *o) { return o ? o->r.seconds : -1.0; }
void oracle_free(oracle_result *o) { delete o; }

int oracle_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int oracle_sample_positions(int n, float *xy) {
    if (n != 8 && n != 16 && n != 32) return 1;
    sample_positions(n, xy);
    return 0;
}

int oracle_pixel_table(int n, uint32_t *out) {
    if (n != 8 && n != 16 && n != 32) return 1;
    make_pixel_table(n, out);
    return 0;
}

void oracle_fetch_pixel(int n, float p0x, float p0y, float p1x, float p1y, uint32_t out[4]) {
    static thread_local Tables T;
    if (T.n != n) {
        T.n = n;
        T.pixel.assign((size_t)65536 * words_for(n), 0);
        make_pixel_table(n, T.pixel.data());
    }
    fetch_pixel(T, p0x, p0y, p1x, p1y, out);
}

void oracle_fetch_pixel_batch(int n, int64_t count, const float *pts, uint32_t *out) {
    Tables T;
    T.n = n;
    T.pixel.assign((size_t)65536 * words_for(n), 0);
    make_pixel_table(n, T.pixel.data());
    for (int64_t i = 0; i < count; ++i)
        fetch_pixel(T, pts[4 * i], pts[4 * i + 1], pts[4 * i + 2], pts[4 * i + 3], out + 4 * i);
}

void oracle_vertical_mask(int n, float y, uint32_t out[4]) { vertical_mask(n, y, out); }
void oracle_vertical_mask2(int n, float y0, float y1, uint32_t out[4]) { vertical_mask2(n, y0, y1, out); }
void oracle_solve_quad(float a, float b, float c, float *r0, float *r1) { solve_quad(a, b, c, r0, r1); }
float oracle_curve_eval(int type, float w, const float *v, float t) { return curve_eval(type, w, v, t); }
void oracle_color_tables(int srgb, float *dec, float *thr) { color_tables(srgb, dec, thr); }

}  // extern "C"
