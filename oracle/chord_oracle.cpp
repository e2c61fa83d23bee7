// chord_oracle.cpp -- CPU restatement of the reference's GPU-generated chord-diagram
// animation (TEST INFRASTRUCTURE ONLY: the checker of sl_chord_frame, never a renderer).
//
// Follows gpu-scanline/src/rasterizer/kernel/animation.cu stage by stage, as the reference
// runs it per diagram:
//   k_evaluate_node / k_evaluate_link (:1890-1986)  eased keyframe interpolation into
//                                                  per-frame node and link arrays
//   k_gen_node_vg (:1988-2162)                      ring sector of every node
//   k_gen_link_vg (:2164-2432)                      ribbon of every link
// and then the host upload's per-curve orientation (unifyCurveVertexOrder,
// ras_base.cpp:523-569) and fill-word rule (ras_base.cpp:144-166), which sl_add_vg applies
// to any scene.  Two documented departures from the literal reference: normalize() of a
// (sin t, cos t) pair is left out (the pair is a unit vector up to rounding, and the
// device code uses the pair as computed), and a node of size 0 maps its links' ends to its
// start instead of dividing by zero.  Sine and cosine are evaluated in double by
// reduction modulo pi/2 (pi/2 split into a 33-bit head and a tail) and Taylor polynomials,
// then rounded to float, so that every IEEE double implementation gives the same bits.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct V2 {
    float x, y;
};

const float kSmallR = 320.f, kLargeR = 350.f;  // SMALL_RADIUS, LARGE_RADIUS (:1722-1723)
const int kKeyMs = 6000;                       // TIME_PER_YEAR (:1708)

// sin x = sin r or cos r (with sign) for x = k pi/2 + r, |r| <= pi/4
void sincos_ref(float xf, float *s_out, float *c_out) {
    const double x = xf;
    const double head = 1.5707963267341256;    // pi/2 rounded to 33 significant bits
    const double tail = 6.077100506506192e-11; // pi/2 - head
    const double k = std::floor(x * 0.6366197723675814 + 0.5);
    const double r = (x - k * head) - k * tail;
    const double z = r * r;
    // Horner forms of the Taylor series, highest term first: inverse factorials with
    // alternating signs, sin to r^21 and cos to r^22
    static const double is[10] = {-1.9572941063391263e-20, 8.22063524662433e-18, -2.8114572543455206e-15,
                                  7.647163731819816e-13,   -1.6059043836821613e-10, 2.505210838544172e-08,
                                  -2.7557319223985893e-06, 0.0001984126984126984, -0.008333333333333333,
                                  0.16666666666666666};
    static const double ic[11] = {8.896791392450574e-22,  -4.110317623312165e-19, 1.5619206968586225e-16,
                                  -4.779477332387385e-14, 1.1470745597729725e-11, -2.08767569878681e-09,
                                  2.755731922398589e-07,  -2.48015873015873e-05,  0.001388888888888889,
                                  -0.041666666666666664,  0.5};
    double ps = is[0];
    for (int i = 1; i < 10; ++i) ps = ps * z + is[i];
    double pc = ic[0];
    for (int i = 1; i < 11; ++i) pc = pc * z + ic[i];
    const double sr = r - r * z * ps, cr = 1.0 - z * pc;
    double s, c;
    switch ((int)((long long)k & 3)) {
    case 0: s = sr; c = cr; break;
    case 1: s = cr; c = -sr; break;
    case 2: s = -sr; c = -cr; break;
    default: s = -cr; c = sr; break;
    }
    *s_out = (float)s;
    *c_out = (float)c;
}

float sin_ref(float x) {
    float s, c;
    sincos_ref(x, &s, &c);
    return s;
}

// GET_ANIMATION_T (:1710-1718) and the clamp of k_evaluate_node (:1918-1920)
float anim_t(int time) {
    float t = time < kKeyMs ? 1.f : (float)(time % kKeyMs) / ((float)kKeyMs * .66666f);
    if (t < 1.f) t = t < 0.5f ? 2 * t * t : 1 - 2 * (1 - t) * (1 - t);
    else t = 1.f;
    t = std::fmax(0.f, t);
    t = std::fmin(1.f, t);
    return t;
}

V2 add(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
V2 sub(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
V2 mul(V2 a, float s) { return {a.x * s, a.y * s}; }

// the seven points of a two-cubic arc on radius R (:2060-2104): ends, middle, and handles
// along the normals n = (-d.y, d.x)
void arc_points(V2 center, float R, float t0, float t1, V2 *p, V2 *dir) {
    float ts[3] = {t0, (float)(t0 * 0.5 + t1 * 0.5), t1};
    const int at[3] = {0, 3, 6};
    for (int i = 0; i < 3; ++i) {
        V2 d;
        sincos_ref(ts[i], &d.x, &d.y);
        dir[at[i]] = d;
        p[at[i]] = add(center, mul(d, R));
    }
    const float len = sin_ref((t1 - t0) * .5f) * .353f * R;
    const V2 n0 = {-dir[0].y, dir[0].x}, n3 = {-dir[3].y, dir[3].x}, n6 = {-dir[6].y, dir[6].x};
    p[1] = sub(p[0], mul(n0, len));
    p[2] = add(p[3], mul(n3, len));
    p[4] = sub(p[3], mul(n3, len));
    p[5] = add(p[6], mul(n6, len));
}

struct Out {
    float *vertex;       // 2 per vertex
    uint32_t *vpath;     // per vertex
    uint32_t *curve;     // 4 per curve: vertex pos, path, type | reversed << 8, weight bits
    uint8_t *fill_rule;  // per path
    uint32_t *fill;      // per path
};

void put(const Out &o, uint32_t c, uint32_t vp, uint32_t path, int type, const V2 *pts) {
    for (int k = 0; k < type; ++k) {
        o.vertex[2 * (vp + k)] = pts[k].x;
        o.vertex[2 * (vp + k) + 1] = pts[k].y;
    }
    o.curve[4 * c] = vp;
    o.curve[4 * c + 1] = path;
    o.curve[4 * c + 2] = (uint32_t)type;
    const float w = -1.f;
    std::memcpy(&o.curve[4 * c + 3], &w, 4);
}

}  // namespace

extern "C" {

void oracle_chord_sincos(float x, float *s, float *c) { sincos_ref(x, s, c); }
float oracle_chord_ease(int time_ms) { return anim_t(time_ms); }

// Keyframe pair and weight of a frame: k = time / 6000 + 1 (chord_animation, :2457-2459),
// resting on the last keyframe after it.
static void frame_keys(int n_keys, int time_ms, int *k, float *t) {
    *k = time_ms / kKeyMs + 1;
    *t = anim_t(time_ms);
    if (*k > n_keys - 1) {
        *k = n_keys - 1;
        *t = 1.f;
    }
}

// Paths of one diagram at time_ms: its nodes and the links of its interval.
int64_t oracle_chord_paths(int n_keys, int n_nodes, const uint32_t *link_offset, int time_ms) {
    int k;
    float t;
    frame_keys(n_keys, time_ms, &k, &t);
    return (int64_t)n_nodes + (link_offset[k] - link_offset[k - 1]);
}

// The scene of one diagram at time_ms, written at paths [pbase, pbase + paths): 6 curves
// and 24 vertex slots per path.  Arrays as in sl_chord (include/scanline.h).
int64_t oracle_chord_diagram(int n_keys, int n_nodes, const float *node_t0, const float *node_t1,
                             const float *node_size, const uint32_t *node_color, const uint32_t *link_offset,
                             const uint32_t *link_node, const float *link_from, const float *link_to, float cx,
                             float cy, int time_ms, uint32_t pbase, float *vertex, uint32_t *vpath, uint32_t *curve,
                             uint8_t *fill_rule, uint32_t *fill) {
    int k;
    float t;
    frame_keys(n_keys, time_ms, &k, &t);
    const V2 center = {cx, cy};
    const Out o{vertex, vpath, curve, fill_rule, fill};
    const int N = n_nodes;
    const uint32_t l0 = link_offset[k - 1], L = link_offset[k] - l0;
    // k_evaluate_node: previous keyframe (k - 1) towards the next (k)
    std::vector<float> at0(N), at1(N), asz(N);
    for (int i = 0; i < N; ++i) {
        const size_t p = (size_t)(k - 1) * N + i, n = (size_t)k * N + i;
        at0[i] = node_t0[p] * (1 - t) + node_t0[n] * t;
        at1[i] = node_t1[p] * (1 - t) + node_t1[n] * t;
        asz[i] = node_size[p] * (1 - t) + node_size[n] * t;
    }
    // k_evaluate_link
    std::vector<float> lpos0(L), lpos1(L), lsz0(L), lsz1(L);
    for (uint32_t i = 0; i < L; ++i) {
        const float *f = link_from + 4 * (size_t)(l0 + i), *g = link_to + 4 * (size_t)(l0 + i);
        lpos0[i] = f[0] * (1 - t) + g[0] * t;
        lpos1[i] = f[1] * (1 - t) + g[1] * t;
        lsz0[i] = f[2] * (1 - t) + g[2] * t;
        lsz1[i] = f[3] * (1 - t) + g[3] * t;
    }
    // k_gen_node_vg
    for (int i = 0; i < N; ++i) {
        const float t0 = at0[i], t1 = at1[i];
        V2 ip[7], op[7], dir[7];
        if (t0 == t1) {
            for (int j = 0; j < 7; ++j) ip[j] = op[j] = {-1000.f, -1000.f};
        } else {
            arc_points(center, kSmallR, t0, t1, ip, dir);
            arc_points(center, kLargeR, t0, t1, op, dir);
        }
        const uint32_t path = pbase + (uint32_t)i, c = path * 6, v = path * 24;
        fill_rule[path] = 0;
        fill[path] = node_color[i];
        for (int j = 0; j < 24; ++j) vpath[v + j] = path;
        for (int j = 20; j < 24; ++j) vertex[2 * (v + j)] = vertex[2 * (v + j) + 1] = -1000.f;
        const V2 c0[2] = {op[0], ip[0]}, c1[2] = {ip[6], op[6]};
        const V2 c2[4] = {ip[0], ip[1], ip[2], ip[3]}, c3[4] = {ip[3], ip[4], ip[5], ip[6]};
        const V2 c4[4] = {op[6], op[5], op[4], op[3]}, c5[4] = {op[3], op[2], op[1], op[0]};
        put(o, c + 0, v + 0, path, 2, c0);
        put(o, c + 1, v + 2, path, 2, c1);
        put(o, c + 2, v + 4, path, 4, c2);
        put(o, c + 3, v + 8, path, 4, c3);
        put(o, c + 4, v + 12, path, 4, c4);
        put(o, c + 5, v + 16, path, 4, c5);
    }
    // k_gen_link_vg
    const float pi = (float)std::acos(-1.0f);
    for (uint32_t i = 0; i < L; ++i) {
        const uint32_t n0 = link_node[2 * (size_t)(l0 + i)], n1 = link_node[2 * (size_t)(l0 + i) + 1];
        const float ns0 = asz[n0], ns1 = asz[n1];
        float lt00 = ns0 != 0.f ? lpos0[i] / ns0 : 0.f, lt01 = ns0 != 0.f ? (lpos0[i] + lsz0[i]) / ns0 : 0.f;
        float lt10 = ns1 != 0.f ? lpos1[i] / ns1 : 0.f, lt11 = ns1 != 0.f ? (lpos1[i] + lsz1[i]) / ns1 : 0.f;
        auto llerp = [](float a, float b, float u) { return a * (1 - u) + b * u; };
        lt00 = llerp(at0[n0], at1[n0], lt00);
        lt01 = llerp(at0[n0], at1[n0], lt01);
        lt10 = llerp(at0[n1], at1[n1], lt10);
        lt11 = llerp(at0[n1], at1[n1], lt11);
        V2 ip0[7], d0[7], ip1[7], d1[7];
        arc_points(center, kSmallR, lt00, lt01, ip0, d0);
        arc_points(center, kSmallR, lt10, lt11, ip1, d1);
        float th_s = std::fabs(lt10 - lt01), th_l = std::fabs(lt11 - lt00);
        if (th_s > pi) th_s -= pi;
        if (th_l > pi) th_l -= pi;
        th_s *= .5f;
        th_l *= .5f;
        const float r_s = kSmallR / (2 * sin_ref(th_s) + 1.f), r_l = kSmallR / (2 * sin_ref(th_l) + 1.f);
        const V2 c00 = add(center, mul(d0[0], r_l)), c01 = add(center, mul(d0[6], r_s));
        const V2 c10 = add(center, mul(d1[0], r_s)), c11 = add(center, mul(d1[6], r_l));
        const uint32_t path = pbase + (uint32_t)N + i, c = path * 6, v = path * 24;
        fill_rule[path] = 0;
        fill[path] = node_color[lsz0[i] > lsz1[i] ? n0 : n1] & 0x7FFFFFFFu;
        for (int j = 0; j < 24; ++j) vpath[v + j] = path;
        const V2 e0[4] = {ip0[6], c01, c10, ip1[0]}, e1[4] = {ip1[6], c11, c00, ip0[0]};
        const V2 e2[4] = {ip0[0], ip0[1], ip0[2], ip0[3]}, e3[4] = {ip0[3], ip0[4], ip0[5], ip0[6]};
        const V2 e4[4] = {ip1[0], ip1[1], ip1[2], ip1[3]}, e5[4] = {ip1[3], ip1[4], ip1[5], ip1[6]};
        put(o, c + 0, v + 0, path, 4, e0);
        put(o, c + 1, v + 4, path, 4, e1);
        put(o, c + 2, v + 8, path, 4, e2);
        put(o, c + 3, v + 12, path, 4, e3);
        put(o, c + 4, v + 16, path, 4, e4);
        put(o, c + 5, v + 20, path, 4, e5);
    }
    // the upload's orientation (ras_base.cpp:523-569) and fill words (ras_base.cpp:144-166)
    const uint32_t P = (uint32_t)N + L;
    for (uint32_t c = pbase * 6; c < (pbase + P) * 6; ++c) {
        const uint32_t vp = curve[4 * c];
        const int vn = (int)(curve[4 * c + 2] & 7);
        uint32_t ixf, iyf, ixl, iyl;
        std::memcpy(&ixf, &vertex[2 * vp], 4);
        std::memcpy(&iyf, &vertex[2 * vp + 1], 4);
        std::memcpy(&ixl, &vertex[2 * (vp + vn - 1)], 4);
        std::memcpy(&iyl, &vertex[2 * (vp + vn - 1) + 1], 4);
        if (!(iyf < iyl || (iyf == iyl && ixf < ixl))) {
            curve[4 * c + 2] |= 1u << 8;
            for (int j = 0; j < vn / 2; ++j) {
                float *a = &vertex[2 * (vp + j)], *b = &vertex[2 * (vp + vn - 1 - j)];
                std::swap(a[0], b[0]);
                std::swap(a[1], b[1]);
            }
        }
    }
    for (uint32_t p = pbase; p < pbase + P; ++p)
        if (!(fill[p] & 0xFF000000u)) fill[p] = 0;
    return P;
}

}  // extern "C"
