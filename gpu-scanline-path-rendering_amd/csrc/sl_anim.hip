// sl_anim.hip -- GPU-side scene generation: the chord-diagram animation that the reference
// generates on the device each frame (vg_animation / chord_animation, animation.cu:
// 2435-2786) instead of uploading a scene.  Each diagram has keyframes: per node its
// angular extent [t0, t1] and size, and per interval between two keyframes the links
// present at either end, each with its two nodes and, at both ends of the interval, the
// sizes and positions of its two ends inside their nodes (genFrame, animation.cu:
// 1200-1330).  A frame at time T (milliseconds) eases from keyframe k - 1 to k,
// k = T / 6000 + 1 (k_evaluate_node / k_evaluate_link, animation.cu:1890-1986), and
// writes straight into the renderer's scene arrays:
//   per node  a ring sector between radii 320 and 350: two radial lines and two cubic
//             arcs a side (k_gen_node_vg, animation.cu:1988-2162);
//   per link  a ribbon: an arc on each end node's inner circle and two cubic chords
//             pulled towards the centre (k_gen_link_vg, animation.cu:2164-2432).
// One thread makes one path, six curves in 24 vertex slots, and also does what the host
// upload does to every curve (sl_add_vg): the vertex order fixed by float bit patterns
// (unifyCurveVertexOrder, ras_base.cpp:523-569), the reversed flag, the packed curve word,
// the path of each vertex, and fill words of alpha 0 cleared.  The reference's year and
// country-name text is not generated (its glyphs come from FreeType fonts of that demo).
//
// Arithmetic is float with contraction off, as in the reference's kernels, except sine and
// cosine: computed in double with plain multiplies and adds (Cody-Waite reduction by pi/2
// in two parts, Taylor polynomials on [-pi/4, pi/4]) and rounded once to float, so that
// the CPU restatement in the oracle reproduces every bit.
#include "sl_internal.h"

namespace sl {
namespace {

constexpr float kInnerR = 320.f, kOuterR = 350.f;  // SMALL_RADIUS, LARGE_RADIUS (animation.cu:1722-1723)
constexpr float kPi = 3.14159274f;                 // acos(-1.f), the float nearest pi

// sin and cos of a float angle (see the file comment); |x| below 1e5 keeps k * kC1 exact.
__device__ __forceinline__ void sincos_fixed(float xf, float &s_out, float &c_out) {
    const double x = (double)xf;
    const double kC1 = 1.5707963267341256;    // pi/2 to 33 significant bits
    const double kC2 = 6.077100506506192e-11; // pi/2 - kC1, rounded
    const double k = floor(x * 0.6366197723675814 + 0.5);
    const double r = (x - k * kC1) - k * kC2;
    const double z = r * r;
    // sin r = r - r z (1/3! - z/5! + ... - z^9/21!), cos r = 1 - z (1/2! - z/4! + ... + z^10/22!)
    double sp = -1.9572941063391263e-20;
    sp = sp * z + 8.22063524662433e-18;
    sp = sp * z - 2.8114572543455206e-15;
    sp = sp * z + 7.647163731819816e-13;
    sp = sp * z - 1.6059043836821613e-10;
    sp = sp * z + 2.505210838544172e-08;
    sp = sp * z - 2.7557319223985893e-06;
    sp = sp * z + 0.0001984126984126984;
    sp = sp * z - 0.008333333333333333;
    sp = sp * z + 0.16666666666666666;
    const double sn = r - r * z * sp;
    double cp = 8.896791392450574e-22;
    cp = cp * z - 4.110317623312165e-19;
    cp = cp * z + 1.5619206968586225e-16;
    cp = cp * z - 4.779477332387385e-14;
    cp = cp * z + 1.1470745597729725e-11;
    cp = cp * z - 2.08767569878681e-09;
    cp = cp * z + 2.755731922398589e-07;
    cp = cp * z - 2.48015873015873e-05;
    cp = cp * z + 0.001388888888888889;
    cp = cp * z - 0.041666666666666664;
    cp = cp * z + 0.5;
    const double cs = 1.0 - z * cp;
    const int q = (int)((long long)k & 3);
    const double s = q == 0 ? sn : (q == 1 ? cs : (q == 2 ? -sn : -cs));
    const double c = q == 0 ? cs : (q == 1 ? -sn : (q == 2 ? -cs : sn));
    s_out = (float)s;
    c_out = (float)c;
}

__device__ __forceinline__ float sin_fixed(float x) {
    float s, c;
    sincos_fixed(x, s, c);
    return s;
}

__device__ __forceinline__ float mix(float a, float b, float t) { return a * (1.f - t) + b * t; }

// A circular arc of radius r about c from angle a0 to a1 as two cubics (k_gen_node_vg):
// p[0], p[3], p[6] at a0, the mean angle and a1 (direction (sin a, cos a)), handles along
// the tangents, of length sin((a1 - a0) / 2) * 0.353 * r.
struct Arc {
    float2 p[7];
    float2 d0, d6;  // directions of the ends
};

__device__ __forceinline__ Arc arc(float2 c, float r, float a0, float a1) {
    Arc A;
    // the middle angle is formed in double and rounded, as the reference's t0 * 0.5 + t1 * 0.5
    const float am = (float)((double)a0 * 0.5 + (double)a1 * 0.5);
    float2 d0, d3, d6;
    sincos_fixed(a0, d0.x, d0.y);
    sincos_fixed(am, d3.x, d3.y);
    sincos_fixed(a1, d6.x, d6.y);
    A.d0 = d0;
    A.d6 = d6;
    const float h = sin_fixed((a1 - a0) * .5f) * .353f * r;
    A.p[0] = make_float2(c.x + d0.x * r, c.y + d0.y * r);
    A.p[3] = make_float2(c.x + d3.x * r, c.y + d3.y * r);
    A.p[6] = make_float2(c.x + d6.x * r, c.y + d6.y * r);
    // p1 = p0 - n0 h, p2 = p3 + n3 h, p4 = p3 - n3 h, p5 = p6 + n6 h with n = (-d.y, d.x)
    A.p[1] = make_float2(A.p[0].x + d0.y * h, A.p[0].y - d0.x * h);
    A.p[2] = make_float2(A.p[3].x - d3.y * h, A.p[3].y + d3.x * h);
    A.p[4] = make_float2(A.p[3].x + d3.y * h, A.p[3].y - d3.x * h);
    A.p[5] = make_float2(A.p[6].x - d6.y * h, A.p[6].y + d6.x * h);
    return A;
}

// One curve of the path being built, into the block's staging arrays (local curve cl and
// vertex slot vl; the words hold the global vertex position vp and path): vertices in the
// fixed order, and the per-curve words the renderer reads.
__device__ __forceinline__ void emit(const ChordScene &S, uint32_t cl, uint32_t vl, uint32_t vp, uint32_t path, int nv,
                                     const float2 *p) {
    const uint32_t ixf = __float_as_uint(p[0].x), iyf = __float_as_uint(p[0].y);
    const uint32_t ixl = __float_as_uint(p[nv - 1].x), iyl = __float_as_uint(p[nv - 1].y);
    const bool rev = !(iyf < iyl || (iyf == iyl && ixf < ixl));
    for (int k = 0; k < nv; ++k) {
        S.vtx[vl + k] = p[rev ? nv - 1 - k : k];
        S.vpath[vl + k] = path;
    }
    S.ctype[cl] = (uint8_t)nv;
    S.cvpos[cl] = vp;
    S.cpath[cl] = path;
    S.cw[cl] = -1.f;
    S.crev[cl] = rev ? 1 : 0;
    S.cinfo[cl] = make_uint4(vp, path, (uint32_t)nv | ((rev ? 1u : 0u) << 8), __float_as_uint(-1.f));
}

__device__ __forceinline__ uint32_t fill_word(uint32_t c) { return (c & 0xFF000000u) ? c : 0u; }

struct ChordLaunch {
    int n_diagrams;
    ChordInterval d[kMaxChordDiagrams];
};

// Node i as path `path`, staged at local path pl.
__device__ void gen_node(const ChordInterval &D, int i, uint32_t path, uint32_t pl, const ChordScene &S) {
    const float t = D.t;
    const float a0 = mix(D.node_t0_a[i], D.node_t0_b[i], t);
    const float a1 = mix(D.node_t1_a[i], D.node_t1_b[i], t);
    Arc in, out;
    if (a0 == a1) {
        for (int k = 0; k < 7; ++k) in.p[k] = out.p[k] = make_float2(-1000.f, -1000.f);
    } else {
        in = arc(D.center, kInnerR, a0, a1);
        out = arc(D.center, kOuterR, a0, a1);
    }
    S.frule[pl] = 0;
    S.finfo[pl] = fill_word(D.node_color[i]);
    const uint32_t c = 6u * pl, v = 24u * pl, vg = 24u * path;
    float2 q[4];
    q[0] = out.p[0]; q[1] = in.p[0];
    emit(S, c + 0, v + 0, vg + 0, path, 2, q);
    q[0] = in.p[6]; q[1] = out.p[6];
    emit(S, c + 1, v + 2, vg + 2, path, 2, q);
    emit(S, c + 2, v + 4, vg + 4, path, 4, in.p);
    emit(S, c + 3, v + 8, vg + 8, path, 4, in.p + 3);
    q[0] = out.p[6]; q[1] = out.p[5]; q[2] = out.p[4]; q[3] = out.p[3];
    emit(S, c + 4, v + 12, vg + 12, path, 4, q);
    q[0] = out.p[3]; q[1] = out.p[2]; q[2] = out.p[1]; q[3] = out.p[0];
    emit(S, c + 5, v + 16, vg + 16, path, 4, q);
    for (int k = 20; k < 24; ++k) {  // slots no curve uses
        S.vtx[v + k] = make_float2(-1000.f, -1000.f);
        S.vpath[v + k] = path;
    }
}

// A link's ends are the stretches [pos, pos + size) of its nodes' sizes, mapped onto the
// nodes' angles.  A node of size 0 maps them to its start (the reference divides by zero).
__device__ void gen_link(const ChordInterval &D, int i, uint32_t path, uint32_t pl, const ChordScene &S) {
    const float t = D.t;
    const float4 fa = D.link_from[i], fb = D.link_to[i];  // {pos0, pos1, size0, size1}
    const float pos0 = mix(fa.x, fb.x, t), pos1 = mix(fa.y, fb.y, t);
    const float sz0 = mix(fa.z, fb.z, t), sz1 = mix(fa.w, fb.w, t);
    const uint2 nd = D.link_node[i];
    const float at0 = mix(D.node_t0_a[nd.x], D.node_t0_b[nd.x], t), at1 = mix(D.node_t1_a[nd.x], D.node_t1_b[nd.x], t);
    const float bt0 = mix(D.node_t0_a[nd.y], D.node_t0_b[nd.y], t), bt1 = mix(D.node_t1_a[nd.y], D.node_t1_b[nd.y], t);
    const float as = mix(D.node_size_a[nd.x], D.node_size_b[nd.x], t);
    const float bs = mix(D.node_size_a[nd.y], D.node_size_b[nd.y], t);
    const float fa0 = as != 0.f ? pos0 / as : 0.f, fa1 = as != 0.f ? (pos0 + sz0) / as : 0.f;
    const float fb0 = bs != 0.f ? pos1 / bs : 0.f, fb1 = bs != 0.f ? (pos1 + sz1) / bs : 0.f;
    const float la0 = mix(at0, at1, fa0), la1 = mix(at0, at1, fa1);
    const float lb0 = mix(bt0, bt1, fb0), lb1 = mix(bt0, bt1, fb1);
    const Arc A = arc(D.center, kInnerR, la0, la1), B = arc(D.center, kInnerR, lb0, lb1);
    float ts = fabsf(lb0 - la1), tl = fabsf(lb1 - la0);
    if (ts > kPi) ts -= kPi;
    if (tl > kPi) tl -= kPi;
    ts *= .5f;
    tl *= .5f;
    const float rs = kInnerR / (2.f * sin_fixed(ts) + 1.f), rl = kInnerR / (2.f * sin_fixed(tl) + 1.f);
    const float2 c = D.center;
    const float2 c00 = make_float2(c.x + A.d0.x * rl, c.y + A.d0.y * rl);
    const float2 c01 = make_float2(c.x + A.d6.x * rs, c.y + A.d6.y * rs);
    const float2 c10 = make_float2(c.x + B.d0.x * rs, c.y + B.d0.y * rs);
    const float2 c11 = make_float2(c.x + B.d6.x * rl, c.y + B.d6.y * rl);
    S.frule[pl] = 0;
    S.finfo[pl] = fill_word(D.node_color[sz0 > sz1 ? nd.x : nd.y] & 0x7FFFFFFFu);
    const uint32_t cv = 6u * pl, v = 24u * pl, vg = 24u * path;
    float2 q[4];
    q[0] = A.p[6]; q[1] = c01; q[2] = c10; q[3] = B.p[0];
    emit(S, cv + 0, v + 0, vg + 0, path, 4, q);
    q[0] = B.p[6]; q[1] = c11; q[2] = c00; q[3] = A.p[0];
    emit(S, cv + 1, v + 4, vg + 4, path, 4, q);
    emit(S, cv + 2, v + 8, vg + 8, path, 4, A.p);
    emit(S, cv + 3, v + 12, vg + 12, path, 4, A.p + 3);
    emit(S, cv + 4, v + 16, vg + 16, path, 4, B.p);
    emit(S, cv + 5, v + 20, vg + 20, path, 4, B.p + 3);
}

// One thread per path; the diagrams' paths follow one another, each diagram's nodes before
// its links (chord_animation, animation.cu:2582-2730).  A block stages its 64 paths' arrays
// in LDS and then writes each array's contiguous stretch with consecutive lanes (a path's
// own 24 vertices and 6 curves, written from one lane, would touch 64 lines a store).
constexpr uint32_t kGenPaths = 64;

__global__ void __launch_bounds__(64) k_chord_frame(ChordLaunch L, uint32_t P, ChordScene G) {
    __shared__ float2 s_vtx[kGenPaths * 24];
    __shared__ uint32_t s_vpath[kGenPaths * 24];
    __shared__ uint4 s_cinfo[kGenPaths * 6];
    __shared__ uint32_t s_cvpos[kGenPaths * 6], s_cpath[kGenPaths * 6], s_finfo[kGenPaths];
    __shared__ float s_cw[kGenPaths * 6];
    __shared__ uint8_t s_ctype[kGenPaths * 6], s_crev[kGenPaths * 6], s_frule[kGenPaths];
    const ChordScene S{s_vtx, s_vpath, s_cvpos, s_cpath, s_finfo, s_ctype, s_frule, s_crev, s_cw, s_cinfo};
    const uint32_t p0 = blockIdx.x * kGenPaths, pl = threadIdx.x, path = p0 + pl;
    if (path < P) {
        uint32_t base = 0;
        for (int d = 0; d < L.n_diagrams; ++d) {
            const ChordInterval &D = L.d[d];
            const uint32_t n = (uint32_t)D.n_nodes + (uint32_t)D.n_links;
            if (path < base + n) {
                const uint32_t i = path - base;
                if (i < (uint32_t)D.n_nodes) gen_node(D, (int)i, path, pl, S);
                else gen_link(D, (int)(i - (uint32_t)D.n_nodes), path, pl, S);
                break;
            }
            base += n;
        }
    }
    __syncthreads();
    const uint32_t np = P - p0 < kGenPaths ? P - p0 : kGenPaths;
    for (uint32_t k = threadIdx.x; k < 24u * np; k += kGenPaths) {
        G.vtx[24u * p0 + k] = s_vtx[k];
        G.vpath[24u * p0 + k] = s_vpath[k];
    }
    for (uint32_t k = threadIdx.x; k < 6u * np; k += kGenPaths) {
        const uint32_t c = 6u * p0 + k;
        G.cinfo[c] = s_cinfo[k];
        G.cvpos[c] = s_cvpos[k];
        G.cpath[c] = s_cpath[k];
        G.cw[c] = s_cw[k];
        G.ctype[c] = s_ctype[k];
        G.crev[c] = s_crev[k];
    }
    if (pl < np) {
        G.frule[path] = s_frule[pl];
        G.finfo[path] = s_finfo[pl];
    }
}

}  // namespace

// Eased interpolation weight of time T in its interval (GET_ANIMATION_T, animation.cu:
// 1710-1718, clamped to [0, 1] as k_evaluate_node does): 1 before the first interval ends,
// else the quadratic ease-in-out of (T mod 6000) / (6000 * 0.66666), held at 1.
float chord_ease(int time_ms) {
    float t = time_ms < kChordKeyMs ? 1.f : (float)(time_ms % kChordKeyMs) / ((float)kChordKeyMs * .66666f);
    if (t < 1.f) t = t < .5f ? 2.f * t * t : 1.f - 2.f * (1.f - t) * (1.f - t);
    else t = 1.f;
    t = t > 0.f ? t : 0.f;
    return t < 1.f ? t : 1.f;
}

void launch_chord_frame(const ChordInterval *d, int n_diagrams, uint32_t P, const ChordScene &out, hipStream_t s) {
    ChordLaunch L{};
    L.n_diagrams = n_diagrams;
    for (int i = 0; i < n_diagrams; ++i) L.d[i] = d[i];
    if (P) hipLaunchKernelGGL(k_chord_frame, dim3((P + kGenPaths - 1) / kGenPaths), dim3(kGenPaths), 0, s, L, P, out);
}

}  // namespace sl
