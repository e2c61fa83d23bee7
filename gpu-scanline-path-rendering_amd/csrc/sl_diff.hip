// sl_diff.hip -- differentiable mode: a forward renderer with continuous coverage and its
// hand-written adjoint (the reference has no such mode).
//
// Model.  A pixel row y holds S sample rows at y + (s + 1/2)/S.  Along each sample row the
// fill indicator of a path (nonzero / even-odd over the signed crossings of the row) is
// integrated exactly over every pixel's [x, x+1), so coverage is a box filter of S rows,
// continuous in the control points:
//     cov_p(x, y) = (1/S) * sum_s |{u in [x, x+1) : fill_p(u, y_s)}|
// and the image is the renderer's src-over onto white, in path order, unquantised:
//     a = paint_a * cov;  rgb = paint * a + rgb * (1 - a);  A = a * a + A * (1 - a).
//
// Forward: crossings of every y-monotone curve piece with the sample rows (half-open rule
// ymin <= y_s < ymax, bisection on t), bucketed by (path, sample row) and sorted by x; the
// fill change at each crossing; coverage per (path, pixel row) by one wavefront; the
// compositing chain per pixel (the colour before each path is kept for the adjoint).
// Adjoint: the chain reversed per pixel gives d/d paint and d/d cov; then every crossing
// with a fill change contributes d cov / d x_i = (fill left - fill right) / S in its pixel,
// and x_i = Bx(t*) with By(t*) = y_s moves with the control points as
//     d x_i / d Pk.x = b_k(t*),   d x_i / d Pk.y = -Bx'(t*) / By'(t*) * b_k(t*).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/scanline.h"
#include "sl_internal.h"

namespace {

template <class T> __device__ __forceinline__ T tmn(T a, T b) { return a < b ? a : b; }
template <class T> __device__ __forceinline__ T tmx(T a, T b) { return a > b ? a : b; }

struct DSeg {
    int type;
    float x[4], y[4];
};

__device__ __forceinline__ DSeg load_seg(const float2 *__restrict__ pts, const uint32_t *__restrict__ cvpos,
                                         const uint8_t *__restrict__ ctype, uint32_t c) {
    DSeg g;
    g.type = ctype[c];
    const int nv = g.type & 7;
    const uint32_t v0 = cvpos[c];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float2 p = pts[v0 + (uint32_t)(k < nv ? k : nv - 1)];
        g.x[k] = p.x;
        g.y[k] = p.y;
    }
    return g;
}

__device__ __forceinline__ float bez(int type, const float *p, float t) {
    const float s = 1.f - t;
    if (type == 2) return s * p[0] + t * p[1];
    if (type == 3) return s * s * p[0] + 2.f * s * t * p[1] + t * t * p[2];
    return s * s * s * p[0] + 3.f * s * s * t * p[1] + 3.f * s * t * t * p[2] + t * t * t * p[3];
}

__device__ __forceinline__ float dbez(int type, const float *p, float t) {
    const float s = 1.f - t;
    if (type == 2) return p[1] - p[0];
    if (type == 3) return 2.f * (s * (p[1] - p[0]) + t * (p[2] - p[1]));
    return 3.f * (s * s * (p[1] - p[0]) + 2.f * s * t * (p[2] - p[1]) + t * t * (p[3] - p[2]));
}

__device__ __forceinline__ void basis(int type, float t, float *b) {
    const float s = 1.f - t;
    b[0] = b[1] = b[2] = b[3] = 0.f;
    if (type == 2) {
        b[0] = s;
        b[1] = t;
    } else if (type == 3) {
        b[0] = s * s;
        b[1] = 2.f * s * t;
        b[2] = t * t;
    } else {
        b[0] = s * s * s;
        b[1] = 3.f * s * s * t;
        b[2] = 3.f * s * t * t;
        b[3] = t * t * t;
    }
}

// t of the y-extrema inside (0, 1), ascending; the piece boundaries are {0, these, 1}
__device__ __forceinline__ int y_extrema(int type, const float *y, float *r) {
    int n = 0;
    if (type == 3) {
        const float den = y[0] - 2.f * y[1] + y[2];
        if (den != 0.f) {
            const float t = (y[0] - y[1]) / den;
            if (t > 0.f && t < 1.f) r[n++] = t;
        }
    } else if (type == 4) {
        const float a = -y[0] + 3.f * y[1] - 3.f * y[2] + y[3];
        const float b = 2.f * (y[0] - 2.f * y[1] + y[2]);
        const float c = y[1] - y[0];
        if (a == 0.f) {
            if (b != 0.f) {
                const float t = -c / b;
                if (t > 0.f && t < 1.f) r[n++] = t;
            }
        } else {
            const float disc = b * b - 4.f * a * c;
            if (disc > 0.f) {
                const float sq = sqrtf(disc);
                float t0 = (-b - sq) / (2.f * a), t1 = (-b + sq) / (2.f * a);
                if (t0 > t1) {
                    const float tt = t0;
                    t0 = t1;
                    t1 = tt;
                }
                if (t0 > 0.f && t0 < 1.f) r[n++] = t0;
                if (t1 > 0.f && t1 < 1.f && t1 != t0) r[n++] = t1;
            }
        }
    }
    return n;
}

// first sample row j with (j + 1/2)/S >= y, clamped to [-1, J + 1] (NaN -> -1)
__device__ __forceinline__ long first_row(float y, int S, long J) {
    float v = ceilf(y * (float)S - 0.5f);
    if (!(v >= -1.f)) v = -1.f;
    if (v > (float)(J + 1)) v = (float)(J + 1);
    return (long)v;
}

// Pieces of a segment and their sample-row ranges [ja, jb); returns the piece count.
struct Pieces {
    int n;
    float t[4];
    int ja[3], jb[3], dir[3];
};

__device__ __forceinline__ Pieces pieces(const DSeg &g, int S, long J) {
    Pieces pc;
    float ts[4] = {0.f, 0.f, 0.f, 0.f};
    const int ne = y_extrema(g.type, g.y, ts + 1);
    ts[ne + 1] = 1.f;
    pc.n = ne + 1;
    float yprev = bez(g.type, g.y, 0.f);
    for (int k = 0; k < 4; ++k) pc.t[k] = ts[k];
    for (int k = 0; k < 3; ++k) {
        pc.ja[k] = pc.jb[k] = 0;
        pc.dir[k] = 0;
        if (k > ne) continue;
        const float y0 = yprev, y1 = bez(g.type, g.y, ts[k + 1]);
        yprev = y1;
        if (y0 == y1) continue;
        const float ylo = fminf(y0, y1), yhi = fmaxf(y0, y1);
        const long ja = tmx(first_row(ylo, S, J), 0L), jb = tmn(first_row(yhi, S, J), J);
        if (jb > ja) {
            pc.ja[k] = (int)ja;
            pc.jb[k] = (int)jb;
        }
        pc.dir[k] = y1 > y0 ? 1 : -1;
    }
    return pc;
}

struct DiffDev {
    int W, H, S, P;
    uint32_t C;
    const uint32_t *cvpos, *cpath;
    const uint8_t *ctype, *frule;
};

__global__ void k_d_count(DiffDev D, const float2 *__restrict__ pts, uint32_t *cnt) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.C) return;
    const DSeg g = load_seg(pts, D.cvpos, D.ctype, c);
    const Pieces pc = pieces(g, D.S, (long)D.H * D.S);
    uint32_t n = 0;
    for (int k = 0; k < pc.n; ++k) n += (uint32_t)(pc.jb[k] - pc.ja[k]);
    cnt[c] = n;
}

struct Cross {
    float x, t;
    uint32_t curve;
    uint32_t row_dir;  // sample row << 1 | (dir > 0)
};

__global__ void k_d_cross(DiffDev D, const float2 *__restrict__ pts, const uint32_t *__restrict__ off, uint32_t NX,
                          Cross *xs, uint32_t *key, uint32_t *bcount) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NX) return;
    // the curve holding crossing i: last c with off[c] <= i
    uint32_t lo = 0, hi = D.C - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid - 1;
    }
    const uint32_t c = lo;
    const DSeg g = load_seg(pts, D.cvpos, D.ctype, c);
    const Pieces pc = pieces(g, D.S, (long)D.H * D.S);
    uint32_t r = i - off[c];
    int k = 0;
    for (; k < pc.n - 1; ++k) {
        const uint32_t m = (uint32_t)(pc.jb[k] - pc.ja[k]);
        if (r < m) break;
        r -= m;
    }
    const int j = pc.ja[k] + (int)r, dir = pc.dir[k];
    const float ys = ((float)j + 0.5f) / (float)D.S;
    float tl = dir > 0 ? pc.t[k] : pc.t[k + 1], th = dir > 0 ? pc.t[k + 1] : pc.t[k];
    for (int it = 0; it < 32; ++it) {
        const float mid = 0.5f * (tl + th);
        if (bez(g.type, g.y, mid) < ys) tl = mid; else th = mid;
    }
    const float t = 0.5f * (tl + th);
    Cross o;
    o.x = bez(g.type, g.x, t);
    o.t = t;
    o.curve = c;
    o.row_dir = ((uint32_t)j << 1) | (dir > 0 ? 1u : 0u);
    xs[i] = o;
    const uint32_t b = D.cpath[c] * (uint32_t)(D.H * D.S) + (uint32_t)j;
    key[i] = b;
    atomicAdd(&bcount[b], 1u);
}

__global__ void k_d_scatter(uint32_t NX, const uint32_t *__restrict__ key, const uint32_t *__restrict__ boff,
                            uint32_t *bcur, uint32_t *order) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NX) return;
    const uint32_t b = key[i];
    order[boff[b] + atomicAdd(&bcur[b], 1u)] = i;
}

__device__ __forceinline__ bool filled(int rule, int w) { return rule ? (w & 1) != 0 : w != 0; }

__device__ __forceinline__ bool cross_less(const Cross &a, const Cross &b) {
    if (a.x != b.x) return a.x < b.x;
    if (a.curve != b.curve) return a.curve < b.curve;
    return a.t < b.t;
}

// One thread per (path, sample row) bucket: sort by (x, curve, t), then the fill change
// at every crossing; the sorted x and direction go to `sx` / `sd` for the coverage pass.
__global__ void k_d_sort(DiffDev D, uint32_t NBK, const uint32_t *__restrict__ boff, const Cross *__restrict__ xs,
                         uint32_t *order, float *sx, int8_t *sd, int8_t *df) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= NBK) return;
    const uint32_t a0 = boff[b], a1 = boff[b + 1];
    if (a1 == a0) return;
    for (uint32_t k = a0 + 1; k < a1; ++k) {
        const uint32_t v = order[k];
        const Cross cv = xs[v];
        uint32_t m = k;
        while (m > a0 && cross_less(cv, xs[order[m - 1]])) {
            order[m] = order[m - 1];
            --m;
        }
        order[m] = v;
    }
    const int rule = D.frule[b / (uint32_t)(D.H * D.S)];
    int w = 0;
    for (uint32_t k = a0; k < a1; ++k) {
        const Cross cv = xs[order[k]];
        const int d = (cv.row_dir & 1u) ? 1 : -1;
        sx[k] = cv.x;
        sd[k] = (int8_t)d;
        df[k] = (int8_t)((filled(rule, w) ? 1 : 0) - (filled(rule, w + d) ? 1 : 0));
        w += d;
    }
}

// One wavefront per (path, pixel row); lanes are pixel columns.
__global__ void __launch_bounds__(256) k_d_coverage(DiffDev D, const uint32_t *__restrict__ boff,
                                                    const float *__restrict__ sx, const int8_t *__restrict__ sd,
                                                    float *cov) {
    const int lane = threadIdx.x & 63;
    const int64_t wid = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wid >= (int64_t)D.P * D.H) return;
    const int p = (int)(wid / D.H), y = (int)(wid - (int64_t)p * D.H);
    const int rule = D.frule[p];
    const uint32_t b0 = (uint32_t)p * (uint32_t)(D.H * D.S) + (uint32_t)(y * D.S);
    const float Wf = (float)D.W;
    float *row = cov + ((size_t)p * D.H + y) * D.W;
    for (int x0 = 0; x0 < D.W; x0 += 64) {
        const int px = x0 + lane;
        const float fpx = (float)px, fpx1 = (float)(px + 1);
        float acc = 0.f;
        for (int s = 0; s < D.S; ++s) {
            const uint32_t a0 = boff[b0 + s], a1 = boff[b0 + s + 1];
            int w = 0;
            float prev = 0.f;
            for (uint32_t k = a0; k < a1; ++k) {
                const float x = sx[k];
                if (k > a0 && filled(rule, w)) {
                    const float u0 = fmaxf(prev, 0.f), u1 = fminf(x, Wf);
                    if (u1 > u0) {
                        const int p0 = (int)floorf(u0);
                        const int p1 = min((int)floorf(u1), D.W - 1);
                        if (px >= p0 && px <= p1) {
                            const float lo = fmaxf(u0, fpx), hi = fminf(u1, fpx1);
                            if (hi > lo) acc += (hi - lo) / (float)D.S;
                        }
                    }
                }
                w += sd[k];
                prev = x;
            }
        }
        if (px < D.W) row[px] = acc;
    }
}

__global__ void k_d_composite(DiffDev D, const float *__restrict__ cov, const float4 *__restrict__ col,
                              float4 *__restrict__ pre, float4 *__restrict__ img) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t NP = (int64_t)D.W * D.H;
    if (i >= NP) return;
    float I0 = 1.f, I1 = 1.f, I2 = 1.f, I3 = 1.f;
    for (int p = 0; p < D.P; ++p) {
        if (pre) pre[(size_t)p * NP + i] = make_float4(I0, I1, I2, I3);
        const float4 c = col[p];
        const float a = c.w * cov[(size_t)p * NP + i];
        const float oma = 1.f - a;
        I0 = c.x * a + I0 * oma;
        I1 = c.y * a + I1 * oma;
        I2 = c.z * a + I2 * oma;
        I3 = a * a + I3 * oma;
    }
    img[i] = make_float4(I0, I1, I2, I3);
}

// Gradient sums are accumulated in 64-bit fixed point, 2^-32 units: each contribution is
// rounded once (to nearest) and added with an integer atomic.  Integer addition is
// associative, so the sums do not depend on the order the atomics land in and every run
// gives the same bits; a float atomicAdd's result depends on that order.  Contributions
// are clamped to |v| <= 2^30 and the sums stay far inside the int64 range for any
// gradient image of moderate size.
constexpr double kFixScale = 4294967296.0;  // 2^32

__device__ __forceinline__ void fix_add(unsigned long long *acc, float v) {
    const double c = fmin(fmax((double)v, -1073741824.0), 1073741824.0);
    atomicAdd(acc, (unsigned long long)__double2ll_rn(c * kFixScale));
}

__global__ void k_d_unfix(const long long *__restrict__ acc, float *__restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (float)((double)acc[i] * (1.0 / kFixScale));
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Adjoint of the compositing chain, one thread per pixel.
__global__ void __launch_bounds__(256) k_d_back_pixels(DiffDev D, const float *__restrict__ cov,
                                                       const float4 *__restrict__ col, const float4 *__restrict__ pre,
                                                       const float4 *__restrict__ gimg, float *gcov,
                                                       unsigned long long *gcol) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t NP = (int64_t)D.W * D.H;
    const bool in = i < NP;
    float4 g = in ? gimg[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    for (int p = D.P - 1; p >= 0; --p) {
        const float4 c = col[p];
        float cv = 0.f, a = 0.f, da = 0.f;
        float4 pr = make_float4(0.f, 0.f, 0.f, 0.f);
        if (in) {
            pr = pre[(size_t)p * NP + i];
            cv = cov[(size_t)p * NP + i];
            a = c.w * cv;
            da = g.x * (c.x - pr.x) + g.y * (c.y - pr.y) + g.z * (c.z - pr.z) + g.w * (2.f * a - pr.w);
            gcov[(size_t)p * NP + i] = da * c.w;
        }
        const float s0 = wave_sum(g.x * a), s1 = wave_sum(g.y * a), s2 = wave_sum(g.z * a), s3 = wave_sum(da * cv);
        if ((threadIdx.x & 63) == 0 && (s0 != 0.f || s1 != 0.f || s2 != 0.f || s3 != 0.f)) {
            fix_add(&gcol[4 * p + 0], s0);
            fix_add(&gcol[4 * p + 1], s1);
            fix_add(&gcol[4 * p + 2], s2);
            fix_add(&gcol[4 * p + 3], s3);
        }
        const float oma = 1.f - a;
        g.x *= oma;
        g.y *= oma;
        g.z *= oma;
        g.w *= oma;
    }
}

// Boundary terms at the crossings, one thread per sorted crossing.
__global__ void k_d_back_cross(DiffDev D, uint32_t NX, const uint32_t *__restrict__ order, const Cross *__restrict__ xs,
                               const int8_t *__restrict__ df, const float2 *__restrict__ pts,
                               const float *__restrict__ gcov, unsigned long long *gpts) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= NX) return;
    const int f = df[k];
    if (!f) return;
    const Cross cr = xs[order[k]];
    if (!(cr.x >= 0.f && cr.x < (float)D.W)) return;
    const uint32_t c = cr.curve;
    const int j = (int)(cr.row_dir >> 1), y = j / D.S;
    const int px = (int)floorf(cr.x);
    const int p = (int)D.cpath[c];
    const size_t NP = (size_t)D.W * D.H;
    const float gx = gcov[(size_t)p * NP + (size_t)y * D.W + px] * (float)f / (float)D.S;
    if (gx == 0.f) return;
    const DSeg g = load_seg(pts, D.cvpos, D.ctype, c);
    float b[4];
    basis(g.type, cr.t, b);
    // dx/dy along the curve; near a y-extremum By' -> 0, bounded at |By'| >= 1e-4
    float dyv = dbez(g.type, g.y, cr.t);
    if (fabsf(dyv) < 1e-4f) dyv = dyv < 0.f ? -1e-4f : 1e-4f;
    const float ratio = dbez(g.type, g.x, cr.t) / dyv;
    const int nv = g.type & 7;
    const uint32_t v0 = D.cvpos[c];
    for (int v = 0; v < nv; ++v) {
        fix_add(&gpts[2 * (size_t)(v0 + v)], gx * b[v]);
        fix_add(&gpts[2 * (size_t)(v0 + v) + 1], -gx * ratio * b[v]);
    }
}

__global__ void k_d_fill(float *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.f;
}

inline unsigned nblk(int64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

struct Buf {
    void *p = nullptr;
    size_t cap = 0;
};

}  // namespace

struct sl_diff {
    int device = 0;
    int W = 0, H = 0, S = 16;
    uint32_t P = 0, C = 0, V = 0;
    Buf cvpos, cpath, ctype, frule;
    Buf cnt, off, xs, key, bcount, boff, bcur, order, sx, sd, df, cov, pre, gcov, scan_tmp;
    Buf gfix;  // fixed-point gradient sums: 2V point components, then 4P colour components
    uint32_t NX = 0;
    bool have_pre = false;
    const float *last_points = nullptr, *last_colors = nullptr;
    // the last forward or backward, possibly still running on last_stream
    hipEvent_t done_ev = nullptr;
    hipStream_t last_stream = nullptr;
    bool pending = false;
    std::string err;
};

namespace {

int dfail(sl_diff *d, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (d) d->err = buf;
    return code;
}

// Forward and backward share the handle's buffers.  Work issued on another stream than
// the last call's waits on the host until that call has finished.
int diff_order(sl_diff *d, hipStream_t s) {
    if (d->pending && s != d->last_stream) {
        if (hipEventSynchronize(d->done_ev) != hipSuccess) return dfail(d, SL_E_HIP, "previous call failed");
        d->pending = false;
    }
    return SL_OK;
}

int diff_issued(sl_diff *d, hipStream_t s) {
    if (!d->done_ev && hipEventCreateWithFlags(&d->done_ev, hipEventDisableTiming) != hipSuccess)
        return dfail(d, SL_E_HIP, "hipEventCreate failed");
    if (hipEventRecord(d->done_ev, s) != hipSuccess) return dfail(d, SL_E_HIP, "hipEventRecord failed");
    d->pending = true;
    d->last_stream = s;
    return SL_OK;
}

int dens(sl_diff *d, Buf &b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return SL_OK;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    const size_t want = bytes + bytes / 4;
    if (hipMalloc(&b.p, want) != hipSuccess) {
        b.p = nullptr;
        (void)hipGetLastError();
        return dfail(d, SL_E_OOM, "hipMalloc(%zu) failed", want);
    }
    b.cap = want;
    return SL_OK;
}

template <class T> T *BP(Buf &b) { return (T *)b.p; }

DiffDev dev_view(sl_diff *d) {
    DiffDev D;
    D.W = d->W; D.H = d->H; D.S = d->S; D.P = (int)d->P; D.C = d->C;
    D.cvpos = BP<uint32_t>(d->cvpos); D.cpath = BP<uint32_t>(d->cpath);
    D.ctype = BP<uint8_t>(d->ctype); D.frule = BP<uint8_t>(d->frule);
    return D;
}

}  // namespace

extern "C" {

int sl_diff_create(int device, int width, int height, int sample_rows, uint32_t n_paths, uint32_t n_curves,
                   uint32_t n_vertices, const uint32_t *curve_vertex_pos, const uint8_t *curve_type,
                   const uint32_t *curve_path_id, const uint8_t *fill_rule, sl_diff **out) {
    if (!out || width <= 0 || height <= 0 || sample_rows <= 0 || sample_rows > 64 || n_paths == 0 ||
        n_curves == 0 || !curve_vertex_pos || !curve_type || !curve_path_id || !fill_rule)
        return SL_E_ARG;
    *out = nullptr;
    if ((uint64_t)n_paths * height * sample_rows >= 0xFFFFFFFFull) return SL_E_UNSUPPORTED;
    if ((uint64_t)n_paths * width * height > (1ull << 28)) return SL_E_UNSUPPORTED;
    for (uint32_t c = 0; c < n_curves; ++c) {
        const int t = curve_type[c];
        if (t != SL_CURVE_LINEAR && t != SL_CURVE_QUADRATIC && t != SL_CURVE_CUBIC) return SL_E_UNSUPPORTED;
        if ((uint64_t)curve_vertex_pos[c] + (t & 7) > n_vertices) return SL_E_ARG;
        if (curve_path_id[c] >= n_paths || (c && curve_path_id[c] < curve_path_id[c - 1])) return SL_E_ARG;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return SL_E_NODEVICE;
    }
    if (device < 0) device = 0;
    if (device >= n || hipSetDevice(device) != hipSuccess) return SL_E_ARG;
    sl_diff *d = new (std::nothrow) sl_diff;
    if (!d) return SL_E_OOM;
    d->device = device;
    d->W = width; d->H = height; d->S = sample_rows;
    d->P = n_paths; d->C = n_curves; d->V = n_vertices;
    int rc;
    if ((rc = dens(d, d->cvpos, (size_t)n_curves * 4)) || (rc = dens(d, d->cpath, (size_t)n_curves * 4)) ||
        (rc = dens(d, d->ctype, n_curves)) || (rc = dens(d, d->frule, n_paths))) {
        sl_diff_destroy(d);
        return rc;
    }
    if (hipMemcpy(d->cvpos.p, curve_vertex_pos, (size_t)n_curves * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d->cpath.p, curve_path_id, (size_t)n_curves * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d->ctype.p, curve_type, n_curves, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d->frule.p, fill_rule, n_paths, hipMemcpyHostToDevice) != hipSuccess) {
        sl_diff_destroy(d);
        return SL_E_HIP;
    }
    *out = d;
    return SL_OK;
}

int sl_diff_destroy(sl_diff *d) {
    if (!d) return SL_OK;
    (void)hipSetDevice(d->device);
    (void)hipDeviceSynchronize();
    Buf *bs[] = {&d->cvpos, &d->cpath, &d->ctype, &d->frule, &d->cnt, &d->off, &d->xs, &d->key, &d->bcount,
                 &d->boff, &d->bcur, &d->order, &d->sx, &d->sd, &d->df, &d->cov, &d->pre, &d->gcov, &d->scan_tmp,
                 &d->gfix};
    for (Buf *b : bs)
        if (b->p) (void)hipFree(b->p);
    if (d->done_ev) (void)hipEventDestroy(d->done_ev);
    delete d;
    return SL_OK;
}

const char *sl_diff_last_error(sl_diff *d) { return d ? d->err.c_str() : "null handle"; }

int sl_diff_forward(sl_diff *d, const float *points, const float *colors, float *image, int keep_for_backward,
                    void *stream) {
    if (!d || !points || !colors || !image) return SL_E_ARG;
    if (hipSetDevice(d->device) != hipSuccess) return dfail(d, SL_E_HIP, "hipSetDevice failed");
    hipStream_t s = (hipStream_t)stream;
    if (int rc0 = diff_order(d, s)) return rc0;
    const DiffDev D = dev_view(d);
    const float2 *pts = (const float2 *)points;
    const uint32_t C = d->C;
    const uint32_t NBK = d->P * (uint32_t)(d->H * d->S);
    const size_t NP = (size_t)d->W * d->H;
    int rc;
    if ((rc = dens(d, d->cnt, ((size_t)C + 1) * 4)) || (rc = dens(d, d->off, ((size_t)C + 1) * 4)) ||
        (rc = dens(d, d->bcount, ((size_t)NBK + 1) * 4)) || (rc = dens(d, d->boff, ((size_t)NBK + 1) * 4)) ||
        (rc = dens(d, d->bcur, (size_t)NBK * 4)) || (rc = dens(d, d->cov, (size_t)d->P * NP * 4)) ||
        (rc = dens(d, d->scan_tmp, sl::scan_tmp_elems_needed(std::max<size_t>(C, NBK) + 1) * 4)))
        return rc;
    hipLaunchKernelGGL(k_d_count, dim3(nblk(C, 256)), dim3(256), 0, s, D, pts, BP<uint32_t>(d->cnt));
    sl::scan_exclusive_u32(BP<uint32_t>(d->cnt), BP<uint32_t>(d->off), C, BP<uint32_t>(d->scan_tmp), s);
    uint32_t NX = 0;
    if (hipMemcpyAsync(&NX, BP<uint32_t>(d->off) + C, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return dfail(d, SL_E_HIP, "crossing count read-back failed");
    d->NX = NX;
    if ((rc = dens(d, d->xs, (size_t)NX * sizeof(Cross))) || (rc = dens(d, d->key, (size_t)NX * 4)) ||
        (rc = dens(d, d->order, (size_t)NX * 4)) || (rc = dens(d, d->sx, (size_t)NX * 4)) ||
        (rc = dens(d, d->sd, NX)) || (rc = dens(d, d->df, NX)))
        return rc;
    sl::launch_fill_u32(BP<uint32_t>(d->bcount), 0u, NBK, s);
    sl::launch_fill_u32(BP<uint32_t>(d->bcur), 0u, NBK, s);
    if (NX) {
        hipLaunchKernelGGL(k_d_cross, dim3(nblk(NX, 256)), dim3(256), 0, s, D, pts, (const uint32_t *)BP<uint32_t>(d->off),
                           NX, BP<Cross>(d->xs), BP<uint32_t>(d->key), BP<uint32_t>(d->bcount));
    }
    sl::scan_exclusive_u32(BP<uint32_t>(d->bcount), BP<uint32_t>(d->boff), NBK, BP<uint32_t>(d->scan_tmp), s);
    if (NX) {
        hipLaunchKernelGGL(k_d_scatter, dim3(nblk(NX, 256)), dim3(256), 0, s, NX, (const uint32_t *)BP<uint32_t>(d->key),
                           (const uint32_t *)BP<uint32_t>(d->boff), BP<uint32_t>(d->bcur), BP<uint32_t>(d->order));
        hipLaunchKernelGGL(k_d_sort, dim3(nblk(NBK, 256)), dim3(256), 0, s, D, NBK, (const uint32_t *)BP<uint32_t>(d->boff),
                           (const Cross *)BP<Cross>(d->xs), BP<uint32_t>(d->order), BP<float>(d->sx), BP<int8_t>(d->sd),
                           BP<int8_t>(d->df));
    }
    hipLaunchKernelGGL(k_d_coverage, dim3(nblk((int64_t)d->P * d->H, 4)), dim3(256), 0, s, D,
                       (const uint32_t *)BP<uint32_t>(d->boff), (const float *)BP<float>(d->sx),
                       (const int8_t *)BP<int8_t>(d->sd), BP<float>(d->cov));
    float4 *pre = nullptr;
    if (keep_for_backward) {
        if ((rc = dens(d, d->pre, (size_t)d->P * NP * 16))) return rc;
        pre = BP<float4>(d->pre);
    }
    hipLaunchKernelGGL(k_d_composite, dim3(nblk((int64_t)NP, 256)), dim3(256), 0, s, D, (const float *)BP<float>(d->cov),
                       (const float4 *)colors, pre, (float4 *)image);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return dfail(d, SL_E_HIP, "forward launch failed: %s", hipGetErrorString(e));
    d->have_pre = keep_for_backward != 0;
    d->last_points = points;
    d->last_colors = colors;
    return diff_issued(d, s);
}

int sl_diff_backward(sl_diff *d, const float *grad_image, float *grad_points, float *grad_colors, void *stream) {
    if (!d || !grad_image || !grad_points || !grad_colors) return SL_E_ARG;
    if (!d->have_pre) return dfail(d, SL_E_STATE, "backward needs a forward with keep_for_backward");
    if (hipSetDevice(d->device) != hipSuccess) return dfail(d, SL_E_HIP, "hipSetDevice failed");
    hipStream_t s = (hipStream_t)stream;
    if (int rc0 = diff_order(d, s)) return rc0;
    const DiffDev D = dev_view(d);
    const size_t NP = (size_t)d->W * d->H;
    int rc;
    const size_t npt = (size_t)d->V * 2, ncol = (size_t)d->P * 4;
    if ((rc = dens(d, d->gcov, (size_t)d->P * NP * 4)) || (rc = dens(d, d->gfix, (npt + ncol) * 8))) return rc;
    unsigned long long *fix_pts = BP<unsigned long long>(d->gfix), *fix_col = fix_pts + npt;
    hipLaunchKernelGGL(k_d_fill, dim3(256), dim3(256), 0, s, BP<float>(d->gfix), (npt + ncol) * 2);
    hipLaunchKernelGGL(k_d_back_pixels, dim3(nblk((int64_t)NP, 256)), dim3(256), 0, s, D, (const float *)BP<float>(d->cov),
                       (const float4 *)d->last_colors, (const float4 *)BP<float4>(d->pre), (const float4 *)grad_image,
                       BP<float>(d->gcov), fix_col);
    if (d->NX)
        hipLaunchKernelGGL(k_d_back_cross, dim3(nblk(d->NX, 256)), dim3(256), 0, s, D, d->NX,
                           (const uint32_t *)BP<uint32_t>(d->order), (const Cross *)BP<Cross>(d->xs),
                           (const int8_t *)BP<int8_t>(d->df), (const float2 *)d->last_points,
                           (const float *)BP<float>(d->gcov), fix_pts);
    hipLaunchKernelGGL(k_d_unfix, dim3(256), dim3(256), 0, s, (const long long *)fix_pts, grad_points, npt);
    hipLaunchKernelGGL(k_d_unfix, dim3(16), dim3(256), 0, s, (const long long *)fix_col, grad_colors, ncol);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return dfail(d, SL_E_HIP, "backward launch failed: %s", hipGetErrorString(e));
    return diff_issued(d, s);
}

int sl_diff_crossings(sl_diff *d, int64_t *n) {
    if (!d || !n) return SL_E_ARG;
    *n = d->NX;
    return SL_OK;
}

}  // extern "C"
