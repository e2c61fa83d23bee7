// sl_geom.hip -- scans, vertex transform/cull, monotone split + fragment counting and
// curve / cut-line intersection (the "ts" array) of the comb-scanline pipeline.
#include "sl_device.h"
#include "sl_internal.h"

namespace sl {
using namespace dev;

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// Exclusive scan of the 256 per-thread values of a block; returns the block total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t &total, uint32_t *lds /*[4]*/) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) lds[wid] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        uint32_t x = lds[w];
        if (w < wid) off += x;
        tot += x;
    }
    __syncthreads();
    total = tot;
    return off + inc - v;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const uint32_t *in, size_t n, uint32_t *tile_sum) {
    __shared__ uint32_t lds[4];
    size_t base = (size_t)blockIdx.x * kScanTile;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        size_t i = base + (size_t)k * kScanThreads + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t tot;
    block_excl_scan(s, tot, lds);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_top(uint32_t *tile_sum, size_t ntiles, uint32_t *out_total) {
    __shared__ uint32_t lds[4];
    uint32_t carry = 0;
    for (size_t c = 0; c < ntiles; c += kScanThreads) {
        size_t i = c + threadIdx.x;
        uint32_t v = i < ntiles ? tile_sum[i] : 0u;
        uint32_t tot;
        uint32_t ex = block_excl_scan(v, tot, lds);
        if (i < ntiles) tile_sum[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *out_total = carry;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_down(const uint32_t *in, uint32_t *out, size_t n,
                                                             const uint32_t *tile_off) {
    __shared__ uint32_t lds[4];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        size_t i = base + k;
        v[k] = i < n ? in[i] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan(s, tot, lds) + tile_off[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        size_t i = base + k;
        if (i < n) out[i] = ex;
        ex += v[k];
    }
}

// The down-sweep for scans of at most kScanDirectTiles tiles: each block sums the tile
// totals before its own (at most 4095 L2-resident words) instead of a separate pass
// scanning them, and the last block writes the grand total to out[n].
constexpr size_t kScanDirectTiles = 4096;

__global__ void __launch_bounds__(kScanThreads) k_scan_down_direct(const uint32_t *in, uint32_t *out, size_t n,
                                                                    const uint32_t *tile_sum) {
    __shared__ uint32_t lds[4];
    uint32_t pre = 0;
    for (uint32_t i = threadIdx.x; i < blockIdx.x; i += kScanThreads) pre += tile_sum[i];
    uint32_t tile_off;
    block_excl_scan(pre, tile_off, lds);
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        size_t i = base + k;
        v[k] = i < n ? in[i] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan(s, tot, lds) + tile_off;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        size_t i = base + k;
        if (i < n) out[i] = ex;
        ex += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = tile_off + tot;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
// two arrays of the same length in one launch
__global__ void k_fill2_u32(uint32_t *p, uint32_t *q, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        p[i] = v;
        q[i] = v;
    }
}
__global__ void k_fill_u8(uint8_t *p, uint8_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// rows of `row_bytes` bytes, `pitch` apart (the bytes between rows are left alone)
__global__ void k_fill_rows_u8(uint8_t *p, uint8_t v, size_t row_bytes, size_t pitch, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[(i / row_bytes) * pitch + i % row_bytes] = v;
}

inline unsigned blocks_for(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// VALU issue-rate probe: every thread runs `iters` rounds of 64 steps of 8 independent fma
// chains, i.e. 512 v_fma_f32 per round and wave, nothing else in the loop body but the
// loop counter (scalar).  Timed over a full grid it gives the wave VALU instructions per
// second the machine issues.
__global__ void __launch_bounds__(256) k_valu_probe(int iters, float seed, float *out) {
    float a0 = seed + threadIdx.x, a1 = a0 + 1.f, a2 = a0 + 2.f, a3 = a0 + 3.f;
    float a4 = a0 + 4.f, a5 = a0 + 5.f, a6 = a0 + 6.f, a7 = a0 + 7.f;
    const float m = 0.999f, c = 1e-3f;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 64; ++k) {
            a0 = __builtin_fmaf(a0, m, c); a1 = __builtin_fmaf(a1, m, c);
            a2 = __builtin_fmaf(a2, m, c); a3 = __builtin_fmaf(a3, m, c);
            a4 = __builtin_fmaf(a4, m, c); a5 = __builtin_fmaf(a5, m, c);
            a6 = __builtin_fmaf(a6, m, c); a7 = __builtin_fmaf(a7, m, c);
        }
    }
    const float r = ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7));
    if (r == 12345.f) out[blockIdx.x] = r;  // keeps the chains live; never true in practice
}

// ------------------------------------------------------------------------------------
// per-path init, vertex transform + region flags (k_transformVertex, .cu:36-105)
// ------------------------------------------------------------------------------------
__global__ void k_path_init(uint32_t P, uint32_t *pflags, int2 *prow) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    pflags[p] = 0u;
    prow[p] = make_int2(INT_MAX, INT_MIN);
}

struct Mat4 { float m[16]; };

__global__ void k_transform(uint32_t V, const float2 *__restrict__ vtx, const uint32_t *__restrict__ vpath,
                            const int4 *__restrict__ pcanvas, Mat4 M, int W, int H, float2 *__restrict__ out,
                            uint32_t *pflags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < V;
    const uint32_t p = live ? vpath[i] : 0u;
    if (pcanvas && live) {
        const int4 cv = pcanvas[p];
        W = cv.x;
        H = cv.y;
    }
    const float2 iv = live ? vtx[i] : make_float2(0.f, 0.f);
    const float ix = iv.x, iy = iv.y, iz = 0.f, iw = 1.f;
    float o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        o[k] = M.m[4 * k + 0] * ix + M.m[4 * k + 1] * iy + M.m[4 * k + 2] * iz + M.m[4 * k + 3] * iw;
    o[0] /= o[3];
    o[1] /= o[3];
    o[2] /= o[3];
    const float w = (float)W, h = (float)H;
    const float z = (float)(W < H ? W : H);
    const float cutoff = -z * 0.5f;
    if (o[2] < cutoff) o[2] = cutoff;
    const float w2 = w * .5f, h2 = h * .5f;
    float ox = (o[0] - w2) * z / (o[2] + z) + w2;
    float oy = (o[1] - h2) * z / (o[2] + z) + h2;
    int xf = ox < 0 ? 0 : (ox <= w ? 1 : 2);
    int yf = oy < 0 ? 0 : (oy <= h ? 1 : 2);
    // region bit per the 3x3 layout of ras_define.h:104-117: regions r = 3 yf + xf take bits
    // 0, 1, 2, 3, -, 4, 5, 6, 7 in order; the inside (r = 4) sets bits 0 and 7
    const int r = 3 * yf + xf;
    const uint32_t bits = r == 4 ? ((1u << 0) | (1u << 7)) : 1u << (r - (r > 4 ? 1 : 0));
    // OR the region bits into the path's flags: one atomic per distinct path of the wave
    const int lane = threadIdx.x & 63;
    uint64_t pend = __ballot(live);
    while (pend) {
        const int leader = __ffsll((long long)pend) - 1;
        const uint32_t lp = (uint32_t)__builtin_amdgcn_readlane((int)p, leader);
        const bool mine = ((pend >> lane) & 1ull) && p == lp;
        const uint64_t grp = __ballot(mine);
        // the group's OR of the eight region bits, one ballot a bit (scalar results, no
        // cross-lane data movement)
        uint32_t acc = 0u;
#pragma unroll
        for (int b = 0; b < 8; ++b) acc |= __ballot(mine && ((bits >> b) & 1u)) ? (1u << b) : 0u;
        if (lane == leader) atomicOr(&pflags[lp], acc);
        pend &= ~grp;
    }
    if (!live) return;
    if (floorf(ox) == ox) ox = ox - 1.f / 256.f;
    out[i] = make_float2(ox, oy);
}

// ------------------------------------------------------------------------------------
// curve loading
// ------------------------------------------------------------------------------------
struct Curve {
    int type;
    float w;
    float x0, x1, x2, x3, y0, y1, y2, y3;
};

// A curve and its monotone cuts from its 64-byte record (k_monotone writes it): one cache
// line per curve for the intersection passes, whose work lists visit curves out of order.
struct CurveRec {
    Curve cu;
    float4 q;       // monotone cuts
    int n;          // cut count
    uint32_t path;
};

__device__ __forceinline__ CurveRec load_rec(const float4 *__restrict__ crec, uint32_t c) {
    CurveRec r;
    const float4 a = crec[4 * (size_t)c], b = crec[4 * (size_t)c + 1];
    r.q = crec[4 * (size_t)c + 2];
    const float4 m = crec[4 * (size_t)c + 3];
    const uint32_t tw = fb(m.x);
    r.cu.type = (int)(tw & 0xFFu);
    r.n = (int)((tw >> 8) & 0xFFu);
    r.cu.w = m.y;
    r.path = fb(m.z);
    r.cu.x0 = a.x; r.cu.y0 = a.y; r.cu.x1 = a.z; r.cu.y1 = a.w;
    r.cu.x2 = b.x; r.cu.y2 = b.y; r.cu.x3 = b.z; r.cu.y3 = b.w;
    return r;
}

__device__ __forceinline__ Curve load_curve(const SceneDev &sc, const float2 *__restrict__ tv, uint32_t c) {
    Curve cu;
    const uint4 ci = sc.cinfo[c];
    cu.type = (int)(ci.z & 0xFFu);
    cu.w = (cu.type == CT_RATIONAL) ? bf(ci.w) : 0.f;
    const uint32_t p = ci.x;
    const int nv = cu.type & 7;
    float2 a = tv[p], b = tv[p + 1];
    float2 d = nv > 2 ? tv[p + 2] : make_float2(0.f, 0.f);
    float2 e = nv > 3 ? tv[p + 3] : make_float2(0.f, 0.f);
    cu.x0 = a.x; cu.y0 = a.y; cu.x1 = b.x; cu.y1 = b.y;
    cu.x2 = d.x; cu.y2 = d.y; cu.x3 = e.x; cu.y3 = e.y;
    return cu;
}
__device__ __forceinline__ float ex(const Curve &c, float t) { return curve_eval(c.type, c.w, c.x0, c.x1, c.x2, c.x3, t); }
__device__ __forceinline__ float ey(const Curve &c, float t) { return curve_eval(c.type, c.w, c.y0, c.y1, c.y2, c.y3, t); }

// Monotone split points of one coordinate (k_make_intersection_0, ras_cut.h:1314-1409)
__device__ __forceinline__ void monotone_axis(const Curve &c, float v0, float v1, float v2, float v3, float *tq, int &n) {
    if (c.type == CT_QUAD) {
        float b = (v0 - v1) + (v2 - v1);
        if (!(fabsf(b) < 1e-6f)) {
            float t = (v0 - v1) / b;
            if (t > 0.f && t < 1.f) { tq[n] = t; n++; }
        }
    } else if (c.type == CT_CUBIC) {
        float r0 = 0.f, r1 = 0.f;
        solve_quad((v3 - v0) + 3.f * (v1 - v2), 2.f * ((v0 - v1) + (v2 - v1)), v1 - v0, r0, r1);
        if (r0 > 0.f && r0 < 1.f) { tq[n] = r0; n++; }
        if (r1 > 0.f && r1 < 1.f && r1 != r0) { tq[n] = r1; n++; }
    } else if (c.type == CT_RATIONAL) {
        const float w0 = 1.f, w1 = c.w, w2 = 1.f;
        float p0 = w0 * v0, p1 = w1 * v1, p2 = w2 * v2;
        float pw01 = p0 * w1 - p1 * w0;
        float pw12 = p1 * w2 - p2 * w1;
        float pw20 = p2 * w0 - p0 * w2;
        float r0 = 0.f, r1 = 0.f;
        solve_quad(pw01 + pw12 + pw20, -(2.f * pw01 + pw20), pw01, r0, r1);
        if (r0 > 0.f && r0 < 1.f) { tq[n] = r0; n++; }
        if (r1 > 0.f && r1 < 1.f && r1 != r0) { tq[n] = r1; n++; }
    }
}

// d_get_xy_begin_end_delta + d_udpate_cut_range (ras_cut.h:693-777)
struct CutRange {
    int32_t xb, xe, yb, ye, nx, ny;
    float dx, dy;
};
__device__ __forceinline__ CutRange cut_range(float v0x, float v0y, float v1x, float v1y, int W, int H) {
    CutRange r;
    if (v0x <= v1x) { r.xb = wadd(floor_even(v0x), 2); r.xe = floor_even(v1x); r.dx = 2.f; }
    else { r.xb = wadd(floor_even(v1x), 2); r.xe = floor_even(v0x); r.dx = -2.f; }
    if (v0y <= v1y) { r.yb = wadd(floor_even(v0y), 2); r.ye = floor_even(v1y); r.dy = 2.f; }
    else { r.yb = wadd(floor_even(v1y), 2); r.ye = floor_even(v0y); r.dy = -2.f; }
    const int32_t xmax = (W & ~1) + 2, ymax = (H & ~1) + 2;
    if ((r.xb < 0 && r.xe < 0) || (r.xb > xmax && r.xe > xmax) || (r.xb > r.xe)) {
        r.nx = 0;
    } else {
        r.xb = tmin(tmax(r.xb, 0), xmax);
        r.xe = tmin(tmax(r.xe, 0), xmax);
        r.nx = tmax((r.xe - r.xb) / 2 + 1, 0);
    }
    if ((r.yb < 0 && r.ye < 0) || (r.yb > ymax && r.ye > ymax) || (r.yb > r.ye)) {
        r.ny = 0;
    } else {
        r.yb = tmin(tmax(r.yb, 0), ymax);
        r.ye = tmin(tmax(r.ye, 0), ymax);
        r.ny = tmax((r.ye - r.yb) / 2 + 1, 0);
    }
    return r;
}

__device__ __forceinline__ float pick5(int i, float a, float b, float c, float d, float e) {
    return i == 0 ? a : (i == 1 ? b : (i == 2 ? c : (i == 3 ? d : e)));
}

// k_make_intersection_0 (ras_cut.h:1254-1482): monotone cuts, then per monotone piece the
// numbers of even x / y cut lines it crosses (d_udpate_cut_range) and the curve's fragment
// count.  Also the number of solve chains (two per piece) of the intersection pass.
__device__ __forceinline__ int wave_min_i(int v) { return wave_reduce(v, INT_MAX, OpMin()); }
__device__ __forceinline__ int wave_max_i(int v) { return wave_reduce(v, INT_MIN, OpMax()); }

// Range of cell rows a visible curve's fragments can fall in: a fragment's row comes from
// its chord midpoint, which lies inside the control-point hull (rationals with w > 0), so
// the hull's rows with one row of margin for rounding, clamped to the canvas rows.
__device__ __forceinline__ void curve_rows(const Curve &cu, int H, int &rlo, int &rhi) {
    const int nv = cu.type & 7;
    float lo = fminf(cu.y0, cu.y1), hi = fmaxf(cu.y0, cu.y1);
    if (nv > 2) { lo = fminf(lo, cu.y2); hi = fmaxf(hi, cu.y2); }
    if (nv > 3) { lo = fminf(lo, cu.y3); hi = fmaxf(hi, cu.y3); }
    const int rmax = (H - 1) >> 1;
    const bool hull = cu.type != CT_RATIONAL || cu.w > 0.f;
    if (hull && fabsf(lo) < 1e30f && fabsf(hi) < 1e30f) {
        rlo = tmax((floor_i32(lo) >> 1) - 1, 0);
        rhi = tmin((floor_i32(hi) >> 1) + 1, rmax);
    } else {
        rlo = 0;
        rhi = rmax;
    }
}

__global__ void __launch_bounds__(256) k_monotone(SceneDev sc, const float2 *__restrict__ tv, const uint32_t *__restrict__ pflags,
                                                   int W, int H, float4 *crec, uint32_t *cfrag,
                                                   uint32_t *pcnt, uint32_t *nchain, int2 *prow) {
    // the block's 64-byte curve records, staged so that they are stored with consecutive lanes
    __shared__ float4 stg[256 * 4];
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = c < sc.C;
    bool vis = false;
    int rlo = 0, rhi = -1;
    uint32_t path = 0u;
    if (live) {
        Curve cu = load_curve(sc, tv, c);
        path = sc.cpath[c];
        vis = region_visible(pflags[path]);
        path_canvas(sc, path, W, H);
        float tq[4] = {0.f, 0.f, 0.f, 0.f};
        int n = 0;
        if (vis) {
            monotone_axis(cu, cu.x0, cu.x1, cu.x2, cu.x3, tq, n);
            monotone_axis(cu, cu.y0, cu.y1, cu.y2, cu.y3, tq, n);
        }
        float q0 = tq[0], q1 = tq[1], q2 = tq[2], q3 = tq[3];
        if (n >= 2) { if (q1 < q0) { float t = q0; q0 = q1; q1 = t; } }
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
        stg[4 * threadIdx.x] = make_float4(cu.x0, cu.y0, cu.x1, cu.y1);
        stg[4 * threadIdx.x + 1] = make_float4(cu.x2, cu.y2, cu.x3, cu.y3);
        stg[4 * threadIdx.x + 2] = make_float4(q0, q1, q2, q3);
        stg[4 * threadIdx.x + 3] = make_float4(bf((uint32_t)cu.type | ((uint32_t)n << 8)), cu.w, bf(path), 0.f);
        uint32_t cnt = 0;
        uint32_t pc[5] = {0u, 0u, 0u, 0u, 0u};
        const int npieces = vis ? n + 1 : 0;
        if (vis) {
            float v0x = cu.x0, v0y = cu.y0;
            for (int ms = 0; ms < npieces; ++ms) {
                float t1 = (ms == n) ? 1.f : pick5(ms, q0, q1, q2, q3, 1.f);
                float v1x = ex(cu, t1), v1y = ey(cu, t1);
                CutRange r = cut_range(v0x, v0y, v1x, v1y, W, H);
                cnt += 1u + (uint32_t)r.nx + (uint32_t)r.ny;
                pc[ms] = (uint32_t)r.nx | ((uint32_t)r.ny << 16);
                v0x = v1x; v0y = v1y;
            }
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) pcnt[5 * (size_t)c + k] = pc[k];
        cfrag[c] = cnt;
        nchain[c] = 2u * (uint32_t)npieces;
        if (vis) curve_rows(cu, H, rlo, rhi);
    }
    __syncthreads();
    const uint32_t c0 = blockIdx.x * blockDim.x;
    const uint32_t nrec = 4u * (sc.C - c0 < blockDim.x ? sc.C - c0 : blockDim.x);
    for (uint32_t k = threadIdx.x; k < nrec; k += blockDim.x) crec[4 * (size_t)c0 + k] = stg[k];
    // the path's row range, one atomic pair per distinct path of the wave
    const int lane = threadIdx.x & 63;
    uint64_t pend = __ballot(vis && rlo <= rhi);
    while (pend) {
        const int leader = __ffsll((long long)pend) - 1;
        const uint32_t lp = (uint32_t)__builtin_amdgcn_readlane((int)path, leader);
        const bool mine = ((pend >> lane) & 1ull) && path == lp;
        const uint64_t grp = __ballot(mine);
        const int mn = wave_min_i(mine ? rlo : INT_MAX), mx = wave_max_i(mine ? rhi : INT_MIN);
        if (lane == leader) {
            atomicMin(&prow[lp].x, mn);
            atomicMax(&prow[lp].y, mx);
        }
        pend &= ~grp;
    }
}

// Work lists.  Pieces (curve << 3 | piece) with nx + ny + 1 >= 128 cuts are "long": their
// two solve chains (curve << 4 | piece << 1 | side, non-empty only) are listed for one
// thread each, and the piece for the wave-parallel merge.  Shorter pieces are listed for
// one thread that solves and walks both sides.  Chains and short pieces are grouped by
// (cubic or closed-form solve, length bin), longest first, so the lanes of a wave run
// loops of the same kind and about the same trip count and the long sequential walks
// start early.  A count pass sizes the classes; the place pass puts every entry at its
// class offset plus a rank taken per block in LDS and one reservation per class and block.
constexpr int kLenBins = 25;
constexpr int kChainClasses = 2 * kLenBins;                     // cubic bins, then closed-form bins
constexpr int kLongPieceClass = kChainClasses;                  // one class
constexpr int kShortClass0 = kChainClasses + 1;                 // cubic bins, then closed-form bins
constexpr int kListClasses = kShortClass0 + 2 * kLenBins;
constexpr uint32_t kLongPiece = 128;
constexpr int kChainListThreads = 1024;
static_assert(2 * kListClasses * sizeof(uint32_t) <= kListCountBytes, "class counts and cursors exceed their buffer");

// Length bins, longest first.  Below 256 a bin's longest entry is at most 4/3 of its
// shortest, so that the lanes of a wave run loops of nearly equal trip counts.  The bin
// of L is the number of lower bounds above L.
__device__ __forceinline__ int len_bin(uint32_t L) {
    constexpr uint32_t lower[kLenBins] = {384, 256, 192, 160, 128, 112, 96, 80, 65, 57, 49, 41, 33,
                                          29,  25,  21,  17,  13,  9,   7,  5,  4,  3,  2,  1};
    int b = 0;
#pragma unroll
    for (int i = 0; i < kLenBins; ++i) b += lower[i] > L ? 1 : 0;
    return b;
}

struct CurvePieces {
    int np;
    int cb;  // 0 for a cubic, kLenBins otherwise
    uint32_t pc[5];
};

__device__ __forceinline__ CurvePieces curve_pieces(const SceneDev &sc, uint32_t C, uint32_t c,
                                                    const uint32_t *__restrict__ nchain,
                                                    const uint32_t *__restrict__ pcnt) {
    CurvePieces r;
    const bool live = c < C;
    r.np = live ? (int)(nchain[c] >> 1) : 0;
    r.cb = (live && sc.ctype[c] == CT_CUBIC) ? 0 : kLenBins;
#pragma unroll
    for (int k = 0; k < 5; ++k) r.pc[k] = k < r.np ? pcnt[5 * (size_t)c + k] : 0u;
    return r;
}

// the classes of piece k's entries: x chain, y chain (-1 when not listed), piece
__device__ __forceinline__ void piece_classes(const CurvePieces &cp, int k, int &qx, int &qy, int &qp) {
    const uint32_t nx = cp.pc[k] & 0xFFFFu, ny = cp.pc[k] >> 16, L = nx + ny + 1;
    if (L >= kLongPiece) {
        qx = nx ? cp.cb + len_bin(nx) : -1;
        qy = ny ? cp.cb + len_bin(ny) : -1;
        qp = kLongPieceClass;
    } else {
        qx = qy = -1;
        qp = kShortClass0 + cp.cb + len_bin(L);
    }
}

// class counts (cnt[kListClasses])
__global__ void __launch_bounds__(kChainListThreads) k_list_count(SceneDev sc, uint32_t C,
                                                                  const uint32_t *__restrict__ nchain,
                                                                  const uint32_t *__restrict__ pcnt, uint32_t *cnt) {
    __shared__ uint32_t h[kListClasses];
    for (int q = threadIdx.x; q < kListClasses; q += blockDim.x) h[q] = 0u;
    __syncthreads();
    const CurvePieces cp = curve_pieces(sc, C, blockIdx.x * blockDim.x + threadIdx.x, nchain, pcnt);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        if (k >= cp.np) break;
        int qx, qy, qp;
        piece_classes(cp, k, qx, qy, qp);
        if (qx >= 0) atomicAdd(&h[qx], 1u);
        if (qy >= 0) atomicAdd(&h[qy], 1u);
        atomicAdd(&h[qp], 1u);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < kListClasses; q += blockDim.x)
        if (h[q]) atomicAdd(&cnt[q], h[q]);
}

// lists from the counts; cursor[kListClasses] starts at zero
__global__ void __launch_bounds__(kChainListThreads) k_list_place(SceneDev sc, uint32_t C,
                                                                  const uint32_t *__restrict__ nchain,
                                                                  const uint32_t *__restrict__ pcnt,
                                                                  const uint32_t *__restrict__ coff,
                                                                  const uint32_t *__restrict__ cnt, uint32_t *cursor,
                                                                  uint4 *clist, uint4 *plist) {
    __shared__ uint32_t h[kListClasses], base[kListClasses];
    for (int q = threadIdx.x; q < kListClasses; q += blockDim.x) h[q] = 0u;
    if (threadIdx.x == 0) {
        // class offsets: within the chain list, and within the piece list (long pieces first)
        uint32_t a = 0u;
        for (int q = 0; q < kChainClasses; ++q) { base[q] = a; a += cnt[q]; }
        a = 0u;
        for (int q = kLongPieceClass; q < kListClasses; ++q) { base[q] = a; a += cnt[q]; }
    }
    __syncthreads();
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    const CurvePieces cp = curve_pieces(sc, C, c, nchain, pcnt);
    uint32_t rk[15];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        if (k >= cp.np) break;
        int qx, qy, qp;
        piece_classes(cp, k, qx, qy, qp);
        rk[3 * k] = qx >= 0 ? atomicAdd(&h[qx], 1u) : 0u;
        rk[3 * k + 1] = qy >= 0 ? atomicAdd(&h[qy], 1u) : 0u;
        rk[3 * k + 2] = atomicAdd(&h[qp], 1u);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < kListClasses; q += blockDim.x)
        if (h[q]) base[q] += atomicAdd(&cursor[q], h[q]);
    __syncthreads();
    uint32_t off = cp.np ? coff[c] : 0u;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        if (k >= cp.np) break;
        int qx, qy, qp;
        piece_classes(cp, k, qx, qy, qp);
        // each entry carries its piece's ts offset and cut counts, so that the intersection
        // passes read only the curve's record besides it
        const uint32_t pb = off, pw = cp.pc[k];
        off += 1u + (pw & 0xFFFFu) + (pw >> 16);
        if (qx >= 0) clist[base[qx] + rk[3 * k]] = make_uint4((c << 4) | ((uint32_t)k << 1), pb, pw, 0u);
        if (qy >= 0) clist[base[qy] + rk[3 * k + 1]] = make_uint4((c << 4) | ((uint32_t)k << 1) | 1u, pb, pw, 0u);
        plist[base[qp] + rk[3 * k + 2]] = make_uint4((c << 3) | (uint32_t)k, pb, pw, 0u);
    }
}

// total entries of classes [q0, q1)
__device__ __forceinline__ uint32_t class_total(const uint32_t *__restrict__ cnt, int q0, int q1) {
    uint32_t t = 0u;
    for (int q = q0; q < q1; ++q) t += cnt[q];
    return t;
}

// t | flags of a monotone piece's end: low bits 2 when the end lies on an integral x,
// else 3 (k_make_intersection_impl_1, ras_cut.h:858-865)
__device__ __forceinline__ float piece_end_flags(float t_raw, float vx) {
    float t = bf(fb(t_raw) & 0xFFFFFFFCu);
    if (floorf(vx) == vx) return bf((fb(t) & 0xFFFFFFFCu) | 2u);
    return bf(fb(t) | 3u);
}

// One cut solve on one coordinate of the curve (ras_cut.h:1004-1106): line and
// quadratic / rational in closed form, cubic by 24-step bisection from t_min.
__device__ __forceinline__ float solve_cut(int type, float w, float a0, float a1, float a2, float a3, float cc,
                                           float t_min, float t1) {
    if (type == CT_LINEAR) return tmin(tmax((cc - a0) * safe_rcp(a1 - a0), t_min), t1);
    if (type == CT_QUAD) return solve_quad_bezier(a0, a1, a2, cc, t_min, t1);
    if (type == CT_RATIONAL) return solve_quad_bezier(a0 - cc, (a1 - cc) * w, a2 - cc, 0.f, t_min, t1);
    // 24 bisection steps.  lo moves only to a midpoint whose f - cc has the sign bit of
    // f(lo) - cc, so that sign bit stays the one of f(t_min) - cc.
    float lo = t_min, hi = t1;
    const float vlo = curve_eval(CT_CUBIC, 0.f, a0, a1, a2, a3, lo);
    float t_solve = lo;
    if (vlo != cc) {
        const uint32_t s0 = fb(vlo - cc) >> 31;
        float last_v = 0.f;
        for (int it = 0; it < 24; ++it) {
            const float tm = (lo + hi) * .5f;
            const float vm = curve_eval(CT_CUBIC, 0.f, a0, a1, a2, a3, tm);
            const bool up = (fb(vm - cc) >> 31) == s0;
            lo = up ? tm : lo;
            hi = up ? hi : tm;
            t_solve = tm;
            last_v = vm;
        }
        if (fabsf(last_v - cc) > 1.f) t_solve = t_min;
    }
    return t_solve;
}

// The cut solve state of one side of a monotone piece.  In the reference's merge walk
// (ras_cut.h:910-1232) the j-th cut on one side is solved from the t the same side
// produced last (or the piece start), so each side's chain is independent of the other
// side's and is solved in order, each t stored as t | side.
struct SideSolver {
    float a0, a1, a2, a3, cc, dd, t_min, t1m, w;
    int type;
    uint32_t side;
    __device__ __forceinline__ float next() {
        const float t = solve_cut(type, w, a0, a1, a2, a3, cc, t_min, t1m);
        const float tm = bf((fb(t) & 0xFFFFFFFCu) | side);
        t_min = tm;
        cc += dd;
        return tm;
    }
};

struct PieceSetup {
    uint32_t c, base, nx, ny;
    float t0m;
    SideSolver sx, sy;
};

// piece k of curve c: its ts offset, cut counts, start flags and both sides' solvers
// (e: the work-list entry {curve and piece, ts offset of the piece, nx | ny << 16})
__device__ __forceinline__ PieceSetup piece_setup(const SceneDev &sc, const float4 *__restrict__ crec,
                                                  uint32_t c, int k, uint4 e, int W, int H) {
    PieceSetup p;
    p.c = c;
    p.base = e.y;
    p.nx = e.z & 0xFFFFu;
    p.ny = e.z >> 16;
    const CurveRec rec = load_rec(crec, c);
    path_canvas(sc, rec.path, W, H);
    const Curve &cu = rec.cu;
    const float4 q = rec.q;
    const int n = rec.n;
    float v0x = cu.x0, v0y = cu.y0, t0m = 0.f;
    if (k > 0) {
        const float tr = pick5(k - 1, q.x, q.y, q.z, q.w, 1.f);
        v0x = ex(cu, tr);
        v0y = ey(cu, tr);
        t0m = piece_end_flags(tr, v0x);
    }
    p.t0m = t0m;
    const float t1r = (k == n) ? 1.f : pick5(k, q.x, q.y, q.z, q.w, 1.f);
    const float v1x = ex(cu, t1r), v1y = ey(cu, t1r);
    const float t1m = piece_end_flags(t1r, v1x);
    const CutRange r = cut_range(v0x, v0y, v1x, v1y, W, H);
    p.sx.a0 = cu.x0; p.sx.a1 = cu.x1; p.sx.a2 = cu.x2; p.sx.a3 = cu.x3;
    p.sx.cc = (float)(r.dx < 0.f ? r.xe : r.xb);
    p.sx.dd = r.dx;
    p.sy.a0 = cu.y0; p.sy.a1 = cu.y1; p.sy.a2 = cu.y2; p.sy.a3 = cu.y3;
    p.sy.cc = (float)(r.dy < 0.f ? r.ye : r.yb);
    p.sy.dd = r.dy;
    p.sx.t_min = p.sy.t_min = t0m;
    p.sx.t1m = p.sy.t1m = t1m;
    p.sx.w = p.sy.w = cu.w;
    p.sx.type = p.sy.type = cu.type;
    p.sx.side = 0u;
    p.sy.side = 1u;
    return p;
}

// Intersection pass 1, long pieces: one thread per (curve, piece, side) chain, the t's
// stored for the merge pass (k_merge_long) with 16-byte stores.
__device__ __forceinline__ void solve_chain(const SceneDev &sc, const float4 *__restrict__ crec, uint4 e, int W,
                                            int H, float *chain_t) {
    const uint32_t d = e.x;
    const int side = (int)(d & 1u);
    PieceSetup p = piece_setup(sc, crec, d >> 4, (int)((d >> 1) & 7u), e, W, H);
    const uint32_t cnt = side ? p.ny : p.nx;
    SideSolver sv = side ? p.sy : p.sx;
    float *out = chain_t + p.base + (side ? p.nx : 0u);
    // scalar stores up to a 16-byte boundary, then whole float4 stores, then the tail
    const uint32_t head = (uint32_t)((4u - (((uintptr_t)out >> 2) & 3u)) & 3u);
    float4 q4 = make_float4(0.f, 0.f, 0.f, 0.f);
    for (uint32_t j = 0; j < cnt; ++j) {
        const float tm = sv.next();
        if (j < head) {
            out[j] = tm;
        } else {
            const uint32_t ph = (j - head) & 3u;
            if (ph == 0) q4.x = tm;
            else if (ph == 1) q4.y = tm;
            else if (ph == 2) q4.z = tm;
            else q4.w = tm;
            if (ph == 3) *(float4 *)(out + j - 3) = q4;
        }
    }
    if (cnt > head) {
        const uint32_t rem = (cnt - head) & 3u;  // values still held in q4
        const uint32_t j0 = cnt - rem;
        if (rem > 0) out[j0] = q4.x;
        if (rem > 1) out[j0 + 1] = q4.y;
        if (rem > 2) out[j0 + 2] = q4.z;
    }
}

// Intersection pass 2 replays the reference's slot walk (ras_cut.h:910-1000) per
// monotone piece over the piece's solved chains: pick the x slot when x <= y, emit its t,
// refill it from the chain or with the exhausted marker 2 | side; equal consecutive t's
// OR their flags into the previous entry.  A piece starts at the previous piece's end
// flags (t = 0 for the first), exactly as the per-curve walk carries them.
struct PieceRef {
    uint32_t c, base, nx, ny;
    float t0m;
};

__device__ __forceinline__ PieceRef piece_ref(const float4 *__restrict__ crec, uint4 e) {
    PieceRef r;
    const uint32_t d = e.x;
    r.c = d >> 3;
    const int k = (int)(d & 7u);
    r.base = e.y;
    r.nx = e.z & 0xFFFFu;
    r.ny = e.z >> 16;
    r.t0m = 0.f;
    if (k > 0) {
        const CurveRec rec = load_rec(crec, r.c);
        const float tr = pick5(k - 1, rec.q.x, rec.q.y, rec.q.z, rec.q.w, 1.f);
        r.t0m = piece_end_flags(tr, ex(rec.cu, tr));
    }
    return r;
}

// Sequential reader of a chain: one 16-byte load per aligned group of four values (the
// chain buffer is 16-byte aligned and padded, so whole groups stay inside it).
struct ChainReader {
    const float *p;
    const float4 *blk;
    float4 buf;
    __device__ __forceinline__ void init(const float *q) {
        p = q;
        blk = nullptr;
    }
    __device__ __forceinline__ float at(uint32_t i) {
        const uintptr_t a = (uintptr_t)(p + i);
        const float4 *b = (const float4 *)(a & ~(uintptr_t)15);
        if (b != blk) {
            blk = b;
            buf = *b;
        }
        const uint32_t c = (uint32_t)(a >> 2) & 3u;
        return c == 0 ? buf.x : (c == 1 ? buf.y : (c == 2 ? buf.z : buf.w));
    }
};

__device__ __forceinline__ void walk_piece(const PieceRef &r, const float *__restrict__ chain_t, int2 *ts) {
    int2 *out = ts + r.base;
    const uint32_t nx = r.nx, ny = r.ny;
    ChainReader X, Y;
    X.init(chain_t + r.base);
    Y.init(chain_t + r.base + nx);
    const float t0m = r.t0m;
    const float two = 2.f, two1 = bf(fb(2.f) | 1u);
    float xs, ys;
    uint32_t xi = 0, yi = 0, rx, ry;
    if (nx == 0) { xs = two; ys = t0m; rx = 0; ry = ny; }
    else if (ny == 0) { xs = t0m; ys = two1; rx = nx; ry = 0; }
    else { xs = X.at(0); xi = 1; rx = nx - 1; ys = t0m; ry = ny; }
    uint32_t last = fb(-1.f);
    const uint32_t n_loop = nx + ny + 1;
    for (uint32_t it = 0; it < n_loop; ++it) {
        float v;
        if (xs <= ys) {
            v = xs;
            if (rx > 0) { xs = X.at(xi++); rx--; } else { xs = two; }
        } else {
            v = ys;
            if (ry > 0) { ys = Y.at(yi++); ry--; } else { ys = two1; }
        }
        uint32_t o = fb(v);
        if ((o & 0xFFFFFFFCu) == (last & 0xFFFFFFFCu)) {
            o |= last;
            out[it - 1].y = (int)o;
        }
        out[it] = make_int2((int)r.c, (int)o);
        last = o;
    }
}

// Short and medium pieces (nx + ny + 1 < 128): one thread solves both sides and replays
// the slot walk as it goes, each slot refilled by the next cut of its side -- the same
// sequence walk_piece reads from the stored chains, without storing them.  The output is
// staged: every lane walks kStageSteps entries of its piece into an LDS column, then the
// wave writes the chunk out 16 lanes to a piece, so each piece's run of up to 128 bytes
// goes out in one store instead of one 8-byte store a step (lines left partly written by
// 64 pieces at once were evicted and written back several times: 1.16 GB written a frame
// for 0.42 GB of ts on the 2048x2048 20k-path scene, 0.58 GB now).  A flag merge into an
// entry already written out (equal t's across a chunk boundary, rare) waits for the
// chunk's stores and then updates that entry in memory.
constexpr int kStageSteps = 16;

__device__ __forceinline__ void walk_pieces_staged(PieceSetup &p, bool live, int2 *ts, uint32_t *buf) {
    const int lane = threadIdx.x & 63;
    const uint32_t nx = live ? p.nx : 0u, ny = live ? p.ny : 0u;
    const uint32_t n_loop = live ? nx + ny + 1 : 0u;
    const uint32_t base = live ? p.base : 0u, curve = live ? p.c : 0u;
    const float two = 2.f, two1 = bf(fb(2.f) | 1u);
    float xs = two, ys = two1;
    uint32_t rx = 0, ry = 0;
    if (live) {
        if (nx == 0) { xs = two; ys = p.t0m; rx = 0; ry = ny; }
        else if (ny == 0) { xs = p.t0m; ys = two1; rx = nx; ry = 0; }
        else { xs = p.sx.next(); rx = nx - 1; ys = p.t0m; ry = ny; }
    }
    const uint32_t nmax = (uint32_t)wave_reduce((int)n_loop, 0, OpMaxU());
    uint32_t last = fb(-1.f);
    for (uint32_t it0 = 0; it0 < nmax; it0 += kStageSteps) {
#pragma unroll
        for (int j = 0; j < kStageSteps; ++j) {
            const uint32_t it = it0 + j;
            if (it < n_loop) {
                float v;
                if (xs <= ys) {
                    v = xs;
                    if (rx > 0) { xs = p.sx.next(); rx--; } else { xs = two; }
                } else {
                    v = ys;
                    if (ry > 0) { ys = p.sy.next(); ry--; } else { ys = two1; }
                }
                uint32_t o = fb(v);
                if ((o & 0xFFFFFFFCu) == (last & 0xFFFFFFFCu)) {
                    o |= last;
                    if (j > 0) {
                        buf[(j - 1) * 64 + lane] = o;
                    } else {
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // the chunk's stores first
                        ts[base + it - 1].y = (int)o;
                    }
                }
                buf[j * 64 + lane] = o;
                last = o;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
#pragma unroll 4
        for (int k = 0; k < 16; ++k) {
            const int L = 4 * k + (lane >> 4), j = lane & 15;
            const uint32_t nL = (uint32_t)__shfl((int)n_loop, L, 64);
            const uint32_t bL = (uint32_t)__shfl((int)base, L, 64), cL = (uint32_t)__shfl((int)curve, L, 64);
            if (it0 + j < nL) ts[bL + it0 + j] = make_int2((int)cL, (int)buf[j * 64 + L]);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    }
}

// Intersection pass 1.  Blocks [0, chain_blocks) solve the listed chains of long pieces,
// longest first; the blocks after them each take 256 short pieces and walk them while
// solving, so the short pieces fill the GPU while the long chains run.
__global__ void __launch_bounds__(256) k_chains(SceneDev sc, const float4 *__restrict__ crec, const uint4 *__restrict__ clist,
                                                 const uint4 *__restrict__ plist, const uint32_t *__restrict__ cnt,
                                                 uint32_t chain_blocks, int W, int H, float *chain_t, int2 *ts) {
    if (blockIdx.x < chain_blocks) {
        const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= class_total(cnt, 0, kChainClasses)) return;
        solve_chain(sc, crec, clist[i], W, H, chain_t);
        return;
    }
    __shared__ uint32_t stage[4][kStageSteps * 64];
    const uint32_t g = (blockIdx.x - chain_blocks) * blockDim.x + threadIdx.x;
    const bool live = g < class_total(cnt, kShortClass0, kListClasses);
    if (!__ballot(live)) return;  // whole wave past the list
    PieceSetup p;
    if (live) {
        const uint4 e = plist[cnt[kLongPieceClass] + g];
        p = piece_setup(sc, crec, e.x >> 3, (int)(e.x & 7u), e, W, H);
    }
    walk_pieces_staged(p, live, ts, stage[threadIdx.x >> 6]);
}

// Long pieces: one two-wave block per piece merges in parallel.  The x chain X and the y
// stream B = [piece start, Y...] are each sorted except that B's start may exceed Y[0]
// by flag bits alone; with B ordered by its running maximum the walk is a stable merge
// (x first on ties): X[i] lands at i + #{B keys < X[i]}, B[j] at j + #{X <= key(B[j])}.
// Runs of equal t then take the OR of their flags up to the next entry of the run.
constexpr int kMergeCap = 1024;
// two waves a piece: the 12 KB window holds a 64-thread block to about three waves per SIMD;
// 128 threads take the headline scene's merge from 258 to 198 us (256: 229 us)
constexpr uint32_t kMergeThreads = 128;

__global__ void __launch_bounds__(kMergeThreads) k_merge_long(const float4 *__restrict__ crec, const float *__restrict__ chain_t,
                                                   const uint4 *__restrict__ plist, const uint32_t *__restrict__ counts,
                                                   int2 *ts) {
    __shared__ float sa[kMergeCap], sb[kMergeCap];
    __shared__ uint32_t sm[kMergeCap];
    const int lane = threadIdx.x;
    const uint32_t n0 = counts[0];
    for (uint32_t g = blockIdx.x; g < n0; g += gridDim.x) {
        const PieceRef r = piece_ref(crec, plist[g]);
        const uint32_t nx = r.nx, ny = r.ny, n = nx + ny + 1;
        if (n > (uint32_t)kMergeCap) {
            // an over-long piece: the walk itself
            if (lane == 0) walk_piece(r, chain_t, ts);
            continue;
        }
        const float *X = chain_t + r.base;
        const float *Y = X + nx;
        if (nx == 0 || ny == 0) {
            // one stream: the piece start, then that chain in order
            const float *Z = nx == 0 ? Y : X;
            for (uint32_t p = lane; p < n; p += kMergeThreads) sm[p] = p == 0 ? fb(r.t0m) : fb(Z[p - 1]);
        } else {
        for (uint32_t i = lane; i < nx; i += kMergeThreads) sa[i] = X[i];
        for (uint32_t j = lane; j <= ny; j += kMergeThreads) sb[j] = j == 0 ? r.t0m : Y[j - 1];
        __syncthreads();
        // running-maximum key of B[j]: max(t0m, Y[j-1]) for j >= 1 (Y is sorted)
        auto bkey = [&](uint32_t j) { return j == 0 ? sb[0] : fmaxf(sb[0], sb[j]); };
        // merge path: each lane merges one contiguous run of outputs [p0, p1), starting
        // from the number of X entries among the first p0 outputs (the first i with
        // X[i] > key(B[p0 - i - 1]), found by binary search)
        const uint32_t nb = ny + 1, seg = (n + kMergeThreads - 1) / kMergeThreads;
        const uint32_t p0 = tmin(n, (uint32_t)lane * seg), p1 = tmin(n, p0 + seg);
        if (p0 < p1) {
            uint32_t lo = p0 > nb ? p0 - nb : 0u, hi = tmin(p0, nx);
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (sa[mid] <= bkey(p0 - mid - 1)) lo = mid + 1; else hi = mid;
            }
            uint32_t i = lo, j = p0 - lo;
            float xv = i < nx ? sa[i] : 0.f, bk = j < nb ? bkey(j) : 0.f;
            for (uint32_t p = p0; p < p1; ++p) {
                const bool takex = i < nx && (j >= nb || xv <= bk);
                if (takex) {
                    sm[p] = fb(xv);
                    ++i;
                    if (i < nx) xv = sa[i];
                } else {
                    sm[p] = fb(sb[j]);
                    ++j;
                    if (j < nb) bk = bkey(j);
                }
            }
        }
        }
        __syncthreads();
        int2 *out = ts + r.base;
        for (uint32_t p = lane; p < n; p += kMergeThreads) {
            const uint32_t v = sm[p], mk = v & 0xFFFFFFFCu;
            uint32_t acc = v;
            for (uint32_t q = p; q > 0 && (sm[q - 1] & 0xFFFFFFFCu) == mk; --q) acc |= sm[q - 1];
            if (p + 1 < n && (sm[p + 1] & 0xFFFFFFFCu) == mk) acc |= sm[p + 1];
            out[p] = make_int2((int)r.c, (int)acc);
        }
        __syncthreads();
    }
}

}  // namespace

size_t scan_tmp_elems_needed(size_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

void scan_exclusive_u32(const uint32_t *in, uint32_t *out, size_t n, uint32_t *tmp, hipStream_t s) {
    size_t ntiles = (n + kScanTile - 1) / kScanTile;
    if (ntiles == 0) {
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, s, tmp, (size_t)0, out + n);
        return;
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)ntiles), dim3(kScanThreads), 0, s, in, n, tmp);
    if (ntiles <= kScanDirectTiles) {
        hipLaunchKernelGGL(k_scan_down_direct, dim3((unsigned)ntiles), dim3(kScanThreads), 0, s, in, out, n,
                           (const uint32_t *)tmp);
        return;
    }
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, s, tmp, ntiles, out + n);
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)ntiles), dim3(kScanThreads), 0, s, in, out, n,
                       (const uint32_t *)tmp);
}



void launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, v, n);
}
void launch_fill2_u32(uint32_t *p, uint32_t *q, uint32_t v, size_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill2_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, q, v, n);
}
void launch_valu_probe(int blocks, int iters, float *out, hipStream_t s) {
    hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(256), 0, s, iters, 0.5f, out);
}

void launch_fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u8, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, v, n);
}

void launch_fill_rows_u8(uint8_t *p, uint8_t v, size_t row_bytes, size_t pitch, size_t rows, hipStream_t s) {
    const size_t n = row_bytes * rows;
    if (!n) return;
    hipLaunchKernelGGL(k_fill_rows_u8, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, v, row_bytes, pitch, n);
}

void launch_path_init(const SceneDev &sc, Frame &f, hipStream_t s) {
    if (!sc.P) return;
    hipLaunchKernelGGL(k_path_init, dim3(blocks_for(sc.P, 256)), dim3(256), 0, s, sc.P, f.pflags, f.prow);
}

void launch_transform(const SceneDev &sc, Frame &f, hipStream_t s) {
    if (!sc.V) return;
    Mat4 M;
    for (int i = 0; i < 16; ++i) M.m[i] = f.m[i];
    hipLaunchKernelGGL(k_transform, dim3(blocks_for(sc.V, 256)), dim3(256), 0, s, sc.V, sc.vtx, sc.vpath, sc.pcanvas, M,
                       f.W, f.H, f.tvtx, f.pflags);
}

void launch_monotone(const SceneDev &sc, Frame &f, hipStream_t s) {
    if (!sc.C) return;
    hipLaunchKernelGGL(k_monotone, dim3(blocks_for(sc.C, 256)), dim3(256), 0, s, sc, (const float2 *)f.tvtx,
                       (const uint32_t *)f.pflags, f.W, f.H, f.crec, f.cfrag, f.pcnt, f.nchain, f.prow);
}

void launch_intersect(const SceneDev &sc, Frame &f, hipStream_t s) {
    if (!sc.C) return;
    uint32_t *cnt = f.lcount, *cursor = f.lcount + kListClasses;
    launch_fill_u32(f.lcount, 0u, 2 * kListClasses, s);
    const dim3 gl(blocks_for(sc.C, kChainListThreads)), bl(kChainListThreads);
    hipLaunchKernelGGL(k_list_count, gl, bl, 0, s, sc, sc.C, (const uint32_t *)f.nchain, (const uint32_t *)f.pcnt, cnt);
    hipLaunchKernelGGL(k_list_place, gl, bl, 0, s, sc, sc.C, (const uint32_t *)f.nchain, (const uint32_t *)f.pcnt,
                       (const uint32_t *)f.coff, (const uint32_t *)cnt, cursor, f.chains, f.plist);
    // chain blocks: enough for every chain; piece blocks: enough for every piece
    const uint32_t npieces = (uint32_t)(f.NCH / 2);
    const uint32_t chain_blocks = blocks_for(f.NCH, 256), piece_blocks = blocks_for(npieces, 256);
    if (f.NCH)
        hipLaunchKernelGGL(k_chains, dim3(chain_blocks + piece_blocks), dim3(256), 0, s, sc, (const float4 *)f.crec,
                           (const uint4 *)f.chains, (const uint4 *)f.plist,
                           (const uint32_t *)cnt, chain_blocks, f.W, f.H, f.chain_t, f.ts);
    // one block per long piece up to 32768 blocks: the pieces' sizes vary widely, and
    // blocks that find no piece exit at once
    if (npieces)
        hipLaunchKernelGGL(k_merge_long, dim3(npieces < 32768u ? npieces : 32768u), dim3(kMergeThreads), 0, s,
                           (const float4 *)f.crec, (const float *)f.chain_t, (const uint4 *)f.plist,
                           (const uint32_t *)(cnt + kLongPieceClass), f.ts);
}

}  // namespace sl
