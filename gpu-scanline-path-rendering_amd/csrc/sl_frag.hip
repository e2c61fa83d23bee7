// sl_frag.hip -- fragment generation (chord, 2x2 cell, scan winding, sign and per-sample
// coverage mask, k_gen_fragment_and_stencil_mask in ras_cut_mask_comb_scanline.cu:112-441)
// and the (path, row) bucketing that replaces the reference's segmented merge sort.
//
// Bucket layout: path p owns buckets [pbase[p], pbase[p+1]) in the order the reference's
// signed int32 key (y+0x7FFF)<<16 | (x+0x7FFF) sorts a path's fragments: cell rows 1..rhi,
// then the out-of-canvas key 0xFFFEFFFE, then row 0.  A path's rows are the exact range
// of its fragment rows, recorded while intersecting.
#include "sl_device.h"
#include "sl_internal.h"

namespace sl {
using namespace dev;

namespace {

template <int SPP> struct Words { static constexpr int n = SPP == 8 ? 1 : (SPP == 16 ? 2 : 4); };

// fetchVerticalTable (ras_qm_mask.h:545-560); 16 spp: 32-bit column word, bit = row*16+s
template <int SPP>
__device__ __forceinline__ void vmask(float y, uint32_t *m) {
    if (SPP == 8) {
        int32_t r = trunc_i32(roundf(y * 16.f));
        uint32_t k = lsr32(0xffffu, wsub(16, r));
        k |= k << 16;
        m[0] = k; m[1] = 0; m[2] = 0; m[3] = 0;
    } else if (SPP == 16) {
        int32_t r = trunc_i32(roundf(y * 32.f));
        uint32_t k = lsr32(0xffffffffu, wsub(32, r));
        m[0] = k; m[1] = k; m[2] = 0; m[3] = 0;
    } else {
        int32_t r = trunc_i32(roundf(y * 64.f));
        uint64_t k = lsr64(~0ull, wsub(64, r));
        uint32_t lo = (uint32_t)k, hi = (uint32_t)(k >> 32);
        m[0] = lo; m[1] = hi; m[2] = lo; m[3] = hi;
    }
}

// fetchVerticalTable2 (ras_qm_mask.h:562-598); (int)(y*K + 0.5) is evaluated in double as
// the reference's double literal promotes it.
template <int SPP>
__device__ __forceinline__ void vmask2(float y0, float y1, uint32_t *m) {
    constexpr int K = SPP == 8 ? 16 : (SPP == 16 ? 32 : 64);
    int32_t s0 = tmax(wsub(K, trunc_i32((double)(y0 * (float)K) + 0.5)), 0);
    int32_t s1 = tmax(wsub(K, trunc_i32((double)(y1 * (float)K) + 0.5)), 0);
    if (SPP == 8) {
        uint32_t k = lsr32(0xffffu, s0) ^ lsr32(0xffffu, s1);
        k |= k << 16;
        m[0] = k; m[1] = 0; m[2] = 0; m[3] = 0;
    } else if (SPP == 16) {
        uint32_t k = lsr32(0xffffffffu, s0) ^ lsr32(0xffffffffu, s1);
        m[0] = k; m[1] = k; m[2] = 0; m[3] = 0;
    } else {
        uint64_t k = lsr64(~0ull, s0) ^ lsr64(~0ull, s1);
        uint32_t lo = (uint32_t)k, hi = (uint32_t)(k >> 32);
        m[0] = lo; m[1] = hi; m[2] = lo; m[3] = hi;
    }
}

// fetchPixelTable (ras_qm_mask.h:519-542): line -> packed sample mask via the 256x256 table
template <int SPP>
__device__ __forceinline__ void fetch_pixel(const uint32_t *__restrict__ tab, float p0x, float p0y, float p1x,
                                            float p1y, uint32_t *m) {
    constexpr int NW = Words<SPP>::n;
    if (p1y < p0y) {
        float t = p0x; p0x = p1x; p1x = t;
        t = p0y; p0y = p1y; p1y = t;
    }
    const float Dx = p1x - p0x, Dy = p1y - p0y;
    const float vx = Dy, vy = -Dx;
    const float il = 1.f / sqrtf(vx * vx + vy * vy);
    const float Nx = vx * il, Ny = vy * il;
    float C = Nx * p0x + Ny * p0y;
    C -= 0.5f * (Nx + Ny);
    const float Cs = bf((fb(C) & 0x80000000u) + 0x3f800000u);
    const float k = tmax(1.f - C * Cs * 0.5f, 0.f) * Cs;
    const float u = (k * Nx) * 0.5f + 0.5f;
    const float v = (k * Ny) * 0.5f + 0.5f;
    const int ix = texel(u), iy = texel(v);
    const uint32_t *t = tab + (size_t)(iy * 256 + ix) * NW;
    if (NW == 4) {
        uint4 q = *(const uint4 *)t;
        m[0] = q.x; m[1] = q.y; m[2] = q.z; m[3] = q.w;
    } else if (NW == 2) {
        uint2 q = *(const uint2 *)t;
        m[0] = q.x; m[1] = q.y; m[2] = 0; m[3] = 0;
    } else {
        m[0] = t[0]; m[1] = 0; m[2] = 0; m[3] = 0;
    }
}

struct Frag {
    int32_t yx;      // the reference's sort key (0xFFFEFFFE when out of canvas rows)
    uint32_t path;
    int wnc;         // winding number change, -1 / 0 / +1
    int swn;         // scan winding number at the cell's midline
    int32_t pos_x;   // cell x (even, clamped to [-2, (W&~1)+2])
    int32_t row;     // cell row (pos_y / 2) or -1
    uint32_t m[4];
};

// Inputs of one fragment, loaded in phases so that a thread can keep the loads of several
// fragments in flight at once: the two intersections, the curve's scalars, its vertices.
struct FragIn {
    int2 e0, e1;
    uint32_t path, vp;
    int type;
    float w;
    bool rev;
    float2 a, b, d, e;
};

__device__ __forceinline__ void frag_load_ts(const int2 *__restrict__ ts, int64_t F, int64_t i, FragIn &in) {
    in.e0 = ts[i];
    in.e1 = (i + 1 != F) ? ts[i + 1] : make_int2(-1, 0x3f800000);
}

__device__ __forceinline__ void frag_load_curve(const SceneDev &sc, FragIn &in) {
    const uint4 ci = sc.cinfo[(uint32_t)in.e0.x];
    in.vp = ci.x;
    in.path = ci.y;
    in.type = (int)(ci.z & 0xFFu);
    in.rev = (ci.z >> 8) != 0u;
    in.w = (in.type == CT_RATIONAL) ? bf(ci.w) : 0.f;
}

__device__ __forceinline__ void frag_load_verts(const float2 *__restrict__ tv, FragIn &in) {
    const int nv = in.type & 7;
    in.a = tv[in.vp];
    in.b = tv[in.vp + 1];
    in.d = nv > 2 ? tv[in.vp + 2] : make_float2(0.f, 0.f);
    in.e = nv > 3 ? tv[in.vp + 3] : make_float2(0.f, 0.f);
}

// What the mask phase needs after the geometry phase: chord ends in cell space, the
// vertical band and the left-edge mask.
struct FragMaskIn {
    float p0x, p0y, p1x, p1y;
    uint32_t lm[4];
    bool live;
};

template <int SPP, bool MASK>
__device__ __forceinline__ Frag frag_geometry(const SceneDev &sc, const FragIn &in, int W, int H, FragMaskIn &mi) {
    Frag o;
    const int2 e0 = in.e0, e1 = in.e1;
    int f0 = e0.y & 1, f1 = e1.y & 1;
    float t0 = bf((uint32_t)e0.y & 0xFFFFFFFCu);
    float t1 = bf((uint32_t)e1.y & 0xFFFFFFFCu);
    if (t0 < 0) t0 = 0.f;
    if (t1 < 0) t1 = 0.f;
    if (t0 == 0.f || t0 == 1.f) f0 = 1;
    if (t1 == 0.f || t1 == 1.f) f1 = 1;
    o.path = in.path;
    path_canvas(sc, o.path, W, H);
    o.yx = (int32_t)0xFFFEFFFE;
    o.row = -1;
    o.pos_x = 0;
    o.wnc = 0;
    o.swn = 0;
    o.m[0] = o.m[1] = o.m[2] = o.m[3] = 0;
    mi.live = false;
    if (e0.x != e1.x) t1 = 1.f;
    if (!(t0 < t1)) return o;

    const int type = in.type;
    const float w = in.w;
    const float2 a = in.a, b = in.b, d = in.d, e = in.e;
    // FRAG_MASK_SUBDIV = 1: t = (1 - fi) * t0 + fi * t1 for fi = 0, 1 (.cu:211-214)
    const float ta = (1.f - 0.f) * t0 + 0.f * t1;
    const float tb = (1.f - 1.f) * t0 + 1.f * t1;
    float p0x = curve_eval(type, w, a.x, b.x, d.x, e.x, ta);
    float p0y = curve_eval(type, w, a.y, b.y, d.y, e.y, ta);
    float p1x = curve_eval(type, w, a.x, b.x, d.x, e.x, tb);
    float p1y = curve_eval(type, w, a.y, b.y, d.y, e.y, tb);
    const bool rev = in.rev;
    const float vfx = rev ? p1x : p0x, vfy = rev ? p1y : p0y;
    const float vlx = rev ? p0x : p1x, vly = rev ? p0y : p1y;
    const int sf = rev ? f1 : f0, sl = rev ? f0 : f1;

    const int32_t rfx = floor_even((vfx + vlx) * 0.5f);
    const int32_t rfy = floor_even((vfy + vly) * 0.5f);
    const int32_t pos_x = tmin(tmax(rfx, -2), (int32_t)((W & ~1) + 2));
    const int32_t pos_y = tmin(tmax(rfy, -2), (int32_t)((H & ~1) + 2));
    o.pos_x = pos_x;
    if ((uint32_t)rfy < (uint32_t)H) {
        o.yx = (int32_t)(((uint32_t)(pos_y + 0x7FFF) << 16) | ((uint32_t)(pos_x + 0x7FFF) & 0xFFFFu));
        o.row = pos_y >> 1;
    }
    const float wn_y = (float)(pos_y + 1);
    int swn = 0;
    if (vfy == vly) swn = 0;
    else if (vfy <= wn_y && wn_y < vly) swn = -1;
    else if (vly <= wn_y && wn_y < vfy) swn = 1;
    o.swn = swn;
    int wnc;
    if (vfy < vly) wnc = -1;
    else if (vfy > vly) wnc = 1;
    else wnc = (vfx < vlx) ? 1 : -1;

    // left boundary (.cu:301-354)
    float ly = 0.f, ry = 0.f;
    int left_flag = 1;
    if (vfx < vlx) { ly = vfy; ry = vly; left_flag = sf; }
    else if (vfx > vlx) { ly = vly; ry = vfy; left_flag = sl; }
    else if (vfx < (float)(pos_x + 1)) {
        if (sf == 0) { ly = vfy; ry = vly; left_flag = 0; }
        else if (sl == 0) { ly = vly; ry = vfy; left_flag = 0; }
    }
    uint32_t lm[4] = {0u, 0u, 0u, 0u};
    if (left_flag == 0) {
        if (MASK) vmask<SPP>((ly - (float)pos_y) * .5f, lm);
        if (ly <= (float)(pos_y + 1)) {
            if (ly >= ry) wnc = -wnc;
        } else {
            if (ly < ry) wnc = -wnc;
#pragma unroll
            for (int k = 0; k < 4; ++k) lm[k] = ~lm[k];
        }
    }
    o.wnc = wnc;
    if (MASK) {
        const float bx = (float)pos_x, by = (float)pos_y;
        mi.p0x = (p0x - bx) * .5f; mi.p0y = (p0y - by) * .5f;
        mi.p1x = (p1x - bx) * .5f; mi.p1y = (p1y - by) * .5f;
#pragma unroll
        for (int k = 0; k < 4; ++k) mi.lm[k] = lm[k];
        mi.live = true;
    }
    return o;
}

// mask = pixel-table mask & vertical band ^ left-edge mask (.cu:330-371)
template <int SPP>
__device__ __forceinline__ void frag_mask(const uint32_t *__restrict__ tab, const FragMaskIn &mi, Frag &o) {
    if (!mi.live) return;
    uint32_t vm[4], tm[4];
    vmask2<SPP>(mi.p0y, mi.p1y, vm);
    fetch_pixel<SPP>(tab, mi.p0x, mi.p0y, mi.p1x, mi.p1y, tm);
#pragma unroll
    for (int k = 0; k < 4; ++k) o.m[k] = (tm[k] & vm[k]) ^ mi.lm[k];
}

template <int SPP, bool MASK>
__device__ __forceinline__ Frag make_fragment(const SceneDev &sc, const uint32_t *__restrict__ tab,
                                              const float2 *__restrict__ tv, const int2 *__restrict__ ts, int64_t F,
                                              int64_t i, int W, int H) {
    FragIn in;
    frag_load_ts(ts, F, i, in);
    frag_load_curve(sc, in);
    frag_load_verts(tv, in);
    FragMaskIn mi;
    Frag o = frag_geometry<SPP, MASK>(sc, in, W, H, mi);
    if (MASK) frag_mask<SPP>(tab, mi, o);
    return o;
}

// bucket of (path, row) inside the path's range; row < 0 is the out-of-canvas key
__device__ __forceinline__ int64_t bucket_of(int2 pr, uint32_t pbase, int32_t row, bool &in_range) {
    const int lo1 = tmax(pr.x, 1);
    const int nhi = (pr.y >= lo1) ? pr.y - lo1 + 1 : 0;
    in_range = true;
    if (row < 0) return (int64_t)pbase + nhi;
    if (row == 0) {
        in_range = (pr.x == 0);
        return (int64_t)pbase + nhi + 1;
    }
    in_range = (row >= lo1 && row <= pr.y);
    return (int64_t)pbase + (row - lo1);
}

__global__ void k_path_buckets(uint32_t P, const int2 *__restrict__ prow, uint32_t *pnb) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int2 pr = prow[p];
    const int lo1 = tmax(pr.x, 1);
    const int nhi = (pr.y >= lo1) ? pr.y - lo1 + 1 : 0;
    pnb[p] = (uint32_t)(nhi + 1 + (pr.x == 0 ? 1 : 0));
}

// per bucket: owning path and row; clears the per-bucket counters.  One wavefront per path,
// its lanes over the path's buckets (coalesced).
__global__ void __launch_bounds__(256) k_bucket_info(uint32_t P, const int2 *__restrict__ prow,
                                                     const uint32_t *__restrict__ pbase, int2 *binfo, uint32_t *bcount,
                                                     uint32_t *bwsum) {
    const uint32_t p = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (p >= P) return;
    const int2 pr = prow[p];
    const int lo1 = tmax(pr.x, 1);
    const int nhi = (pr.y >= lo1) ? pr.y - lo1 + 1 : 0;
    const uint32_t b0 = pbase[p], nb = pbase[p + 1] - b0;
    for (uint32_t j = lane; j < nb; j += 64) {
        const int row = (int)j < nhi ? lo1 + (int)j : ((int)j == nhi ? -1 : 0);
        binfo[b0 + j] = make_int2((int)p, row);
        bcount[b0 + j] = 0;
        bwsum[b0 + j] = 0;
    }
}

// Lanes of a wave that hit the same bucket share one atomic: the loop takes the first
// pending lane's bucket as the leader's, serves every lane with that bucket, and repeats
// (a wave of consecutive fragments walks a few curves, so it touches few buckets).
__device__ __forceinline__ uint64_t lanes_lt() {
    const int lane = threadIdx.x & 63;
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// Fragments per thread of the count and scatter passes.  A thread's fragments are 64
// apart (lane-major), so each of its K fragments is coalesced across the wave, and their
// loads are issued phase by phase to overlap the dependent memory round trips.
constexpr int kFragPerThread = 2;

template <int K>
__device__ __forceinline__ void load_fragments(const SceneDev &sc, const float2 *__restrict__ tv,
                                               const int2 *__restrict__ ts, int64_t F, int64_t i0,
                                               const int2 *__restrict__ prow, const uint32_t *__restrict__ pbase,
                                               FragIn *in, bool *live, int2 *pr, uint32_t *pb) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
        live[j] = i0 + 64 * j < F;
        if (live[j]) frag_load_ts(ts, F, i0 + 64 * j, in[j]);
    }
#pragma unroll
    for (int j = 0; j < K; ++j)
        if (live[j]) frag_load_curve(sc, in[j]);
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (live[j]) {
            frag_load_verts(tv, in[j]);
            pr[j] = prow[in[j].path];
            pb[j] = pbase[in[j].path];
        }
    }
}

// Count pass: every fragment that lands in a bucket (a valid row, or an out-of-canvas row
// whose scan winding still feeds the winding carry) takes its rank in the bucket from the
// bucket counter, which lets the scatter pass place records without atomics.  A block's
// fragments are consecutive, so they belong to a few neighbouring paths and their buckets
// form a short window of bucket ids: the block counts them in LDS (the LDS atomic returns
// the rank inside the block), then one thread per non-empty bucket of the window reserves
// the block's share with one returning device atomic.  Ranks inside a bucket are in no
// particular order; the resolve pass orders a bucket's records by cell and its results do
// not depend on the order of records within a cell.  Windows wider than the LDS array
// fall back to one atomic per distinct bucket of a wave.  The same window sums the
// buckets' scan windings (mod 2^32) for the winding carry.
constexpr uint32_t kCountWindow = 2048;

__device__ __forceinline__ uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
__device__ __forceinline__ uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) { return (uint32_t)wave_reduce((int)v, -1, OpMinU()); }
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) { return (uint32_t)wave_reduce((int)v, 0, OpMaxU()); }

template <int SPP>
__global__ void __launch_bounds__(256) k_frag_count(SceneDev sc, const uint32_t *__restrict__ tab,
                                                     const float2 *__restrict__ tv, const int2 *__restrict__ ts, int64_t F,
                                                     int W, int H, const int2 *__restrict__ prow,
                                                     const uint32_t *__restrict__ pbase, uint32_t *bcount, uint32_t *bwsum,
                                                     uint32_t *rank, uint32_t *err) {
    constexpr int K = kFragPerThread;
    __shared__ uint32_t s_cnt[kCountWindow], s_sw[kCountWindow];
    __shared__ uint32_t s_lo[4], s_hi[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = (int)(blockDim.x >> 6);
    const int64_t i0 = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64 * K + lane;
    FragIn in[K];
    bool live[K];
    int2 pr[K];
    uint32_t pb[K];
    load_fragments<K>(sc, tv, ts, F, i0, prow, pbase, in, live, pr, pb);
    bool act[K];
    uint32_t b[K];
    int sw[K];
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        act[j] = false;
        b[j] = 0;
        sw[j] = 0;
        if (live[j]) {
            FragMaskIn mi;
            const Frag f = frag_geometry<SPP, false>(sc, in[j], W, H, mi);
            bool ok;
            b[j] = (uint32_t)bucket_of(pr[j], pb[j], f.row, ok);
            if (!ok) atomicOr(err, 1u);
            act[j] = ok && (f.row >= 0 || f.swn != 0);
            sw[j] = act[j] ? (int)f.swn : 0;
            if (act[j]) {
                lo = umin(lo, b[j]);
                hi = umax(hi, b[j]);
            }
        }
    }
    lo = wave_min_u32(lo);
    hi = wave_max_u32(hi);
    if (lane == 0) {
        s_lo[wv] = lo;
        s_hi[wv] = hi;
    }
    __syncthreads();
    lo = s_lo[0];
    hi = s_hi[0];
    for (int k = 1; k < nwv; ++k) {
        lo = umin(lo, s_lo[k]);
        hi = umax(hi, s_hi[k]);
    }
    if (lo <= hi && hi - lo < kCountWindow) {
        const uint32_t span = hi - lo + 1;
        for (uint32_t i = threadIdx.x; i < span; i += blockDim.x) {
            s_cnt[i] = 0u;
            s_sw[i] = 0u;
        }
        __syncthreads();
        uint32_t lr[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            lr[j] = act[j] ? atomicAdd(&s_cnt[b[j] - lo], 1u) : 0u;
            if (sw[j]) atomicAdd(&s_sw[b[j] - lo], (uint32_t)sw[j]);
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < span; i += blockDim.x) {
            const uint32_t c = s_cnt[i], w = s_sw[i];
            if (c) s_cnt[i] = atomicAdd(&bcount[lo + i], c);
            if (w) atomicAdd(&bwsum[lo + i], w);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (live[j]) rank[i0 + 64 * j] = act[j] ? s_cnt[b[j] - lo] + lr[j] : 0xFFFFFFFFu;
        return;
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (sw[j]) atomicAdd(&bwsum[b[j]], (uint32_t)sw[j]);
        uint64_t pend = __ballot(act[j]);
        uint32_t base = 0, r = 0;
        int my_leader = 0;
        while (pend) {
            const int leader = __ffsll((long long)pend) - 1;
            const uint32_t lb = (uint32_t)__builtin_amdgcn_readlane((int)b[j], leader);
            const bool mine = ((pend >> lane) & 1ull) && b[j] == lb;
            const uint64_t grp = __ballot(mine);
            if (mine) {
                my_leader = leader;
                r = (uint32_t)__popcll(grp & lanes_lt());
            }
            if (lane == leader) base = atomicAdd(&bcount[lb], (uint32_t)__popcll(grp));
            pend &= ~grp;
        }
        base = (uint32_t)__shfl((int)base, my_leader, 64);
        if (live[j]) rank[i0 + 64 * j] = act[j] ? base + r : 0xFFFFFFFFu;
    }
}

// Scatter pass: records at boff[bucket] + rank.  A path's fragments are neighbours in the
// input and land in the same few buckets, so the XCD-contiguous block order keeps their
// partial-line writes in one L2.
template <int SPP>
__global__ void __launch_bounds__(256) k_frag_scatter(SceneDev sc, const uint32_t *__restrict__ tab,
                                                       const float2 *__restrict__ tv, const int2 *__restrict__ ts,
                                                       int64_t F, int W, int H, const int2 *__restrict__ prow,
                                                       const uint32_t *__restrict__ pbase,
                                                       const uint32_t *__restrict__ boff, const uint32_t *__restrict__ rank,
                                                       uint32_t *rec_meta, uint32_t *rec_mask) {
    constexpr int NW = Words<SPP>::n;
    constexpr int K = kFragPerThread;
    const int lane = threadIdx.x & 63;
    const int64_t i0 = ((int64_t)xcd_chunk(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64 * K + lane;
    FragIn in[K];
    bool live[K];
    int2 pr[K];
    uint32_t pb[K];
    load_fragments<K>(sc, tv, ts, F, i0, prow, pbase, in, live, pr, pb);
    uint32_t rk[K];
#pragma unroll
    for (int j = 0; j < K; ++j) rk[j] = live[j] ? rank[i0 + 64 * j] : 0xFFFFFFFFu;
    Frag f[K];
    FragMaskIn mi[K];
    uint32_t b[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        b[j] = 0;
        if (rk[j] != 0xFFFFFFFFu) {
            f[j] = frag_geometry<SPP, true>(sc, in[j], W, H, mi[j]);
            bool ok;
            b[j] = (uint32_t)bucket_of(pr[j], pb[j], f[j].row, ok);
            if (f[j].row < 0) mi[j].live = false;  // winding only: no mask needed
        }
    }
#pragma unroll
    for (int j = 0; j < K; ++j)
        if (rk[j] != 0xFFFFFFFFu) frag_mask<SPP>(tab, mi[j], f[j]);
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (rk[j] == 0xFFFFFFFFu) continue;
        const uint32_t slot = boff[b[j]] + rk[j];
        if (slot >= boff[b[j] + 1]) continue;
        const Frag &g = f[j];
        const uint32_t cx = (uint32_t)((g.pos_x + 2) >> 1);
        rec_meta[slot] = cx | (((uint32_t)g.swn & 0xFFu) << 16) | ((g.wnc < 0 ? 1u : 0u) << 24);
        if (NW == 4) {
            *(uint4 *)(rec_mask + (size_t)slot * 4) = make_uint4(g.m[0], g.m[1], g.m[2], g.m[3]);
        } else if (NW == 2) {
            *(uint2 *)(rec_mask + (size_t)slot * 2) = make_uint2(g.m[0], g.m[1]);
        } else {
            rec_mask[slot] = g.m[0];
        }
    }
}

template <int SPP>
__global__ void __launch_bounds__(256) k_debug_fragments(SceneDev sc, const uint32_t *__restrict__ tab,
                                                          const float2 *__restrict__ tv, const int2 *__restrict__ ts,
                                                          int64_t F, int W, int H, int32_t *yx, int32_t *pid, int8_t *wn,
                                                          uint32_t *mask4) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F) return;
    Frag f = make_fragment<SPP, true>(sc, tab, tv, ts, F, i, W, H);
    yx[i] = f.yx;
    pid[i] = (int32_t)(f.path | ((uint32_t)f.wnc << 30));
    wn[i] = (int8_t)f.swn;
    *(uint4 *)(mask4 + 4 * i) = make_uint4(f.m[0], f.m[1], f.m[2], f.m[3]);
}

inline unsigned blocks_for(int64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

void launch_path_buckets(const SceneDev &sc, Frame &f, hipStream_t s) {
    if (!sc.P) return;
    hipLaunchKernelGGL(k_path_buckets, dim3(blocks_for(sc.P, 256)), dim3(256), 0, s, sc.P, (const int2 *)f.prow, f.pnb);
}

void launch_bucket_info(const SceneDev &sc, Frame &f, hipStream_t s) {
    if (!sc.P) return;
    hipLaunchKernelGGL(k_bucket_info, dim3(blocks_for(sc.P, 4)), dim3(256), 0, s, sc.P, (const int2 *)f.prow,
                       (const uint32_t *)f.pbase, f.binfo, f.bcount, f.bwsum);
}

void launch_frag_count(const SceneDev &sc, const Tables &t, Frame &f, hipStream_t s) {
    if (!f.F) return;
    const unsigned fb = f.frag_block > 0 && f.frag_block <= 256 ? (unsigned)f.frag_block : 256u;
    dim3 g(blocks_for(f.F, fb * kFragPerThread)), b(fb);
#define SL_LAUNCH(N)                                                                                              \
    hipLaunchKernelGGL(k_frag_count<N>, g, b, 0, s, sc, t.pixel, (const float2 *)f.tvtx, (const int2 *)f.ts, f.F, \
                       f.W, f.H, (const int2 *)f.prow, (const uint32_t *)f.pbase, f.bcount, f.bwsum, f.frank, \
                       f.err_flag)
    if (t.spp == 8) SL_LAUNCH(8);
    else if (t.spp == 16) SL_LAUNCH(16);
    else SL_LAUNCH(32);
#undef SL_LAUNCH
}

void launch_frag_scatter(const SceneDev &sc, const Tables &t, Frame &f, hipStream_t s) {
    if (!f.F) return;
    const unsigned fb = f.frag_block > 0 && f.frag_block <= 256 ? (unsigned)f.frag_block : 256u;
    dim3 g(blocks_for(f.F, fb * kFragPerThread)), b(fb);
#define SL_LAUNCH(N)                                                                                                \
    hipLaunchKernelGGL(k_frag_scatter<N>, g, b, 0, s, sc, t.pixel, (const float2 *)f.tvtx, (const int2 *)f.ts, f.F, \
                       f.W, f.H, (const int2 *)f.prow, (const uint32_t *)f.pbase, (const uint32_t *)f.boff,         \
                       (const uint32_t *)f.frank, f.rec_meta, f.rec_mask)
    if (t.spp == 8) SL_LAUNCH(8);
    else if (t.spp == 16) SL_LAUNCH(16);
    else SL_LAUNCH(32);
#undef SL_LAUNCH
}

void launch_bucket_carry(Frame &f, hipStream_t s) {
    // winding carry of every bucket: the exclusive sum (mod 2^32) of the scan windings of
    // all records in the buckets before it, bucket order being the reference's fragment order
    scan_exclusive_u32(f.bwsum, f.bcarry, (size_t)f.NB, f.scan_tmp, s);
}

void launch_debug_fragments(const SceneDev &sc, const Tables &t, Frame &f, int32_t *yx, int32_t *pid, int8_t *wn,
                            uint32_t *mask4, hipStream_t s) {
    if (!f.F) return;
    dim3 g(blocks_for(f.F, 256)), b(256);
#define SL_LAUNCH(N)                                                                                                   \
    hipLaunchKernelGGL(k_debug_fragments<N>, g, b, 0, s, sc, t.pixel, (const float2 *)f.tvtx, (const int2 *)f.ts, f.F, \
                       f.W, f.H, yx, pid, wn, mask4)
    if (t.spp == 8) SL_LAUNCH(8);
    else if (t.spp == 16) SL_LAUNCH(16);
    else SL_LAUNCH(32);
#undef SL_LAUNCH
}

}  // namespace sl
