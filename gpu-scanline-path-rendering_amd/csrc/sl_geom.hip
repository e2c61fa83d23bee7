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

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u8(uint8_t *p, uint8_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

inline unsigned blocks_for(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

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
                            Mat4 M, int W, int H, float2 *__restrict__ out, uint32_t *pflags) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    float2 iv = vtx[i];
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
    // region bit per the 3x3 layout of ras_define.h:104-117; inside sets bits 0 and 7
    uint32_t bits;
    switch ((yf << 4) | xf) {
    case 0x00: bits = 1u << 0; break;
    case 0x01: bits = 1u << 1; break;
    case 0x02: bits = 1u << 2; break;
    case 0x10: bits = 1u << 3; break;
    case 0x11: bits = (1u << 0) | (1u << 7); break;
    case 0x12: bits = 1u << 4; break;
    case 0x20: bits = 1u << 5; break;
    case 0x21: bits = 1u << 6; break;
    default: bits = 1u << 7; break;
    }
    uint32_t p = vpath[i];
    if ((__hip_atomic_load(&pflags[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bits) != bits)
        atomicOr(&pflags[p], bits);
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

__device__ __forceinline__ Curve load_curve(const SceneDev &sc, const float2 *__restrict__ tv, uint32_t c) {
    Curve cu;
    cu.type = sc.ctype[c];
    cu.w = (cu.type == CT_RATIONAL) ? sc.cw[c] : 0.f;
    const uint32_t p = sc.cvpos[c];
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

// k_make_intersection_0 (ras_cut.h:1254-1482): cuts + fragment count per curve
__global__ void __launch_bounds__(256) k_monotone(SceneDev sc, const float2 *__restrict__ tv, const uint32_t *__restrict__ pflags,
                                                   int W, int H, float4 *ccuts, uint8_t *cncut, uint32_t *cfrag) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= sc.C) return;
    Curve cu = load_curve(sc, tv, c);
    const bool vis = region_visible(pflags[sc.cpath[c]]);
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
    ccuts[c] = make_float4(q0, q1, q2, q3);
    cncut[c] = (uint8_t)n;
    uint32_t cnt = 0;
    if (vis) {
        const int npieces = n + 1;
        float v0x = cu.x0, v0y = cu.y0;
        for (int ms = 0; ms < npieces; ++ms) {
            float t1 = (ms == n) ? 1.f : pick5(ms, q0, q1, q2, q3, 1.f);
            float v1x = ex(cu, t1), v1y = ey(cu, t1);
            CutRange r = cut_range(v0x, v0y, v1x, v1y, W, H);
            cnt += 1u + (uint32_t)r.nx + (uint32_t)r.ny;
            v0x = v1x; v0y = v1y;
        }
    }
    cfrag[c] = cnt;
}

// k_make_intersection_1 (ras_cut.h:837-1241, 1484-1562): one thread per curve writes
// the curve's {curve id, t | flags} entries.  While emitting, it evaluates each
// fragment's chord midpoint row (as the fragment kernel will) to record the per-path
// range of valid cell rows.
__global__ void __launch_bounds__(256) k_intersect(SceneDev sc, const float2 *__restrict__ tv,
                                                    const float4 *__restrict__ ccuts, const uint8_t *__restrict__ cncut,
                                                    const uint32_t *__restrict__ coff, int W, int H, int2 *ts, int2 *prow) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= sc.C) return;
    const uint32_t base = coff[c];
    const int64_t cap = (int64_t)(coff[c + 1] - base);
    if (cap == 0) return;
    Curve cu = load_curve(sc, tv, c);
    const float4 q = ccuts[c];
    const int n = cncut[c];
    int2 *out = ts + base;
    int rmin = INT_MAX, rmax = INT_MIN;
    bool have_prev = false;
    float prev_t = 0.f, prev_y = 0.f;
    auto emit_row = [&](float t) {
        // fragment (prev_t, t) exists iff prev_t < t; its cell row comes from the
        // midpoint of its chord (k_gen_fragment_and_stencil_mask, .cu:247-262)
        float y = ey(cu, t);
        if (have_prev && prev_t < t) {
            int32_t ry = floor_even((prev_y + y) * 0.5f);
            if ((uint32_t)ry < (uint32_t)H) {
                int r = ry >> 1;
                rmin = tmin(rmin, r);
                rmax = tmax(rmax, r);
            }
        }
        have_prev = true;
        prev_t = t;
        prev_y = y;
    };

    float t0_ms = 0.f;
    float v0x = cu.x0, v0y = cu.y0;
    int64_t p = 0;
    for (int ms = 0; ms <= n; ++ms) {
        float t1_ms = (ms == n) ? 1.f : pick5(ms, q.x, q.y, q.z, q.w, 1.f);
        float v1x = ex(cu, t1_ms), v1y = ey(cu, t1_ms);
        t1_ms = bf(fb(t1_ms) & 0xFFFFFFFCu);
        if (floorf(v1x) == v1x) t1_ms = bf((fb(t1_ms) & 0xFFFFFFFCu) | 2u);
        else t1_ms = bf(fb(t1_ms) | 3u);
        CutRange r = cut_range(v0x, v0y, v1x, v1y, W, H);
        int nx = r.nx, ny = r.ny;
        const int n_loop = nx + ny + 1;
        float x = (float)(r.dx < 0.f ? r.xe : r.xb);
        float y = (float)(r.dy < 0.f ? r.ye : r.yb);
        float slot_x = t0_ms, slot_y = t0_ms;
        uint32_t last = fb(-1.f);
        for (int i = -1; i < n_loop; ++i) {
            float t_solve = 0.f, cc = 0.f;
            int side = 0;
            float t_min;
            if (i == -1) {
                t_min = slot_x;
                if (nx == 0) { side = 0; t_solve = 2.f; }
                else if (ny == 0) { side = 1; t_solve = 2.f; }
                else { side = 0; nx--; cc = x; x += r.dx; }
            } else if (slot_x <= slot_y) {
                t_min = slot_x;
                side = 0;
                if (nx > 0) { nx--; cc = x; x += r.dx; } else { t_solve = 2.f; }
            } else {
                t_min = slot_y;
                side = 1;
                if (ny > 0) { ny--; cc = y; y += r.dy; } else { t_solve = 2.f; }
            }
            if (i >= 0 && p < cap) {
                uint32_t o = fb(t_min);
                if ((o & 0xFFFFFFFCu) == (last & 0xFFFFFFFCu)) {
                    o |= last;
                    out[p - 1].y = (int)o;
                }
                out[p] = make_int2((int)c, (int)o);
                last = o;
                p++;
                float tm = bf(o & 0xFFFFFFFCu);
                if (tm < 0.f) tm = 0.f;
                emit_row(tm);
            }
            if (t_solve < 2.f) {
                const float a0 = side ? cu.y0 : cu.x0, a1 = side ? cu.y1 : cu.x1;
                const float a2 = side ? cu.y2 : cu.x2, a3 = side ? cu.y3 : cu.x3;
                if (cu.type == CT_LINEAR) {
                    t_solve = tmin(tmax((cc - a0) * safe_rcp(a1 - a0), t_min), t1_ms);
                } else if (cu.type == CT_QUAD) {
                    t_solve = solve_quad_bezier(a0, a1, a2, cc, t_min, t1_ms);
                } else if (cu.type == CT_RATIONAL) {
                    t_solve = solve_quad_bezier(a0 - cc, (a1 - cc) * cu.w, a2 - cc, 0.f, t_min, t1_ms);
                } else {
                    float lo = t_min, hi = t1_ms;
                    float vlo = curve_eval(CT_CUBIC, 0.f, a0, a1, a2, a3, lo);
                    t_solve = lo;
                    if (vlo != cc) {
                        const float raw = lo;
                        float last_v = 0.f;
                        for (int it = 0; it < 24; ++it) {
                            float tm = (lo + hi) * .5f;
                            float vm = curve_eval(CT_CUBIC, 0.f, a0, a1, a2, a3, tm);
                            t_solve = tm;
                            last_v = vm;
                            if ((int32_t)(fb(vm - cc) ^ fb(vlo - cc)) >= 0) { lo = tm; vlo = vm; }
                            else { hi = tm; }
                        }
                        if (fabsf(last_v - cc) > 1.f) t_solve = raw;
                    }
                }
            }
            const float nv = bf((fb(t_solve) & 0xFFFFFFFCu) | (uint32_t)side);
            if (side) slot_y = nv; else slot_x = nv;
        }
        t0_ms = t1_ms;
        v0x = v1x; v0y = v1y;
    }
    emit_row(1.f);  // the curve's last fragment ends at t = 1 (.cu:148)
    if (rmin <= rmax) {
        const uint32_t pth = sc.cpath[c];
        atomicMin(&prow[pth].x, rmin);
        atomicMax(&prow[pth].y, rmax);
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
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, s, tmp, ntiles, out + n);
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)ntiles), dim3(kScanThreads), 0, s, in, out, n, (const uint32_t *)tmp);
}

void launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, v, n);
}
void launch_fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u8, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, v, n);
}

void launch_path_init(const SceneDev &sc, Frame &f, hipStream_t s) {
    if (!sc.P) return;
    hipLaunchKernelGGL(k_path_init, dim3(blocks_for(sc.P, 256)), dim3(256), 0, s, sc.P, f.pflags, f.prow);
}

void launch_transform(const SceneDev &sc, Frame &f, hipStream_t s) {
    if (!sc.V) return;
    Mat4 M;
    for (int i = 0; i < 16; ++i) M.m[i] = f.m[i];
    hipLaunchKernelGGL(k_transform, dim3(blocks_for(sc.V, 256)), dim3(256), 0, s, sc.V, sc.vtx, sc.vpath, M, f.W, f.H,
                       f.tvtx, f.pflags);
}

void launch_monotone(const SceneDev &sc, Frame &f, hipStream_t s) {
    if (!sc.C) return;
    hipLaunchKernelGGL(k_monotone, dim3(blocks_for(sc.C, 256)), dim3(256), 0, s, sc, (const float2 *)f.tvtx,
                       (const uint32_t *)f.pflags, f.W, f.H, f.ccuts, f.cncut, f.cfrag);
}

void launch_intersect(const SceneDev &sc, Frame &f, hipStream_t s) {
    if (!sc.C) return;
    hipLaunchKernelGGL(k_intersect, dim3(blocks_for(sc.C, 256)), dim3(256), 0, s, sc, (const float2 *)f.tvtx,
                       (const float4 *)f.ccuts, (const uint8_t *)f.cncut, (const uint32_t *)f.coff, f.W, f.H, f.ts, f.prow);
}

}  // namespace sl
