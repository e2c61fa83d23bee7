// sl_resolve.hip -- per-(path, row) resolve: order a bucket's fragments by cell x, run
// the winding scan, merge fragments of a cell into one per-sample stencil mask and
// emit interior spans.  Together these replace the reference's segmented sort
// (cuda_seg_sort_int_by_int), int8 winding scan (thrust_impl_scan.cu:16-19),
// k_mark_merged_fragment_and_span (.cu:462-535), k_gen_stencil_cuda_sm /
// k_gen_stencil_block_boundary (.cu:554-786) and k_gen_gl_merged_fragment_and_span
// (.cu:789-868).
//
// Item encoding (one u32 per item; a bucket's items are disjoint and stored in x order,
// i.e. the span that ends at a cell precedes the cell's merged fragment -- the reference's
// draw list emits the pair the other way round, which the draw-list dump restores):
//   bits 0-15  x of the first pixel
//   bits 16-31 span width in pixels, 0 for a merged 2x2 fragment (its mask is stored at
//              the same slot of `imask`).
#include "sl_device.h"
#include "sl_internal.h"

#include <algorithm>

namespace sl {
using namespace dev;

namespace {

template <int SPP> struct Words { static constexpr int n = SPP == 8 ? 1 : (SPP == 16 ? 2 : 4); };

__device__ __forceinline__ int wave_incl_scan_i(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}
__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = tmin(v, __shfl_xor(v, d, 64));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = tmax(v, __shfl_xor(v, d, 64));
    return v;
}

// fill-rule test of k_mark_merged_fragment_and_span (.cu:511)
__device__ __forceinline__ bool fill_test(int rule, int8_t w) {
    return (rule == 0 && w != 0) || (rule == 1 && (w & 1));
}

// Per-sample stencil of one merged fragment in the reference's 7-bit byte-lane counters:
// acc = init_pos - init_neg + sum(+/- ((mask >> b) & 0x01010101)) (mod 2^32), result bit =
// ((0x80808080 + acc) & 0x7f7f7f7f) tested per byte (k_gen_stencil_cuda_sm, .cu:665-716).
template <int NW>
struct Stencil {
    uint32_t acc[NW][8];
    __device__ __forceinline__ void init(int8_t w) {
        uint32_t ip = 0u, in = 0u;
        if (w < 0) in = 0x01010101u * (uint32_t)(-(int)w);
        else ip = 0x01010101u * (uint32_t)w;
        const uint32_t v = ip - in;
#pragma unroll
        for (int k = 0; k < NW; ++k)
#pragma unroll
            for (int b = 0; b < 8; ++b) acc[k][b] = v;
    }
    __device__ __forceinline__ void add(const uint32_t *m, bool neg) {
#pragma unroll
        for (int k = 0; k < NW; ++k)
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                uint32_t a = (m[k] >> b) & 0x01010101u;
                acc[k][b] = neg ? acc[k][b] - a : acc[k][b] + a;
            }
    }
    __device__ __forceinline__ void finish(int rule, uint32_t *out) const {
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            uint32_t ret = 0u;
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                uint32_t v = (0x80808080u + acc[k][b]) & 0x7f7f7f7fu;
                if (rule == 0) { v |= v >> 4; v |= v >> 2; v |= v >> 1; }
                ret |= (v & 0x01010101u) << b;
            }
            out[k] = ret;
        }
    }
};

template <int NW>
__device__ __forceinline__ void store_mask(uint32_t *imask, size_t slot, const uint32_t *m) {
    if (NW == 4) *(uint4 *)(imask + slot * 4) = make_uint4(m[0], m[1], m[2], m[3]);
    else if (NW == 2) *(uint2 *)(imask + slot * 2) = make_uint2(m[0], m[1]);
    else imask[slot] = m[0];
}
template <int NW>
__device__ __forceinline__ void load_mask(const uint32_t *p, size_t slot, uint32_t *m) {
    if (NW == 4) { uint4 q = *(const uint4 *)(p + slot * 4); m[0] = q.x; m[1] = q.y; m[2] = q.z; m[3] = q.w; }
    else if (NW == 2) { uint2 q = *(const uint2 *)(p + slot * 2); m[0] = q.x; m[1] = q.y; }
    else m[0] = p[slot];
}

struct ResolveArgs {
    int64_t NB;
    const int2 *binfo;
    const uint32_t *boff;
    const uint32_t *bcarry;
    const uint8_t *frule;
    const uint32_t *rec_meta;
    const uint32_t *rec_mask;
    uint32_t *scratch;
    int W;
    uint32_t *items;
    uint32_t *imask;
    uint32_t *bicnt;
    uint32_t *bmcnt;
    uint32_t *bext;
    uint32_t *big_list;
    uint32_t *big_count;
    uint32_t *err;
};

// Wave-wide inclusive scans / reductions with DPP row shifts and row broadcasts.
template <class Op>
__device__ __forceinline__ int dpp_scan(int x, int ident, Op op) {
    x = op(x, __builtin_amdgcn_update_dpp(ident, x, 0x111, 0xf, 0xf, false));  // row_shr:1
    x = op(x, __builtin_amdgcn_update_dpp(ident, x, 0x112, 0xf, 0xf, false));  // row_shr:2
    x = op(x, __builtin_amdgcn_update_dpp(ident, x, 0x114, 0xf, 0xf, false));  // row_shr:4
    x = op(x, __builtin_amdgcn_update_dpp(ident, x, 0x118, 0xf, 0xf, false));  // row_shr:8
    x = op(x, __builtin_amdgcn_update_dpp(ident, x, 0x142, 0xa, 0xf, false));  // row_bcast:15
    x = op(x, __builtin_amdgcn_update_dpp(ident, x, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return x;
}
struct OpAdd { __device__ int operator()(int a, int b) const { return a + b; } };
struct OpMin { __device__ int operator()(int a, int b) const { return a < b ? a : b; } };
struct OpMax { __device__ int operator()(int a, int b) const { return a > b ? a : b; } };
__device__ __forceinline__ int scan_add(int x) { return dpp_scan(x, 0, OpAdd()); }
__device__ __forceinline__ int lane63(int x) { return __builtin_amdgcn_readlane(x, 63); }

// Resolve one (path, row) bucket of n <= 64*S fragments with one wavefront: stage the
// records in LDS, rank them by (cell x, index), then walk the sorted order in chunks of
// 64 carrying the winding number and item count.  Per cell group, the first lane sums
// the members' masks into the per-sample stencil; spans open between groups whose
// winding passes the fill rule.
template <int SPP, int S>
__device__ __forceinline__ void resolve_bucket(const ResolveArgs &A, int64_t b, uint32_t base, int n, uint32_t *key,
                                               uint32_t *order, uint32_t *meta, uint32_t *msk) {
    constexpr int NW = Words<SPP>::n;
    const int lane = threadIdx.x & 63;
    uint32_t mykey[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int i = s * 64 + lane;
        mykey[s] = 0xFFFFFFFFu;
        if (i < n) {
            const uint32_t m = A.rec_meta[base + i];
            meta[i] = m;
            uint32_t mk[NW];
            load_mask<NW>(A.rec_mask, (size_t)base + i, mk);
#pragma unroll
            for (int k = 0; k < NW; ++k) msk[i * NW + k] = mk[k];
            mykey[s] = ((m & 0xFFFFu) << 10) | (uint32_t)i;
            key[i] = mykey[s];
        }
    }
    __syncthreads();
    int rk[S];
#pragma unroll
    for (int s = 0; s < S; ++s) rk[s] = 0;
    if (S == 1) {
        for (int j = 0; j < n; ++j) {
            const uint32_t kj = (uint32_t)__builtin_amdgcn_readlane((int)mykey[0], j);
            rk[0] += (kj < mykey[0]) ? 1 : 0;
        }
    } else {
        for (int j = 0; j < n; ++j) {
            const uint32_t kj = key[j];
#pragma unroll
            for (int s = 0; s < S; ++s) rk[s] += (kj < mykey[s]) ? 1 : 0;
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s)
        if (s * 64 + lane < n) order[rk[s]] = (uint32_t)(s * 64 + lane);
    __syncthreads();

    const int rule = A.frule[A.binfo[b].x];
    uint32_t wcarry = A.bcarry[b];
    int icarry = 0, nmerged = 0, xmin = INT_MAX, xmax = INT_MIN;
    const size_t ib = 2 * (size_t)base;
#pragma unroll
    for (int c = 0; c < S; ++c) {
        if (c * 64 >= n) break;
        const int p = c * 64 + lane;
        const bool v = p < n;
        const uint32_t i = v ? order[p] : 0u;
        const uint32_t m = v ? meta[i] : 0u;
        const int cx = (int)(m & 0xFFFFu);
        const int swn = v ? (int)(int8_t)(m >> 16) : 0;
        const int prev_cx = (v && p > 0) ? (int)(meta[order[p - 1]] & 0xFFFFu) : -1;
        const bool head = v && (p == 0 || cx != prev_cx);
        const int incl = scan_add(swn);
        const int8_t w_at = (int8_t)(int)(wcarry + (uint32_t)(incl - swn));
        wcarry += (uint32_t)lane63(incl);
        const int x = 2 * cx - 2;
        const bool merged = head && x >= 0 && x < A.W;
        uint32_t res[NW];
        if (merged) {
            Stencil<NW> st;
            st.init(w_at);
            for (int q = p; q < n; ++q) {
                const uint32_t iq = order[q];
                const uint32_t mq = meta[iq];
                if ((int)(mq & 0xFFFFu) != cx) break;
                st.add(msk + iq * NW, ((mq >> 24) & 1u) != 0u);
            }
            st.finish(rule, res);
        }
        const int prev_x = 2 * prev_cx - 2;
        const bool span = head && p > 0 && (prev_x + 2 < x) && fill_test(rule, w_at);
        const int cnt = (merged ? 1 : 0) + (span ? 1 : 0);
        const int ic = scan_add(cnt);
        const int idx = icarry + ic - cnt;
        icarry += lane63(ic);
        if (span) {
            const int x0 = prev_x + 2;
            A.items[ib + idx] = (uint32_t)x0 | ((uint32_t)(x - x0) << 16);
        }
        if (merged) {
            A.items[ib + idx + (span ? 1 : 0)] = (uint32_t)x;
            store_mask<NW>(A.imask, ib + idx + (span ? 1 : 0), res);
        }
        const int lo = span ? prev_x + 2 : (merged ? x : INT_MAX);
        const int hi = merged ? x + 2 : (span ? x : INT_MIN);
        xmin = tmin(xmin, lane63(dpp_scan(lo, INT_MAX, OpMin())));
        xmax = tmax(xmax, lane63(dpp_scan(hi, INT_MIN, OpMax())));
        nmerged += lane63(scan_add(merged ? 1 : 0));
    }
    if (lane == 0) {
        A.bicnt[b] = (uint32_t)icarry;
        A.bmcnt[b] = (uint32_t)nmerged;
        A.bext[b] = icarry ? ((uint32_t)xmin | ((uint32_t)xmax << 16)) : 0xFFFFu;
    }
    __syncthreads();
}

constexpr int kResolveBlocks = 16384;

// Every bucket: empty ones are cleared, buckets of at most 64 fragments are resolved,
// larger ones are listed for the size-class passes below.
template <int SPP>
__global__ void __launch_bounds__(64) k_resolve_first(ResolveArgs A) {
    constexpr int NW = Words<SPP>::n;
    __shared__ uint32_t key[64], order[64], meta[64], msk[64 * NW];
    const int lane = threadIdx.x;
    for (int64_t b = blockIdx.x; b < A.NB; b += gridDim.x) {
        const uint32_t base = A.boff[b];
        const int n = (int)(A.boff[b + 1] - base);
        if (n == 0) {
            if (lane == 0) { A.bicnt[b] = 0; A.bmcnt[b] = 0; A.bext[b] = 0xFFFFu; }
            continue;
        }
        if (n > 64) {
            if (lane == 0) A.big_list[atomicAdd(A.big_count, 1u)] = (uint32_t)b;
            continue;
        }
        resolve_bucket<SPP, 1>(A, b, base, n, key, order, meta, msk);
    }
}

// Buckets of 64*S/2 < n <= 64*S fragments from the list.
template <int SPP, int S>
__global__ void __launch_bounds__(64) k_resolve_class(ResolveArgs A) {
    constexpr int NW = Words<SPP>::n;
    __shared__ uint32_t key[64 * S], order[64 * S], meta[64 * S], msk[64 * S * NW];
    const uint32_t nbig = *A.big_count;
    for (uint32_t bi = blockIdx.x; bi < nbig; bi += gridDim.x) {
        const int64_t b = A.big_list[bi];
        const uint32_t base = A.boff[b];
        const int n = (int)(A.boff[b + 1] - base);
        if (n <= 32 * S || n > 64 * S) continue;
        resolve_bucket<SPP, S>(A, b, base, n, key, order, meta, msk);
    }
}

// Buckets of more than 512 fragments (rare): bitonic-sort (cx << 16 | index) keys in a
// global scratch region, then walk the sorted order in chunks of 64 reading the records
// from global memory.
template <int SPP>
__global__ void __launch_bounds__(64) k_resolve_huge(ResolveArgs A) {
    constexpr int NW = Words<SPP>::n;
    const int lane = threadIdx.x;
    const uint32_t nbig = *A.big_count;
    for (uint32_t bi = blockIdx.x; bi < nbig; bi += gridDim.x) {
        const int64_t b = A.big_list[bi];
        const uint32_t base = A.boff[b];
        const int n = (int)(A.boff[b + 1] - base);
        if (n <= 512) continue;
        if (n > 65536) {
            if (lane == 0) { atomicOr(A.err, 2u); A.bicnt[b] = 0; A.bmcnt[b] = 0; A.bext[b] = 0xFFFFu; }
            continue;
        }
        int P2 = 1;
        while (P2 < n) P2 <<= 1;
        uint32_t *keys = A.scratch + 2 * (size_t)base;
        for (int i = lane; i < P2; i += 64)
            keys[i] = (i < n) ? (((A.rec_meta[base + i] & 0xFFFFu) << 16) | (uint32_t)i) : 0xFFFFFFFFu;
        __syncthreads();
        for (int k = 2; k <= P2; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = lane; i < P2; i += 64) {
                    const int l = i ^ j;
                    if (l > i) {
                        const uint32_t a = keys[i], c = keys[l];
                        const bool up = (i & k) == 0;
                        if ((a > c) == up) { keys[i] = c; keys[l] = a; }
                    }
                }
                __syncthreads();
            }
        }
        const int rule = A.frule[A.binfo[b].x];
        uint32_t wcarry = A.bcarry[b];
        int icarry = 0, nmerged = 0, xmin = INT_MAX, xmax = INT_MIN;
        const size_t ib = 2 * (size_t)base;
        for (int c0 = 0; c0 < n; c0 += 64) {
            const int p = c0 + lane;
            const bool v = p < n;
            const uint32_t kk = v ? keys[p] : 0xFFFFFFFFu;
            const int cx = (int)(kk >> 16);
            const uint32_t meta = v ? A.rec_meta[base + (kk & 0xFFFFu)] : 0u;
            const int swn = v ? (int)(int8_t)(meta >> 16) : 0;
            const int prev_cx = (v && p > 0) ? (int)(keys[p - 1] >> 16) : -1;
            const bool head = v && (p == 0 || cx != prev_cx);
            const int incl = scan_add(swn);
            const int8_t w_at = (int8_t)(int)(wcarry + (uint32_t)(incl - swn));
            wcarry += (uint32_t)lane63(incl);
            const int x = 2 * cx - 2;
            const bool merged = head && x >= 0 && x < A.W;
            uint32_t res[NW];
            if (merged) {
                Stencil<NW> st;
                st.init(w_at);
                for (int q = p; q < n; ++q) {
                    const uint32_t kq = keys[q];
                    if ((int)(kq >> 16) != cx) break;
                    const uint32_t j = kq & 0xFFFFu;
                    uint32_t m2[NW];
                    load_mask<NW>(A.rec_mask, (size_t)base + j, m2);
                    st.add(m2, ((A.rec_meta[base + j] >> 24) & 1u) != 0u);
                }
                st.finish(rule, res);
            }
            const int prev_x = 2 * prev_cx - 2;
            const bool span = head && p > 0 && (prev_x + 2 < x) && fill_test(rule, w_at);
            const int cnt = (merged ? 1 : 0) + (span ? 1 : 0);
            const int ic = scan_add(cnt);
            const int idx = icarry + ic - cnt;
            icarry += lane63(ic);
            if (span) {
                const int x0 = prev_x + 2;
                A.items[ib + idx] = (uint32_t)x0 | ((uint32_t)(x - x0) << 16);
            }
            if (merged) {
                A.items[ib + idx + (span ? 1 : 0)] = (uint32_t)x;
                store_mask<NW>(A.imask, ib + idx + (span ? 1 : 0), res);
            }
            const int lo = span ? prev_x + 2 : (merged ? x : INT_MAX);
            const int hi = merged ? x + 2 : (span ? x : INT_MIN);
            xmin = tmin(xmin, lane63(dpp_scan(lo, INT_MAX, OpMin())));
            xmax = tmax(xmax, lane63(dpp_scan(hi, INT_MIN, OpMax())));
            nmerged += lane63(scan_add(merged ? 1 : 0));
        }
        if (lane == 0) {
            A.bicnt[b] = (uint32_t)icarry;
            A.bmcnt[b] = (uint32_t)nmerged;
            A.bext[b] = icarry ? ((uint32_t)xmin | ((uint32_t)xmax << 16)) : 0xFFFFu;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------
// compositing order: non-empty buckets, stably sorted by cell row (path order kept)
// ---------------------------------------------------------------------------------
__global__ void k_nz_flags(int64_t NB, const uint32_t *__restrict__ bicnt, uint32_t *flag) {
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < NB) flag[b] = bicnt[b] ? 1u : 0u;
}
__global__ void k_nz_compact(int64_t NB, const uint32_t *__restrict__ flag, const uint32_t *__restrict__ off,
                             const int2 *__restrict__ binfo, uint32_t *keys, uint32_t *vals) {
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= NB || !flag[b]) return;
    const uint32_t o = off[b];
    keys[o] = (uint32_t)binfo[b].y;
    vals[o] = (uint32_t)b;
}

constexpr int kRsThreads = 256, kRsItems = 8, kRsTile = kRsThreads * kRsItems;

__global__ void __launch_bounds__(kRsThreads) k_rs_hist(const uint32_t *__restrict__ keys, const uint32_t *n_ptr,
                                                         int shift, uint32_t ntiles, uint32_t *hist) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t n = *n_ptr;
    const size_t base = (size_t)blockIdx.x * kRsTile;
    for (int k = 0; k < kRsItems; ++k) {
        size_t i = base + (size_t)k * kRsThreads + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

__global__ void __launch_bounds__(kRsThreads) k_rs_scatter(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                            uint32_t *kout, uint32_t *vout, const uint32_t *n_ptr, int shift,
                                                            uint32_t ntiles, const uint32_t *__restrict__ hoff) {
    __shared__ uint32_t wcnt[4][256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int d = threadIdx.x; d < 4 * 256; d += kRsThreads) (&wcnt[0][0])[d] = 0;
    __syncthreads();
    const uint32_t n = *n_ptr;
    const size_t base = (size_t)blockIdx.x * kRsTile + (size_t)w * (kRsTile / 4);
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t lr[kRsItems], dg[kRsItems], kk[kRsItems], vv[kRsItems];
#pragma unroll
    for (int r = 0; r < kRsItems; ++r) {
        const size_t i = base + (size_t)r * 64 + lane;
        const bool valid = i < n;
        kk[r] = valid ? kin[i] : 0u;
        vv[r] = valid ? vin[i] : 0u;
        const uint32_t d = (kk[r] >> shift) & 255u;
        dg[r] = d;
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const unsigned long long bb = __ballot((d >> bit) & 1u);
            m &= ((d >> bit) & 1u) ? bb : ~bb;
        }
        const uint32_t before = valid ? wcnt[w][d] : 0u;
        const int rk = __popcll(m & lt);
        lr[r] = before + (uint32_t)rk;
        const int tot = __popcll(m);
        __builtin_amdgcn_wave_barrier();
        if (valid && rk == tot - 1) wcnt[w][d] = before + (uint32_t)tot;
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    {
        const int d = threadIdx.x;
        uint32_t run = 0;
        for (int ww = 0; ww < 4; ++ww) {
            uint32_t t = wcnt[ww][d];
            wcnt[ww][d] = run;
            run += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRsItems; ++r) {
        const size_t i = base + (size_t)r * 64 + lane;
        if (i < n) {
            const uint32_t d = dg[r];
            const uint32_t pos = hoff[(size_t)d * ntiles + blockIdx.x] + wcnt[w][d] + lr[r];
            kout[pos] = kk[r];
            vout[pos] = vv[r];
        }
    }
}

__global__ void k_row_bounds(const uint32_t *__restrict__ keys, const uint32_t *n_ptr, uint32_t *rowbeg, uint32_t *rowend) {
    const uint32_t n = *n_ptr;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = keys[i];
    if (i == 0 || keys[i - 1] != r) rowbeg[r] = i;
    if (i + 1 == n || keys[i + 1] != r) rowend[r] = i + 1;
}

// ---------------------------------------------------------------------------------
// compositor: one wavefront per 64x2 pixel tile; lane = pixel column, two rows.  For each
// bucket of the tile's cell row (path order), the lanes blend the item covering their
// pixels: src-over into an 8-bit framebuffer re-quantised after each blend, coverage
// alpha = paint alpha * popcount(mask)/N (output_{8,32}.frag.glsl), or per-sample
// blending + resolve in multisample mode (ms_output_{8,32}.frag.glsl).
// ---------------------------------------------------------------------------------
struct CompArgs {
    int W, H, tiles_x;
    const uint32_t *rowbeg, *rowend, *rowvals;
    const int2 *binfo;
    const uint32_t *finfo, *bicnt, *boff, *bext, *items, *imask;
    const float *dec_c, *thr_c, *dec_a, *thr_a;
    const uint8_t *lut_c, *lut_a;
    uint8_t *image;
    int pitch;
};

struct LdsTables {
    float dec_c[256], thr_c[256], dec_a[256], thr_a[256];
    uint8_t lut_c[4096], lut_a[4096];
};

__device__ __forceinline__ uint32_t enc8(const float *thr, const uint8_t *lut, float v) {
    if (!(v >= 0.f)) return 0u;
    if (v >= 1.f) return 255u;
    uint32_t k = lut[(int)(v * 4096.f)];
    if (k < 255u && v >= thr[k + 1]) ++k;
    return k;
}

__device__ __forceinline__ uint32_t blend_px(const LdsTables &T, uint32_t dst, const float *sc, float a) {
    const float oma = 1.f - a;
    uint32_t out = 0u;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        float o = sc[ch] * a + T.dec_c[(dst >> (8 * ch)) & 0xFFu] * oma;
        out |= enc8(T.thr_c, T.lut_c, o) << (8 * ch);
    }
    float oa = a * a + T.dec_a[dst >> 24] * oma;
    out |= enc8(T.thr_a, T.lut_a, oa) << 24;
    return out;
}

template <int SPP>
__device__ __forceinline__ uint32_t pixel_bits(const uint32_t *m, int dx, int dy) {
    if (SPP == 8) return (m[0] >> ((dx * 2 + dy) * 8)) & 0xFFu;
    if (SPP == 16) return (m[dx] >> (dy * 16)) & 0xFFFFu;
    return m[dx * 2 + dy];
}

template <int SPP, bool MS>
__global__ void __launch_bounds__(256) k_composite(CompArgs A) {
    constexpr int NW = Words<SPP>::n;
    constexpr uint32_t FULL = SPP == 32 ? 0xFFFFFFFFu : ((1u << SPP) - 1u);
    __shared__ LdsTables T;
    for (int i = threadIdx.x; i < 256; i += 256) {
        T.dec_c[i] = A.dec_c[i]; T.thr_c[i] = A.thr_c[i];
        T.dec_a[i] = A.dec_a[i]; T.thr_a[i] = A.thr_a[i];
    }
    for (int i = threadIdx.x; i < 4096 / 4; i += 256) {
        ((uint32_t *)T.lut_c)[i] = ((const uint32_t *)A.lut_c)[i];
        ((uint32_t *)T.lut_a)[i] = ((const uint32_t *)A.lut_a)[i];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int R = (A.H + 1) >> 1;
    if (tile >= R * A.tiles_x) return;
    const int r = tile / A.tiles_x;
    const int tx0 = (tile - r * A.tiles_x) * 64;
    const int x = tx0 + lane;
    constexpr int NS = MS ? SPP : 1;
    uint32_t st0[NS], st1[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) { st0[s] = 0xFFFFFFFFu; st1[s] = 0xFFFFFFFFu; }

    const uint32_t q0 = A.rowbeg[r], q1 = A.rowend[r];
    for (uint32_t qc = q0; qc < q1; qc += 64) {
        const uint32_t q = qc + lane;
        const bool inq = q < q1;
        const uint32_t bb = inq ? A.rowvals[q] : 0u;
        const uint32_t ext = inq ? A.bext[bb] : 0xFFFFu;
        const int exlo = (int)(ext & 0xFFFFu), exhi = (int)(ext >> 16);
        const bool ov = inq && exlo < tx0 + 64 && exhi > tx0;
        unsigned long long om = __ballot(ov);
        uint32_t col = 0u, icnt = 0u, ibase = 0u;
        if (ov) {
            col = A.finfo[A.binfo[bb].x];
            icnt = A.bicnt[bb];
            ibase = 2u * A.boff[bb];
        }
        while (om) {
            const int j = __ffsll((long long)om) - 1;
            om &= om - 1;
            const uint32_t c = __shfl(col, j, 64);
            const int cnt = (int)__shfl(icnt, j, 64);
            const uint32_t ib = __shfl(ibase, j, 64);
            float sc[3];
            sc[0] = T.dec_c[c & 0xFFu];
            sc[1] = T.dec_c[(c >> 8) & 0xFFu];
            sc[2] = T.dec_c[(c >> 16) & 0xFFu];
            const float sa = T.dec_a[c >> 24];
            for (int ic = 0; ic < cnt; ic += 64) {
                const int m = tmin(64, cnt - ic);
                const uint32_t it = lane < m ? A.items[ib + ic + lane] : 0u;
                const int ix0 = (int)(it & 0xFFFFu);
                const int iw = (int)(it >> 16);
                const int iend = ix0 + (iw ? iw : 2);
                const int first = __shfl(ix0, 0, 64);
                const int lastend = __shfl(iend, m - 1, 64);
                if (lastend <= tx0) continue;
                if (first >= tx0 + 64) break;
                // last item with x0 <= x (items are disjoint and sorted)
                int pos = -1;
#pragma unroll
                for (int step = 64; step >= 1; step >>= 1) {
                    const int cand = pos + step;
                    const int cx0 = __shfl(ix0, cand < 64 ? cand : 63, 64);
                    if (cand < m && cx0 <= x) pos = cand;
                }
                const int pend = __shfl(iend, pos < 0 ? 0 : pos, 64);
                const int pw = __shfl(iw, pos < 0 ? 0 : pos, 64);
                const int px0 = __shfl(ix0, pos < 0 ? 0 : pos, 64);
                const bool cov = pos >= 0 && x < pend && x < A.W;
                if (!cov) continue;
                uint32_t b0 = FULL, b1 = FULL;
                if (pw == 0) {
                    uint32_t mk[NW];
                    load_mask<NW>(A.imask, (size_t)ib + ic + pos, mk);
                    b0 = pixel_bits<SPP>(mk, x - px0, 0);
                    b1 = pixel_bits<SPP>(mk, x - px0, 1);
                }
                if (!MS) {
                    const float a0 = sa * ((float)__popc(b0) / (float)SPP);
                    const float a1 = sa * ((float)__popc(b1) / (float)SPP);
                    st0[0] = blend_px(T, st0[0], sc, a0);
                    st1[0] = blend_px(T, st1[0], sc, a1);
                } else {
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        if ((b0 >> s) & 1u) st0[s] = blend_px(T, st0[s], sc, sa);
                        if ((b1 >> s) & 1u) st1[s] = blend_px(T, st1[s], sc, sa);
                    }
                }
            }
        }
    }
    if (x >= A.W) return;
    uint32_t o0, o1;
    if (!MS) {
        o0 = st0[0];
        o1 = st1[0];
    } else {
        uint32_t outs[2];
#pragma unroll
        for (int rr = 0; rr < 2; ++rr) {
            const uint32_t *stp = rr ? st1 : st0;
            uint32_t o = 0u;
#pragma unroll
            for (int ch = 0; ch < 4; ++ch) {
                const float *dec = ch < 3 ? T.dec_c : T.dec_a;
                float sum = 0.f;
#pragma unroll
                for (int s = 0; s < NS; ++s) sum += dec[(stp[s] >> (8 * ch)) & 0xFFu];
                const float avg = sum / (float)SPP;
                o |= (ch < 3 ? enc8(T.thr_c, T.lut_c, avg) : enc8(T.thr_a, T.lut_a, avg)) << (8 * ch);
            }
            outs[rr] = o;
        }
        o0 = outs[0];
        o1 = outs[1];
    }
    const int y0 = 2 * r;
    *(uint32_t *)(A.image + (size_t)y0 * A.pitch + (size_t)x * 4) = o0;
    if (y0 + 1 < A.H) *(uint32_t *)(A.image + (size_t)(y0 + 1) * A.pitch + (size_t)x * 4) = o1;
}

// draw list + merged masks in the reference's order and layout (.cu:789-868, 1195-1266)
template <int SPP>
__global__ void k_debug_draw(int64_t NB, const int2 *__restrict__ binfo, const uint32_t *__restrict__ finfo,
                             const uint32_t *__restrict__ boff, const uint32_t *__restrict__ bicnt,
                             const uint32_t *__restrict__ items, const uint32_t *__restrict__ imask,
                             const uint32_t *__restrict__ moff, const uint32_t *__restrict__ ioff, int4 *draw,
                             uint32_t *mmask4) {
    constexpr int NW = Words<SPP>::n;
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= NB) return;
    const uint32_t cnt = bicnt[b];
    if (!cnt) return;
    const int2 info = binfo[b];
    const uint32_t fill = finfo[info.x];
    const int y = 2 * info.y;
    const size_t ib = 2 * (size_t)boff[b];
    uint32_t mi = moff[b];
    const uint32_t io = ioff[b];
    uint32_t o = io;
    auto put_merged = [&](uint32_t j, int x0) {
        draw[o++] = make_int4((int32_t)(((uint32_t)y << 16) | (uint32_t)x0), 2, (int)fill, (int)(mi + 1));
        uint32_t mk[4] = {0u, 0u, 0u, 0u};
        load_mask<NW>(imask, ib + j, mk);
        *(uint4 *)(mmask4 + 4 * (size_t)mi) = make_uint4(mk[0], mk[1], mk[2], mk[3]);
        ++mi;
    };
    for (uint32_t j = 0; j < cnt; ++j) {
        const uint32_t it = items[ib + j];
        const int x0 = (int)(it & 0xFFFFu), w = (int)(it >> 16);
        if (w == 0) {
            put_merged(j, x0);
            continue;
        }
        // a span followed by the merged fragment it ends at: the reference emits that
        // fragment first (.cu:836-862)
        const bool pair = j + 1 < cnt && (items[ib + j + 1] >> 16) == 0u &&
                          (int)(items[ib + j + 1] & 0xFFFFu) == x0 + w;
        if (pair) put_merged(j + 1, x0 + w);
        draw[o++] = make_int4((int32_t)(((uint32_t)y << 16) | (uint32_t)x0), w, (int)fill, 0);
        if (pair) ++j;
    }
}

__global__ void k_count_reduce(const uint32_t *__restrict__ a, size_t n, unsigned long long *out) {
    unsigned long long s = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) s += a[i];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, s);
}

inline unsigned blocks_for(int64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

void launch_resolve(const SceneDev &sc, const Tables &t, Frame &f, hipStream_t s) {
    if (!f.NB) return;
    ResolveArgs A;
    A.NB = f.NB; A.binfo = f.binfo; A.boff = f.boff; A.bcarry = f.bcarry; A.frule = sc.frule;
    A.rec_meta = f.rec_meta; A.rec_mask = f.rec_mask; A.scratch = f.scratch_keys; A.W = f.W;
    A.items = f.items; A.imask = f.imask; A.bicnt = f.bicnt; A.bmcnt = f.bmcnt; A.bext = f.bext;
    A.big_list = f.big_list; A.big_count = f.big_count; A.err = f.err_flag;
    launch_fill_u32(f.big_count, 0u, 1, s);
    const unsigned g1 = (unsigned)std::min<int64_t>(f.NB, kResolveBlocks);
    dim3 b64(64);
#define SL_LAUNCH(N)                                                                  \
    do {                                                                              \
        hipLaunchKernelGGL(k_resolve_first<N>, dim3(g1), b64, 0, s, A);               \
        hipLaunchKernelGGL((k_resolve_class<N, 2>), dim3(4096), b64, 0, s, A);        \
        hipLaunchKernelGGL((k_resolve_class<N, 4>), dim3(4096), b64, 0, s, A);        \
        hipLaunchKernelGGL((k_resolve_class<N, 8>), dim3(4096), b64, 0, s, A);        \
        hipLaunchKernelGGL(k_resolve_huge<N>, dim3(1024), b64, 0, s, A);              \
    } while (0)
    if (t.spp == 8) SL_LAUNCH(8);
    else if (t.spp == 16) SL_LAUNCH(16);
    else SL_LAUNCH(32);
#undef SL_LAUNCH
}

void launch_order_rows(const SceneDev &sc, Frame &f, hipStream_t s) {
    (void)sc;
    const int R = (f.H + 1) / 2;
    launch_fill_u32(f.rowbeg, 0u, (size_t)R, s);
    launch_fill_u32(f.rowend, 0u, (size_t)R, s);
    if (!f.NB) return;
    hipLaunchKernelGGL(k_nz_flags, dim3(blocks_for(f.NB, 256)), dim3(256), 0, s, f.NB, (const uint32_t *)f.bicnt, f.nz_flag);
    scan_exclusive_u32(f.nz_flag, f.nz_off, (size_t)f.NB, f.scan_tmp, s);
    hipLaunchKernelGGL(k_nz_compact, dim3(blocks_for(f.NB, 256)), dim3(256), 0, s, f.NB, (const uint32_t *)f.nz_flag,
                       (const uint32_t *)f.nz_off, (const int2 *)f.binfo, f.nz_keys[0], f.nz_vals[0]);
    const uint32_t *n_ptr = f.nz_off + f.NB;
    const uint32_t ntiles = (uint32_t)((f.NB + kRsTile - 1) / kRsTile);
    int bits = 0;
    while ((1 << bits) < R) ++bits;
    const int passes = bits <= 8 ? 1 : (bits + 7) / 8;
    int cur = 0;
    for (int pss = 0; pss < passes; ++pss) {
        const int shift = 8 * pss;
        hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(kRsThreads), 0, s, (const uint32_t *)f.nz_keys[cur], n_ptr, shift,
                           ntiles, f.rs_hist);
        scan_exclusive_u32(f.rs_hist, f.rs_hoff, (size_t)256 * ntiles, f.scan_tmp, s);
        hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(kRsThreads), 0, s, (const uint32_t *)f.nz_keys[cur],
                           (const uint32_t *)f.nz_vals[cur], f.nz_keys[cur ^ 1], f.nz_vals[cur ^ 1], n_ptr, shift, ntiles,
                           (const uint32_t *)f.rs_hoff);
        cur ^= 1;
    }
    if (cur != 0) {
        // keep the sorted lists in slot 0 for the compositor
        uint32_t *k = f.nz_keys[0], *v = f.nz_vals[0];
        f.nz_keys[0] = f.nz_keys[1]; f.nz_vals[0] = f.nz_vals[1];
        f.nz_keys[1] = k; f.nz_vals[1] = v;
    }
    hipLaunchKernelGGL(k_row_bounds, dim3(blocks_for(f.NB, 256)), dim3(256), 0, s, (const uint32_t *)f.nz_keys[0], n_ptr,
                       f.rowbeg, f.rowend);
}

void launch_composite(const SceneDev &sc, const Tables &t, Frame &f, hipStream_t s) {
    CompArgs A;
    A.W = f.W; A.H = f.H; A.tiles_x = (f.W + 63) / 64;
    A.rowbeg = f.rowbeg; A.rowend = f.rowend; A.rowvals = f.nz_vals[0];
    A.binfo = f.binfo; A.finfo = sc.finfo; A.bicnt = f.bicnt; A.boff = f.boff; A.bext = f.bext;
    A.items = f.items; A.imask = f.imask;
    A.dec_c = t.dec_c; A.thr_c = t.thr_c; A.dec_a = t.dec_a; A.thr_a = t.thr_a; A.lut_c = t.lut_c; A.lut_a = t.lut_a;
    A.image = f.image; A.pitch = f.pitch;
    const int R = (f.H + 1) / 2;
    const int ntiles = R * A.tiles_x;
    dim3 g(blocks_for(ntiles, 4)), b(256);
#define SL_LAUNCH(N, MS) hipLaunchKernelGGL((k_composite<N, MS>), g, b, 0, s, A)
    if (!f.ms) {
        if (t.spp == 8) SL_LAUNCH(8, false);
        else if (t.spp == 16) SL_LAUNCH(16, false);
        else SL_LAUNCH(32, false);
    } else {
        if (t.spp == 8) SL_LAUNCH(8, true);
        else if (t.spp == 16) SL_LAUNCH(16, true);
        else SL_LAUNCH(32, true);
    }
#undef SL_LAUNCH
}

void launch_debug_draw_list(const SceneDev &sc, const Tables &t, Frame &f, const uint32_t *moff, const uint32_t *ioff,
                            int4 *draw, uint32_t *mmask4, hipStream_t s) {
    if (!f.NB) return;
    dim3 g(blocks_for(f.NB, 256)), b(256);
    if (t.spp == 8)
        hipLaunchKernelGGL(k_debug_draw<8>, g, b, 0, s, f.NB, (const int2 *)f.binfo, sc.finfo, (const uint32_t *)f.boff,
                           (const uint32_t *)f.bicnt, (const uint32_t *)f.items, (const uint32_t *)f.imask, moff, ioff, draw, mmask4);
    else if (t.spp == 16)
        hipLaunchKernelGGL(k_debug_draw<16>, g, b, 0, s, f.NB, (const int2 *)f.binfo, sc.finfo, (const uint32_t *)f.boff,
                           (const uint32_t *)f.bicnt, (const uint32_t *)f.items, (const uint32_t *)f.imask, moff, ioff, draw, mmask4);
    else
        hipLaunchKernelGGL(k_debug_draw<32>, g, b, 0, s, f.NB, (const int2 *)f.binfo, sc.finfo, (const uint32_t *)f.boff,
                           (const uint32_t *)f.bicnt, (const uint32_t *)f.items, (const uint32_t *)f.imask, moff, ioff, draw, mmask4);
}

void launch_count_reduce(const uint32_t *a, size_t n, unsigned long long *out, hipStream_t s) {
    hipLaunchKernelGGL(k_count_reduce, dim3(256), dim3(256), 0, s, a, n, out);
}

}  // namespace sl
