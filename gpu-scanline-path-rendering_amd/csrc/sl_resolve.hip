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
    // per byte lane: v = (0x80 + counter) & 0x7f; nonzero keeps v != 0, even-odd keeps
    // bit 0.  v <= 0x7f, so v + 0x7f carries into bit 7 of its own byte exactly when v != 0.
    __device__ __forceinline__ void finish(int rule, uint32_t *out) const {
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            uint32_t ret = 0u;
            if (rule == 0) {
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    const uint32_t v = (0x80808080u + acc[k][b]) & 0x7f7f7f7fu;
                    ret |= ((v + 0x7f7f7f7fu) & 0x80808080u) >> (7 - b);
                }
            } else {
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    const uint32_t v = (0x80808080u + acc[k][b]) & 0x7f7f7f7fu;
                    ret |= (v & 0x01010101u) << b;
                }
            }
            out[k] = ret;
        }
    }
};

// The same result for small groups, bit-sliced.  The byte lanes of the stencil word start
// at w * 0x01010101 and gain +/- 0x01 at set bits, so the word is sum_j c_j * 256^j
// (mod 2^32) with c_j the sample's true count w + d_j.  While every c_j lies in
// [-128, 127], adding 0x80808080 leaves byte j = c_j + 128 with no carry between lanes,
// and the tests read c_j mod 128 != 0 (nonzero) or c_j & 1 (even-odd).  For g <= 7
// members and |w| <= 120, c_j lies in [-127, 127], so it is enough to count d_j in
// [-7, 7] in four two's-complement bit planes: nonzero is d_j != -w, even-odd is
// d_j ^ w in plane 0.  Groups outside that range take the byte-lane stencil.
constexpr int kSlicedMaxGroup = 7;
constexpr int kSlicedMaxWinding = 120;

__device__ __forceinline__ bool sliced_ok(int g, int8_t w) {
    return g <= kSlicedMaxGroup && w >= -kSlicedMaxWinding && w <= kSlicedMaxWinding;
}

template <int NW>
struct Sliced {
    uint32_t d[4][NW];
    // d = +1 or -1 (1111 in four planes) at the set bits of one member
    __device__ __forceinline__ void init_one(const uint32_t *m, bool neg) {
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            d[0][k] = m[k];
#pragma unroll
            for (int p = 1; p < 4; ++p) d[p][k] = neg ? m[k] : 0u;
        }
    }
    // d += 1 (or -= 1) at the set bits: ripple carry (borrow) through the planes
    __device__ __forceinline__ void add(const uint32_t *m, bool neg) {
        const uint32_t nm = neg ? ~0u : 0u;
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            uint32_t c = m[k];
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const uint32_t t = (d[p][k] ^ nm) & c;
                d[p][k] ^= c;
                c = t;
            }
        }
    }
    __device__ __forceinline__ void finish(int rule, int8_t w, uint32_t *out) const {
        if (rule != 0) {
            const uint32_t odd = (w & 1) ? ~0u : 0u;
#pragma unroll
            for (int k = 0; k < NW; ++k) out[k] = d[0][k] ^ odd;
            return;
        }
        const int t = -(int)w;
        const bool hit = t >= -8 && t <= 7;  // d_j in [-7, 7] can equal -w only then
        uint32_t tp[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) tp[p] = ((t >> p) & 1) ? ~0u : 0u;
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            uint32_t ne = 0u;
#pragma unroll
            for (int p = 0; p < 4; ++p) ne |= d[p][k] ^ tp[p];
            out[k] = hit ? ne : ~0u;
        }
    }
};

// Sample counts of the four pixels of a resolved cell mask, 6 bits each at bit 6*(dx*2+dy).
template <int SPP>
__device__ __forceinline__ uint32_t cell_counts(const uint32_t *m) {
    uint32_t c = 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint32_t bits;
        if (SPP == 8) bits = (m[0] >> (q * 8)) & 0xFFu;
        else if (SPP == 16) bits = (m[q >> 1] >> ((q & 1) * 16)) & 0xFFFFu;
        else bits = m[q];
        c |= (uint32_t)__popc(bits) << (6 * q);
    }
    return c;
}

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
    const int4 *pcanvas;     // batched scenes: per-path canvas (W, H, row base, scene)
    const uint32_t *scarry;  // batched scenes: winding-scan carry at each scene's first bucket
    uint32_t *items;
    uint32_t *imask;         // null when no consumer needs the merged masks
    uint32_t *icov;
    uint32_t *bicnt;
    uint32_t *bmcnt;
    uint32_t *bext;
    uint64_t *btile;         // [NB] 64-pixel tiles holding items, bit k = the k-th tile from the
                             // one holding the extent's start (bit 63: that tile or any beyond)
    uint32_t *big_list;
    uint32_t *big_count;
    uint32_t *err;
};

// 64-pixel tiles spanned by [lo, hi) as bits of a 64-bit mask (all tiles when the canvas
// is wider than 64 tiles), OR-reduced over the wave into lane 63.
__device__ __forceinline__ uint64_t tile_bits(int lo, int hi, int tb) {
    if (lo >= hi) return 0ull;
    const int t0 = tmin((lo >> 6) - tb, 63), t1 = tmin(((hi - 1) >> 6) - tb, 63);
    return (~0ull << t0) & (~0ull >> (63 - t1));
}

// The bit a tile starting at tx0 tests in the bitmap of a bucket whose extent starts at
// exlo (tiles at or beyond the 63rd from the extent's first all test bit 63).
__device__ __forceinline__ int tile_bit(int tx0, int exlo) { return tmax(tmin((tx0 >> 6) - (exlo >> 6), 63), 0); }
__device__ __forceinline__ uint64_t wave_or64(uint64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane(dpp_scan((int)(uint32_t)v, 0, OpOr()), 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane(dpp_scan((int)(uint32_t)(v >> 32), 0, OpOr()), 63);
    return (uint64_t)lo | ((uint64_t)hi << 32);
}
__device__ __forceinline__ int scan_add(int x) { return dpp_scan(x, 0, OpAdd()); }
__device__ __forceinline__ int lane63(int x) { return __builtin_amdgcn_readlane(x, 63); }

// Merged mask of the cell group at sorted positions [p, e) (records staged in LDS).  The
// first member is the head lane's own record i0 (meta m0): most groups have no other.
template <int NW>
__device__ __forceinline__ void group_mask(const uint32_t *order, const uint32_t *meta, const uint32_t *msk, int p,
                                           int e, uint32_t i0, uint32_t m0, int8_t w_at, int rule, uint32_t *res) {
    uint32_t first[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) first[k] = msk[i0 * NW + k];
    const bool neg0 = ((m0 >> 24) & 1u) != 0u;
    if (sliced_ok(e - p, w_at)) {
        Sliced<NW> st;
        st.init_one(first, neg0);
        for (int q = p + 1; q < e; ++q) {
            const uint32_t iq = order[q];
            st.add(msk + iq * NW, ((meta[iq] >> 24) & 1u) != 0u);
        }
        st.finish(rule, w_at, res);
    } else {
        Stencil<NW> st;
        st.init(w_at);
        st.add(first, neg0);
        for (int q = p + 1; q < e; ++q) {
            const uint32_t iq = order[q];
            st.add(msk + iq * NW, ((meta[iq] >> 24) & 1u) != 0u);
        }
        st.finish(rule, res);
    }
}

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
#pragma unroll 8
        for (int j = 0; j < n; ++j) {
            const uint32_t kj = (uint32_t)__builtin_amdgcn_readlane((int)mykey[0], j);
            rk[0] += (kj < mykey[0]) ? 1 : 0;
        }
    } else {
#pragma unroll 8
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
    int W = A.W;
    if (A.pcanvas) {
        // each scene of a batch scans its winding numbers from zero
        const int4 cv = A.pcanvas[A.binfo[b].x];
        W = cv.x;
        wcarry -= A.scarry[cv.w];
    }
    int icarry = 0, nmerged = 0, xmin = INT_MAX, xmax = INT_MIN;
    uint64_t tmask = 0ull;
    const size_t ib = 2 * (size_t)base;
    // group heads of every chunk as wave masks: a group ends at the next head
    uint64_t hm[S];
#pragma unroll
    for (int c = 0; c < S; ++c) {
        const int p = c * 64 + lane;
        bool h = false;
        if (p < n) h = p == 0 || (meta[order[p]] & 0xFFFFu) != (meta[order[p - 1]] & 0xFFFFu);
        hm[c] = __ballot(h);
    }
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
        const bool merged = head && x >= 0 && x < W;
        uint32_t res[NW];
        if (merged) {
            const uint64_t later = lane < 63 ? (hm[c] & (~0ull << (lane + 1))) : 0ull;
            int e = n;
            if (later) {
                e = c * 64 + (int)__builtin_ctzll(later);
            } else {
                for (int c2 = c + 1; c2 < S; ++c2)
                    if (hm[c2]) { e = c2 * 64 + (int)__builtin_ctzll(hm[c2]); break; }
            }
            group_mask<NW>(order, meta, msk, p, e, i, m, w_at, rule, res);
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
            if (A.imask) store_mask<NW>(A.imask, ib + idx + (span ? 1 : 0), res);
            A.icov[ib + idx + (span ? 1 : 0)] = cell_counts<SPP>(res);
        }
        const int lo = span ? prev_x + 2 : (merged ? x : INT_MAX);
        const int hi = merged ? x + 2 : (span ? x : INT_MIN);
        xmin = tmin(xmin, lane63(dpp_scan(lo, INT_MAX, OpMin())));
        xmax = tmax(xmax, lane63(dpp_scan(hi, INT_MIN, OpMax())));
        tmask |= wave_or64(tile_bits(lo, hi, xmin >> 6));
        nmerged += lane63(scan_add(merged ? 1 : 0));
    }
    if (lane == 0) {
        A.bicnt[b] = (uint32_t)icarry;
        A.bmcnt[b] = (uint32_t)nmerged;
        A.bext[b] = icarry ? ((uint32_t)xmin | ((uint32_t)xmax << 16)) : 0xFFFFu;
        A.btile[b] = tmask;
    }
    __syncthreads();
}

constexpr int kResolveBlocks = 65536;

// One small bucket's records, loaded a bucket ahead of its resolve (the grid-stride loop
// keeps the next bucket's records in flight while the current one resolves); the bucket
// scalars come from per-lane registers (see k_resolve_first).
template <int NW>
struct SmallIn {
    uint32_t meta, mk[NW];
    uint32_t carry;
    int rule, W;
};

template <int NW>
__device__ __forceinline__ void small_load(const ResolveArgs &A, uint32_t base, uint32_t end, SmallIn<NW> &o) {
    const int lane = threadIdx.x & 63;
    const int n = (int)(end - base);
    o.meta = 0u;
#pragma unroll
    for (int k = 0; k < NW; ++k) o.mk[k] = 0u;
    if (n > 0 && n <= 64 && lane < n) {
        o.meta = A.rec_meta[base + lane];
        load_mask<NW>(A.rec_mask, (size_t)base + lane, o.mk);
    }
}

// resolve_bucket<SPP, 1> with the records already in registers
template <int SPP>
__device__ __forceinline__ void resolve_small(const ResolveArgs &A, int64_t b, uint32_t base, int n,
                                              const SmallIn<Words<SPP>::n> &in, uint32_t *order, uint32_t *meta,
                                              uint32_t *msk) {
    constexpr int NW = Words<SPP>::n;
    const int lane = threadIdx.x & 63;
    uint32_t mykey = 0xFFFFFFFFu;
    if (lane < n) {
        meta[lane] = in.meta;
#pragma unroll
        for (int k = 0; k < NW; ++k) msk[lane * NW + k] = in.mk[k];
        mykey = ((in.meta & 0xFFFFu) << 10) | (uint32_t)lane;
    }
    int rk = 0;
#pragma unroll 8
    for (int j = 0; j < n; ++j) {
        const uint32_t kj = (uint32_t)__builtin_amdgcn_readlane((int)mykey, j);
        rk += (kj < mykey) ? 1 : 0;
    }
    if (lane < n) order[rk] = (uint32_t)lane;
    __syncthreads();

    const int rule = in.rule;
    uint32_t wcarry = in.carry;
    const int W = in.W;
    const size_t ib = 2 * (size_t)base;
    const int p = lane;
    const bool v = p < n;
    const uint32_t i = v ? order[p] : 0u;
    const uint32_t m = v ? meta[i] : 0u;
    const int cx = (int)(m & 0xFFFFu);
    const int swn = v ? (int)(int8_t)(m >> 16) : 0;
    const int prev_cx = (v && p > 0) ? (int)(meta[order[p - 1]] & 0xFFFFu) : -1;
    const bool head = v && (p == 0 || cx != prev_cx);
    const int incl = scan_add(swn);
    const int8_t w_at = (int8_t)(int)(wcarry + (uint32_t)(incl - swn));
    const int x = 2 * cx - 2;
    const bool merged = head && x >= 0 && x < W;
    // the group of a head lane ends at the next head (or the bucket end)
    const uint64_t heads = __ballot(head);
    uint32_t res[NW];
    if (merged) {
        const uint64_t later = p < 63 ? (heads & (~0ull << (p + 1))) : 0ull;
        const int e = later ? (int)__builtin_ctzll(later) : n;
        group_mask<NW>(order, meta, msk, p, e, i, m, w_at, rule, res);
    }
    const int prev_x = 2 * prev_cx - 2;
    const bool span = head && p > 0 && (prev_x + 2 < x) && fill_test(rule, w_at);
    const int cnt = (merged ? 1 : 0) + (span ? 1 : 0);
    const int ic = scan_add(cnt);
    const int idx = ic - cnt;
    if (span) {
        const int x0 = prev_x + 2;
        A.items[ib + idx] = (uint32_t)x0 | ((uint32_t)(x - x0) << 16);
    }
    if (merged) {
        A.items[ib + idx + (span ? 1 : 0)] = (uint32_t)x;
        if (A.imask) store_mask<NW>(A.imask, ib + idx + (span ? 1 : 0), res);
        A.icov[ib + idx + (span ? 1 : 0)] = cell_counts<SPP>(res);
    }
    const int lo = span ? prev_x + 2 : (merged ? x : INT_MAX);
    const int hi = merged ? x + 2 : (span ? x : INT_MIN);
    const int xmin = lane63(dpp_scan(lo, INT_MAX, OpMin()));
    const int xmax = lane63(dpp_scan(hi, INT_MIN, OpMax()));
    const int nmerged = lane63(scan_add(merged ? 1 : 0));
    const int icnt = lane63(ic);
    const uint64_t tmask = wave_or64(tile_bits(lo, hi, xmin >> 6));
    if (lane == 0) {
        A.bicnt[b] = (uint32_t)icnt;
        A.bmcnt[b] = (uint32_t)nmerged;
        A.bext[b] = icnt ? ((uint32_t)xmin | ((uint32_t)xmax << 16)) : 0xFFFFu;
        A.btile[b] = tmask;
    }
    __syncthreads();
}

// Buckets of at most S (16 or 32) fragments, 64 / S of them in one wave: lane l works on
// segment l / S, whose records, rank sort, scans and reductions stay inside its S lanes
// (segmented DPP scans), so one pass resolves the whole pack.  ks[s] is the s-th bucket's
// slot in the per-lane scalars (lane k holds bucket b0 + k G), -1 for an empty segment.
template <int SPP, int S>
__device__ __forceinline__ void resolve_pack(const ResolveArgs &A, int64_t b0, int64_t G, const int *ks, int vbase,
                                             int vend, int vcarry, int vrule, int vW, uint32_t *order, uint32_t *meta,
                                             uint32_t *msk) {
    constexpr int NW = Words<SPP>::n;
    constexpr int P = 64 / S;
    const int lane = threadIdx.x & 63;
    const int seg = lane / S, i = lane % S, sb = seg * S;
    int k = -1;
#pragma unroll
    for (int s = 0; s < P; ++s)
        if (seg == s) k = ks[s];
    // every lane takes part in the cross-lane reads (a lane outside the exec mask would
    // supply zeros to the lanes reading from it)
    const int src = k < 0 ? 0 : k;
    const uint32_t base = (uint32_t)__shfl(vbase, src, 64);
    const int end = __shfl(vend, src, 64);
    const uint32_t wcarry = (uint32_t)__shfl(vcarry, src, 64);
    const int rule = __shfl(vrule, src, 64);
    const int W = __shfl(vW, src, 64);
    const int n = k < 0 ? 0 : end - (int)base;
    uint32_t mykey = 0xFFFFFFFFu;
    if (i < n) {
        const uint32_t m = A.rec_meta[base + i];
        uint32_t mk[NW];
        load_mask<NW>(A.rec_mask, (size_t)base + i, mk);
        meta[lane] = m;
#pragma unroll
        for (int q = 0; q < NW; ++q) msk[lane * NW + q] = mk[q];
        mykey = ((m & 0xFFFFu) << 10) | (uint32_t)i;
    }
    int rk = 0;
#pragma unroll
    for (int j = 0; j < S; ++j) rk += ((uint32_t)__shfl((int)mykey, sb + j, 64) < mykey) ? 1 : 0;
    if (i < n) order[sb + rk] = (uint32_t)lane;  // LDS index of the record
    __syncthreads();

    const int p = i;
    const bool v = p < n;
    const uint32_t i0 = v ? order[sb + p] : 0u;
    const uint32_t m = v ? meta[i0] : 0u;
    const int cx = (int)(m & 0xFFFFu);
    const int swn = v ? (int)(int8_t)(m >> 16) : 0;
    const int prev_cx = (v && p > 0) ? (int)(meta[order[sb + p - 1]] & 0xFFFFu) : -1;
    const bool head = v && (p == 0 || cx != prev_cx);
    const int incl = dpp_scan_seg<S>(swn, 0, OpAdd());
    const int8_t w_at = (int8_t)(int)(wcarry + (uint32_t)(incl - swn));
    const int x = 2 * cx - 2;
    const bool merged = head && x >= 0 && x < W;
    // the group of a head lane ends at the next head of its segment (or the bucket end)
    const uint64_t heads = __ballot(head);
    uint32_t res[NW];
    if (merged) {
        const uint64_t segm = (S == 64) ? ~0ull : (((1ull << S) - 1ull) << sb);
        const uint64_t later = (lane < 63 ? (heads & (~0ull << (lane + 1))) : 0ull) & segm;
        const int e = later ? (int)__builtin_ctzll(later) - sb : n;
        group_mask<NW>(order + sb, meta, msk, p, e, i0, m, w_at, rule, res);
    }
    const int prev_x = 2 * prev_cx - 2;
    const bool span = head && p > 0 && (prev_x + 2 < x) && fill_test(rule, w_at);
    const int cnt = (merged ? 1 : 0) + (span ? 1 : 0);
    const int ic = dpp_scan_seg<S>(cnt, 0, OpAdd());
    const int idx = ic - cnt;
    const size_t ib = 2 * (size_t)base;
    if (span) {
        const int x0 = prev_x + 2;
        A.items[ib + idx] = (uint32_t)x0 | ((uint32_t)(x - x0) << 16);
    }
    if (merged) {
        A.items[ib + idx + (span ? 1 : 0)] = (uint32_t)x;
        if (A.imask) store_mask<NW>(A.imask, ib + idx + (span ? 1 : 0), res);
        A.icov[ib + idx + (span ? 1 : 0)] = cell_counts<SPP>(res);
    }
    const int lo = span ? prev_x + 2 : (merged ? x : INT_MAX);
    const int hi = merged ? x + 2 : (span ? x : INT_MIN);
    const int xmin = dpp_scan_seg<S>(lo, INT_MAX, OpMin());
    const int xmax = dpp_scan_seg<S>(hi, INT_MIN, OpMax());
    const int nmerged = dpp_scan_seg<S>(merged ? 1 : 0, 0, OpAdd());
    const uint64_t tb = tile_bits(lo, hi, xmin >> 6);  // a lane's prefix minimum is its segment's first item
    const uint32_t tlo = (uint32_t)dpp_scan_seg<S>((int)(uint32_t)tb, 0, OpOr());
    const uint32_t thi = (uint32_t)dpp_scan_seg<S>((int)(uint32_t)(tb >> 32), 0, OpOr());
    if (k >= 0 && i == S - 1) {
        // the segment's last lane holds its bucket's totals
        const int64_t b = b0 + (int64_t)k * G;
        A.bicnt[b] = (uint32_t)ic;
        A.bmcnt[b] = (uint32_t)nmerged;
        A.bext[b] = ic ? ((uint32_t)xmin | ((uint32_t)xmax << 16)) : 0xFFFFu;
        A.btile[b] = (uint64_t)tlo | ((uint64_t)thi << 32);
    }
    __syncthreads();
}

// Every bucket: empty ones are cleared, buckets of at most 128 fragments are resolved,
// larger ones are listed for the size-class passes below.
template <int SPP>
__global__ void __launch_bounds__(64, 6) k_resolve_first(ResolveArgs A) {
    constexpr int NW = Words<SPP>::n;
    __shared__ uint32_t key[128], order[128], meta[128], msk[128 * NW];
    const int lane = threadIdx.x;
    const int64_t G = gridDim.x;
    // Strided buckets keep the waves' loads balanced (a path's neighbouring rows have
    // similar sizes); the blocks of one XCD take consecutive b0, so neighbouring buckets,
    // whose records share cache lines, are resolved under one L2.
    for (int64_t b0 = xcd_chunk(blockIdx.x, (uint32_t)G); b0 < A.NB; b0 += 64 * G) {
        // The wave's next 64 buckets b0 + k G: lane k loads their offsets and scalars with
        // vector loads, read back per bucket by readlane.  (Scalar loads would share the
        // LDS wait counter, so every LDS wait of a resolve would also wait for them.)
        const int64_t bl = b0 + (int64_t)lane * G;
        int vbase = 0, vend = 0, vcarry = 0, vrule = 0, vW = A.W;
        if (bl < A.NB) {
            const int2 info = A.binfo[bl];
            vbase = (int)A.boff[bl];
            // the out-of-canvas bucket holds records only for the winding carry
            vend = info.y < 0 ? vbase : (int)A.boff[bl + 1];
            vrule = A.frule[info.x];
            uint32_t carry = A.bcarry[bl];
            if (A.pcanvas) {
                const int4 cv = A.pcanvas[info.x];
                vW = cv.x;
                carry -= A.scarry[cv.w];
            }
            vcarry = (int)carry;
        }
        const int64_t left = (A.NB - b0 + G - 1) / G;
        const int nk = (int)(left < 64 ? left : 64);
        // buckets of up to 16 fragments four to a wave, of 17..32 two to a wave
        const int myn = lane < nk ? vend - vbase : -1;
        uint64_t m16 = __ballot(myn > 0 && myn <= 16);
        uint64_t m32 = __ballot(myn > 16 && myn <= 32);
        uint64_t rest = __ballot(lane < nk) & ~(m16 | m32);
        while (m16) {
            int ks[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                ks[s] = m16 ? (int)__builtin_ctzll(m16) : -1;
                m16 &= m16 - 1ull;
            }
            resolve_pack<SPP, 16>(A, b0, G, ks, vbase, vend, vcarry, vrule, vW, order, meta, msk);
        }
        while (m32) {
            int ks[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                ks[s] = m32 ? (int)__builtin_ctzll(m32) : -1;
                m32 &= m32 - 1ull;
            }
            resolve_pack<SPP, 32>(A, b0, G, ks, vbase, vend, vcarry, vrule, vW, order, meta, msk);
        }
        // the rest one by one, the next bucket's records loading during the current one
        SmallIn<NW> in0, in1;
        if (rest) {
            const int k0 = (int)__builtin_ctzll(rest);
            small_load<NW>(A, (uint32_t)__builtin_amdgcn_readlane(vbase, k0),
                           (uint32_t)__builtin_amdgcn_readlane(vend, k0), in0);
        }
        while (rest) {
            const int k = (int)__builtin_ctzll(rest);
            rest &= rest - 1ull;
            const int64_t b = b0 + (int64_t)k * G;
            const uint32_t base = (uint32_t)__builtin_amdgcn_readlane(vbase, k);
            const int n = __builtin_amdgcn_readlane(vend, k) - (int)base;
            if (rest) {
                const int kn = (int)__builtin_ctzll(rest);
                small_load<NW>(A, (uint32_t)__builtin_amdgcn_readlane(vbase, kn),
                               (uint32_t)__builtin_amdgcn_readlane(vend, kn), in1);
            }
            if (n == 0) {
                if (lane == 0) { A.bicnt[b] = 0; A.bmcnt[b] = 0; A.bext[b] = 0xFFFFu; A.btile[b] = 0ull; }
            } else if (n > 128) {
                if (lane == 0) A.big_list[atomicAdd(A.big_count, 1u)] = (uint32_t)b;
            } else if (n > 64) {
                resolve_bucket<SPP, 2>(A, b, base, n, key, order, meta, msk);
            } else {
                in0.carry = (uint32_t)__builtin_amdgcn_readlane(vcarry, k);
                in0.rule = __builtin_amdgcn_readlane(vrule, k);
                in0.W = __builtin_amdgcn_readlane(vW, k);
                resolve_small<SPP>(A, b, base, n, in0, order, meta, msk);
            }
            in0 = in1;
        }
    }
}

// The listed buckets (more than 128 fragments), one per block, in one launch so that the
// few long ones overlap: up to 1024 fragments staged in LDS, more in k_resolve_huge.
constexpr int kBigStage = 16;  // chunks of 64 fragments staged in LDS

template <int SPP>
__global__ void __launch_bounds__(64) k_resolve_big(ResolveArgs A) {
    constexpr int NW = Words<SPP>::n;
    __shared__ uint32_t key[64 * kBigStage], order[64 * kBigStage], meta[64 * kBigStage], msk[64 * kBigStage * NW];
    const uint32_t nbig = *A.big_count;
    for (uint32_t bi = blockIdx.x; bi < nbig; bi += gridDim.x) {
        const int64_t b = A.big_list[bi];
        const uint32_t base = A.boff[b];
        const int n = (int)(A.boff[b + 1] - base);
        if (n <= 256) resolve_bucket<SPP, 4>(A, b, base, n, key, order, meta, msk);
        else if (n <= 512) resolve_bucket<SPP, 8>(A, b, base, n, key, order, meta, msk);
        else if (n <= 64 * kBigStage) resolve_bucket<SPP, kBigStage>(A, b, base, n, key, order, meta, msk);
    }
}

// Buckets of more than 1024 fragments (rare): bitonic-sort 64-bit (cell x, index) keys,
// their high words in the bucket's stretch of the global scratch and their low words in
// its stretch of the item array (2n words each; the padded power of two is at most 2n),
// copy the sorted indices into the scratch (the items are written over next), then walk
// the sorted order in chunks of 64 reading the records from global memory.  Any bucket
// size works.
template <int SPP>
__global__ void __launch_bounds__(64) k_resolve_huge(ResolveArgs A) {
    constexpr int NW = Words<SPP>::n;
    const int lane = threadIdx.x;
    const uint32_t nbig = *A.big_count;
    for (uint32_t bi = blockIdx.x; bi < nbig; bi += gridDim.x) {
        const int64_t b = A.big_list[bi];
        const uint32_t base = A.boff[b];
        const int n = (int)(A.boff[b + 1] - base);
        if (n <= 64 * kBigStage) continue;
        int P2 = 1;
        while (P2 < n) P2 <<= 1;
        uint32_t *khi = A.scratch + 2 * (size_t)base;
        uint32_t *klo = A.items + 2 * (size_t)base;
        for (int i = lane; i < P2; i += 64) {
            khi[i] = (i < n) ? (A.rec_meta[base + i] & 0xFFFFu) : 0xFFFFFFFFu;
            klo[i] = (i < n) ? (uint32_t)i : 0xFFFFFFFFu;
        }
        __syncthreads();
        for (int k = 2; k <= P2; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = lane; i < P2; i += 64) {
                    const int l = i ^ j;
                    if (l > i) {
                        const uint32_t ah = khi[i], al = klo[i], ch = khi[l], cl = klo[l];
                        const bool gt = ah > ch || (ah == ch && al > cl);
                        const bool up = (i & k) == 0;
                        if (gt == up) {
                            khi[i] = ch; klo[i] = cl;
                            khi[l] = ah; klo[l] = al;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (int i = lane; i < n; i += 64) khi[i] = klo[i];
        __syncthreads();
        const uint32_t *order = khi;  // sorted fragment indices; a fragment's cell x is in its record
        auto cell_x = [&](int q) { return (int)(A.rec_meta[base + order[q]] & 0xFFFFu); };
        const int rule = A.frule[A.binfo[b].x];
        uint32_t wcarry = A.bcarry[b];
        int W = A.W;
        if (A.pcanvas) {
            const int4 cv = A.pcanvas[A.binfo[b].x];
            W = cv.x;
            wcarry -= A.scarry[cv.w];
        }
        int icarry = 0, nmerged = 0, xmin = INT_MAX, xmax = INT_MIN;
        uint64_t tmask = 0ull;
        const size_t ib = 2 * (size_t)base;
        for (int c0 = 0; c0 < n; c0 += 64) {
            const int p = c0 + lane;
            const bool v = p < n;
            const uint32_t meta = v ? A.rec_meta[base + order[p]] : 0u;
            const int cx = (int)(meta & 0xFFFFu);
            const int swn = v ? (int)(int8_t)(meta >> 16) : 0;
            const int prev_cx = (v && p > 0) ? cell_x(p - 1) : -1;
            const bool head = v && (p == 0 || cx != prev_cx);
            const int incl = scan_add(swn);
            const int8_t w_at = (int8_t)(int)(wcarry + (uint32_t)(incl - swn));
            wcarry += (uint32_t)lane63(incl);
            const int x = 2 * cx - 2;
            const bool merged = head && x >= 0 && x < W;
            uint32_t res[NW];
            if (merged) {
                Stencil<NW> st;
                st.init(w_at);
                for (int q = p; q < n; ++q) {
                    const uint32_t j = order[q];
                    const uint32_t mq = A.rec_meta[base + j];
                    if ((int)(mq & 0xFFFFu) != cx) break;
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
                if (A.imask) store_mask<NW>(A.imask, ib + idx + (span ? 1 : 0), res);
                A.icov[ib + idx + (span ? 1 : 0)] = cell_counts<SPP>(res);
            }
            const int lo = span ? prev_x + 2 : (merged ? x : INT_MAX);
            const int hi = merged ? x + 2 : (span ? x : INT_MIN);
            xmin = tmin(xmin, lane63(dpp_scan(lo, INT_MAX, OpMin())));
            xmax = tmax(xmax, lane63(dpp_scan(hi, INT_MIN, OpMax())));
            tmask |= wave_or64(tile_bits(lo, hi, xmin >> 6));
            nmerged += lane63(scan_add(merged ? 1 : 0));
        }
        if (lane == 0) {
            A.bicnt[b] = (uint32_t)icarry;
            A.bmcnt[b] = (uint32_t)nmerged;
            A.bext[b] = icarry ? ((uint32_t)xmin | ((uint32_t)xmax << 16)) : 0xFFFFu;
            A.btile[b] = tmask;
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
                             const int2 *__restrict__ binfo, const int4 *__restrict__ pcanvas, uint32_t *keys,
                             uint32_t *vals) {
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= NB || !flag[b]) return;
    const uint32_t o = off[b];
    const int2 bi = binfo[b];
    keys[o] = (uint32_t)bi.y + (pcanvas ? (uint32_t)pcanvas[bi.x].z : 0u);
    vals[o] = (uint32_t)b;
}

// Winding-scan carry at the first bucket of every scene of a batch.
__global__ void k_scene_carry(int n, const uint32_t *__restrict__ sfirst, const uint32_t *__restrict__ pbase,
                              const uint32_t *__restrict__ bcarry, uint32_t *scarry) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) scarry[i] = bcarry[pbase[sfirst[i]]];
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
// gradient paint (output_32.vert.glsl:55-97, output_32.frag.glsl:45-51; the multisample
// shaders differ only in the x offset of the line ends).  The reference evaluates the
// gradient coordinates at the two ends of each item's line, (x0 + off, y + 1/2) and
// (x0 + w + off, y + 1/2) per pixel row, the rasterizer interpolates them linearly to the
// pixel centre (t = (x - x0 + 1/2) / w), and the fragment shader maps them to the ramp
// parameter c = min(|d|^2 / (sqrt(B^2 + |d|^2 f.z) - B), 1), d = g - f, B = d . f, and
// samples the 1024-wide ramp with linear filtering and clamped edges, sRGB-decoding the
// texels when sRGB correction is on.  The same float operations, in the same order, are
// restated in the oracle.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ float2 grad_coord(const GradDev &G, const float4 &w0, const float4 &w1, float sx, float sy) {
    const float *k = G.ctx;
    const float rdx = (k[0] * sx + k[3] * sy) + k[6];
    const float rdy = (k[1] * sx + k[4] * sy) + k[7];
    const float rdz = (k[2] * sx + k[5] * sy) + k[8];
    const float q = k[12] / rdz;
    const float ox = k[9] + q * rdx, oy = k[10] + q * rdy;
    return make_float2((ox * w0.x + oy * w0.y) + w0.z, (ox * w0.w + oy * w1.x) + w1.y);
}

// A fill word without alpha names gradient slot c - 1; it paints when the slot's ramp scale
// is positive (the shaders' gradient branch), and draws nothing otherwise.
__device__ __forceinline__ bool grad_paints(const GradDev &G, uint32_t c) {
    const uint32_t gi = c - 1u;
    return gi < G.nslots && G.table[3 * gi + 1].w > 0.f;
}

__device__ __forceinline__ float4 ramp_texel(const GradDev &G, int i, int j, const float *dc, const float *da) {
    const uint32_t t = G.ramp[(size_t)j * 1024 + i];
    return make_float4(dc[t & 0xFFu], dc[(t >> 8) & 0xFFu], dc[(t >> 16) & 0xFFu], da[t >> 24]);
}

// Paint of gradient word c at pixel (x, yrow) of an item covering columns [x0, x0 + w) of
// the cell row holding yrow.  dc / da decode a texel's colour and alpha codes.
__device__ __forceinline__ float4 grad_color(const GradDev &G, uint32_t c, int x0, int w, int x, int yrow, const float *dc,
                             const float *da) {
    const uint32_t gi = c - 1u;
    const float4 w0 = G.table[3 * gi], w1 = G.table[3 * gi + 1], w2 = G.table[3 * gi + 2];
    const float sy = (float)(yrow & ~1) + ((yrow & 1) ? 1.5f : 0.5f);
    const float2 ga = grad_coord(G, w0, w1, (float)x0 + G.xoff, sy);
    const float2 gb = grad_coord(G, w0, w1, (float)(x0 + w) + G.xoff, sy);
    const float t = ((float)(x - x0) + 0.5f) / (float)w;
    const float omt = 1.f - t;
    const float gx = ga.x * omt + gb.x * t, gy = ga.y * omt + gb.y * t;
    const float dx = gx - w2.x, dy = gy - w2.y;
    const float A = dx * dx + dy * dy, B = dx * w2.x + dy * w2.y;
    const float den = sqrtf(B * B + A * w2.z) - B;
    float cc = A * (den > 1e-6f ? 1.f / den : 0.f);
    cc = cc < 1.f ? cc : 1.f;
    if (G.rows <= 0) return make_float4(0.f, 0.f, 0.f, 1.f);
    const int ri = __float_as_int(w1.z);
    const float u = ((float)(ri & 1023) + 0.5f) * (1.f / 1024.f) + cc * w1.w;
    const float v = ((float)(ri >> 10) + 0.5f) * G.irh;
    const float tu = u * 1024.f - 0.5f, tv = v * (float)G.rows - 0.5f;
    const float fu = floorf(tu), fv = floorf(tv);
    const float au = tu - fu, av = tv - fv, bu = 1.f - au, bv = 1.f - av;
    const int iu = (int)fu, jv = (int)fv;
    const int i0 = tmin(tmax(iu, 0), 1023), i1 = tmin(tmax(iu + 1, 0), 1023);
    const int j0 = tmin(tmax(jv, 0), G.rows - 1), j1 = tmin(tmax(jv + 1, 0), G.rows - 1);
    const float4 t00 = ramp_texel(G, i0, j0, dc, da), t10 = ramp_texel(G, i1, j0, dc, da);
    const float4 t01 = ramp_texel(G, i0, j1, dc, da), t11 = ramp_texel(G, i1, j1, dc, da);
    float4 o;
    o.x = (t00.x * bu + t10.x * au) * bv + (t01.x * bu + t11.x * au) * av;
    o.y = (t00.y * bu + t10.y * au) * bv + (t01.y * bu + t11.y * au) * av;
    o.z = (t00.z * bu + t10.z * au) * bv + (t01.z * bu + t11.z * au) * av;
    o.w = (t00.w * bu + t10.w * au) * bv + (t01.w * bu + t11.w * au) * av;
    return o;
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------------------------
// coverage-mode compositor.  The framebuffer state of each pixel is carried as the four
// decoded floats of its 8-bit codes; every blend re-quantises to a code and decodes it
// again, exactly as an 8-bit framebuffer does:
//   UNORM8: code = largest k with v >= (k - 0.5) / 255, i.e. with 510 v - (2k - 1) >= 0 --
//           one fma gives that sign exactly, applied to an estimate that is never too
//           small; decode = RN(k / 255) by a corrected product.
//   sRGB8:  a 4096-entry table gives the code k0 of the interval start floor(4096 v)/4096;
//           no interval holds two thresholds, so the code is k0 or k0 + 1, and one 16-byte
//           step entry {thr[k0+1], dec[k0], dec[k0+1]} decides and decodes.
// ---------------------------------------------------------------------------------
// The code is m = floor(255 v + 0.5) (clamped).  The estimate trunc(RN(255 v + 0.5)) is m or
// m + 1: rounding is monotone and the integers m, m + 1 are floats, so RN never crosses below
// m.  One exact sign test, RN(510 v + 1 - 2k) < 0 iff v < thr[k], takes the extra one back.
__device__ __forceinline__ int unorm_code(float v) {
    int k = tmin(tmax((int)__builtin_fmaf(v, 255.f, 0.5f), 0), 255);
    k -= (__builtin_fmaf(v, 510.f, (float)(1 - 2 * k)) < 0.f) ? 1 : 0;  // v < thr[k]
    return tmax(k, 0);
}

__device__ __forceinline__ float unorm_dec(int k) {
    const float fk = (float)k, r = 1.f / 255.f;
    const float q = fk * r;
    return __builtin_fmaf(__builtin_fmaf(-q, 255.f, fk), r, q);
}

// lut[i] is the code of i/4096; at most one threshold lies in each 1/4096 interval, so the
// code of v is lut[i] or lut[i] + 1 for i = floor(4096 v).
__device__ __forceinline__ int srgb_code(const float4 *step, const uint8_t *lut, float v, float &dec) {
    const int i = tmin(tmax((int)(v * 4096.f), 0), 4095);
    const int k0 = lut[i];
    const float4 st = step[k0];
    const bool up = v >= st.x;
    dec = up ? st.z : st.y;
    return k0 + (up ? 1 : 0);
}

template <bool SRGB>
__device__ __forceinline__ void blend_dec(const float4 *step, const uint8_t *lut, float *d, const float4 &src, float a) {
    const float oma = 1.f - a;
    const float v0 = src.x * a + d[0] * oma;
    const float v1 = src.y * a + d[1] * oma;
    const float v2 = src.z * a + d[2] * oma;
    const float va = a * a + d[3] * oma;
    if (SRGB) {
        srgb_code(step, lut, v0, d[0]);
        srgb_code(step, lut, v1, d[1]);
        srgb_code(step, lut, v2, d[2]);
    } else {
        d[0] = unorm_dec(unorm_code(v0));
        d[1] = unorm_dec(unorm_code(v1));
        d[2] = unorm_dec(unorm_code(v2));
    }
    d[3] = unorm_dec(unorm_code(va));
}

template <bool SRGB>
__device__ __forceinline__ uint32_t pack_dec(const float4 *step, const uint8_t *lut, const float *d) {
    uint32_t o = 0u;
    float t;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) o |= (uint32_t)(SRGB ? srgb_code(step, lut, d[ch], t) : unorm_code(d[ch])) << (8 * ch);
    return o | ((uint32_t)unorm_code(d[3]) << 24);
}

struct CovArgs {
    int W, H, tiles_x;
    const uint32_t *rowbeg, *rowend;
    const uint4 *rrec;          // per row-list entry: extent, colour, item count, item base
    const uint64_t *rtile;      // per row-list entry: 64-pixel tiles holding items
    const uint32_t *items, *icov;
    const uint32_t *imask;      // merged masks (multisample output)
    const uint32_t *rpath;      // row-list path ids (multisample output)
    const uint32_t *ptab;       // per-path blend tables (multisample output)
    const float *dec_c;         // [256] colour decode, for the paint colours
    const float *dec_a;         // [256] alpha decode (gradient ramp texels)
    const float4 *step_c;       // [256] sRGB steps {thr[k+1], dec[k], dec[k+1], 0}
    const uint8_t *lut_c;       // [4096] sRGB code of i/4096
    uint8_t *image;
    int pitch;
    GradDev G;
    // batched scenes: per scene {W, H, first tile, first cell row} and image byte offset
    int nscenes;
    int64_t ntiles;
    const int4 *stile;
    const uint64_t *simg;
};

constexpr int kCovList = 64;    // per-wave list of the tile's buckets

// Scatter the starts of a long bucket's items (more than 64) into the column slots.
__device__ __noinline__ void scatter_big(const uint32_t *items, uint32_t ib, int cnt, int tx0, uint32_t *slot) {
    const int lane = threadIdx.x & 63;
    for (int ic = 0; ic < cnt; ic += 64) {
        if (ic + lane < cnt) {
            const uint32_t it = items[ib + ic + lane];
            const int x0 = (int)(it & 0xFFFFu), iw = (int)(it >> 16);
            const int end = x0 + (iw ? iw : 2);
            if (end > tx0 && x0 < tx0 + 64) atomicMax(&slot[tmax(x0 - tx0, 0)], (uint32_t)(ic + lane + 1));
        }
    }
}

// One wavefront per 64x2 pixel tile (lane = pixel column, rows 2r and 2r+1).  The cell
// row's non-empty buckets, in path order, are filtered by extent into a per-wave list
// (paint colour decoded, item base and count).  The list is resolved kCovGroup buckets at
// a time: each item writes its index into the LDS slot of its first column in the tile
// (items of a bucket are disjoint and sorted by x), an inclusive max-scan over the lanes
// then names each column's covering candidate, and the lanes blend in path order:
// coverage alpha = paint alpha * covered samples / N (output_{8,32}.frag.glsl).
// The kernel is bound by VALU issue, LDS and load latency together; eight waves per SIMD
// (registers capped at 64, a few spilled) hide the latency best.  Groups of 8 buckets with
// the sRGB tables take about 23 KiB of LDS a workgroup, so six workgroups (six waves per
// SIMD) fit in the CU's 160 KiB, and they get the registers of that occupancy (measured
// on the 4096-scene batch: 32.7 ms a render against 35.5 ms when capped at four).
template <int SPP, bool SRGB, int kCovGroup, bool BATCH, bool GRAD>
__global__ void __launch_bounds__(256, (kCovGroup == 4 || !SRGB) ? 8 : 6) k_composite_cov(CovArgs A) {
    __shared__ float4 s_step[SRGB ? 256 : 1];
    __shared__ uint8_t s_lut[SRGB ? 4096 : 4];
    __shared__ float s_dec[256];
    __shared__ float s_deca[GRAD ? 256 : 1];
    __shared__ float4 l_src[4][kCovList];
    __shared__ uint2 l_ref[4][kCovList];
    __shared__ uint32_t slot[4][kCovGroup][64];
    for (int i = threadIdx.x; i < 256; i += 256) {
        if (SRGB) s_step[i] = A.step_c[i];
        s_dec[i] = A.dec_c[i];
        if (GRAD) s_deca[i] = A.dec_a[i];
    }
    for (int i = threadIdx.x; i < (SRGB ? 1024 : 0); i += 256) ((uint32_t *)s_lut)[i] = ((const uint32_t *)A.lut_c)[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int W = A.W, H = A.H, tiles_x = A.tiles_x, row0 = 0, pitch = A.pitch;
    int tile;
    if (!BATCH) {
        // Workgroup b runs on XCD b % 8 (round-robin deal): XCD x takes the cell rows
        // x, x + 8, x + 16, ... whole, its workgroups walking them tile by tile, so each
        // row's items are fetched into one XCD's L2 and every XCD gets rows from the whole
        // image (the per-row work is uneven, so contiguous bands would not balance).
        const int xcd = (int)(blockIdx.x & 7u), t = (int)(blockIdx.x >> 3) * 4 + w;
        const int rr = t / tiles_x, row = xcd + 8 * rr;
        if (row >= (H + 1) / 2) return;
        tile = row * tiles_x + (t - rr * tiles_x);
    } else {
        tile = (int)blockIdx.x * 4 + w;
        if (tile >= A.ntiles) return;
    }
    uint8_t *image = A.image;
    if (BATCH) {
        // the scene whose tile range holds this tile: last one with first tile <= tile
        int lo = 0, hi = A.nscenes - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (A.stile[mid].z <= tile) lo = mid; else hi = mid - 1;
        }
        const int4 st = A.stile[lo];
        W = st.x;
        H = st.y;
        tiles_x = (W + 63) >> 6;
        tile -= st.z;
        row0 = st.w;
        pitch = 4 * W;
        image = A.image + A.simg[lo];
    }
    const int r = tile / tiles_x;
    const int tx0 = (tile - r * tiles_x) * 64;
    const int tx1 = tmin(tx0 + 64, W);
    const int x = tx0 + lane;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    float d0[4] = {1.f, 1.f, 1.f, 1.f}, d1[4] = {1.f, 1.f, 1.f, 1.f};

    auto flush = [&](int nlist) {
        wave_sync_lds();
        for (int g0 = 0; g0 < nlist; g0 += kCovGroup) {
            const int ng = tmin(kCovGroup, nlist - g0);
            uint32_t it[kCovGroup], cv[kCovGroup];
#pragma unroll
            for (int g = 0; g < kCovGroup; ++g) {
                slot[w][g][lane] = 0u;
                const uint2 ref = g < ng ? l_ref[w][g0 + g] : make_uint2(0u, 0u);
                const bool ld = lane < (int)ref.y;
                it[g] = ld ? A.items[ref.x + lane] : 0u;
                cv[g] = ld ? A.icov[ref.x + lane] : 0u;
            }
            wave_sync_lds();
#pragma unroll
            for (int g = 0; g < kCovGroup; ++g) {
                if (g >= ng) break;
                const uint2 ref = l_ref[w][g0 + g];
                if (ref.y > 64u) {
                    scatter_big(A.items, ref.x, (int)ref.y, tx0, slot[w][g]);
                } else if (lane < (int)ref.y) {
                    const int x0 = (int)(it[g] & 0xFFFFu), iw = (int)(it[g] >> 16);
                    const int end = x0 + (iw ? iw : 2);
                    if (end > tx0 && x0 < tx0 + 64) atomicMax(&slot[w][g][tmax(x0 - tx0, 0)], (uint32_t)(lane + 1));
                }
            }
            wave_sync_lds();
            uint32_t sv[kCovGroup];
#pragma unroll
            for (int g = 0; g < kCovGroup; ++g) sv[g] = slot[w][g][lane];
#pragma unroll
            for (int g = 0; g < kCovGroup; ++g) sv[g] = (uint32_t)dpp_scan((int)sv[g], 0, OpMax());
            // Each lane lists the group's buckets covering its column, in path order, as
            // {n0, n1, g}; the blend passes then run as many times as the deepest lane's
            // list, each lane blending its own next entry (same per-pixel order as one
            // pass per bucket, fewer passes where the buckets do not overlap).
            uint32_t ent[kCovGroup], ent2[GRAD ? kCovGroup : 1];
            int nent = 0;
#pragma unroll
            for (int g = 0; g < kCovGroup; ++g) {
                ent[g] = 0u;
                if (g >= ng) break;
                const uint2 ref = l_ref[w][g0 + g];
                const int p = (int)sv[g] - 1;
                uint32_t pit, pcv;
                if (ref.y > 64u) {
                    pit = p >= 0 ? A.items[ref.x + p] : 0u;
                    pcv = p >= 0 ? A.icov[ref.x + p] : 0u;
                } else {
                    pit = (uint32_t)__shfl((int)it[g], tmax(p, 0), 64);
                    pcv = (uint32_t)__shfl((int)cv[g], tmax(p, 0), 64);
                }
                const int px0 = (int)(pit & 0xFFFFu), pw = (int)(pit >> 16);
                const int pend = px0 + (pw ? pw : 2);
                const bool covered = p >= 0 && x < pend && x < W;
                uint32_t n0 = SPP, n1 = SPP;
                if (pw == 0) {
                    n0 = (pcv >> (6 * ((x - px0) * 2))) & 63u;
                    n1 = (pcv >> (6 * ((x - px0) * 2 + 1))) & 63u;
                }
                const uint32_t e = n0 | (n1 << 8) | ((uint32_t)g << 16);
                const uint32_t e2 = (uint32_t)px0 | ((uint32_t)(pw ? pw : 2) << 16);
#pragma unroll
                for (int j = 0; j <= g; ++j) {
                    ent[j] = (covered && nent == j) ? e : ent[j];
                    if (GRAD) ent2[j] = (covered && nent == j) ? e2 : ent2[j];
                }
                nent += covered ? 1 : 0;
            }
#pragma unroll
            for (int j = 0; j < kCovGroup; ++j) {
                if (!__ballot(nent > j)) break;
                if (nent > j) {
                    const uint32_t e = ent[j];
                    const float4 src = l_src[w][g0 + (int)(e >> 16)];
                    if (GRAD && src.w < 0.f) {
                        // a gradient paint: its colour at each of the lane's two pixels
                        const uint32_t c = __float_as_uint(src.x);
                        const int ix0 = (int)(ent2[j] & 0xFFFFu), iw = (int)(ent2[j] >> 16);
                        const float4 p0 = grad_color(A.G, c, ix0, iw, x, 2 * r, s_dec, s_deca);
                        const float4 p1 = grad_color(A.G, c, ix0, iw, x, 2 * r + 1, s_dec, s_deca);
                        blend_dec<SRGB>(s_step, s_lut, d0, p0, p0.w * ((float)(e & 63u) / (float)SPP));
                        blend_dec<SRGB>(s_step, s_lut, d1, p1, p1.w * ((float)((e >> 8) & 63u) / (float)SPP));
                    } else {
                        blend_dec<SRGB>(s_step, s_lut, d0, src, src.w * ((float)(e & 63u) / (float)SPP));
                        blend_dec<SRGB>(s_step, s_lut, d1, src, src.w * ((float)((e >> 8) & 63u) / (float)SPP));
                    }
                }
            }
            wave_sync_lds();
        }
    };

    const uint32_t q0 = A.rowbeg[row0 + r], q1 = A.rowend[row0 + r];
    int nlist = 0;
    for (uint32_t qc = q0; qc < q1; qc += 64) {
        const bool inq = qc + lane < q1;
        const uint4 rec = inq ? A.rrec[qc + lane] : make_uint4(0xFFFFu, 0u, 0u, 0u);
        const uint64_t rt = inq ? A.rtile[qc + lane] : 0ull;
        const int exlo = (int)(rec.x & 0xFFFFu), exhi = (int)(rec.x >> 16);
        // a word without alpha paints nothing, unless it names a gradient that paints
        const bool paints = (rec.y & 0xFF000000u) || (GRAD && rec.y != 0u && grad_paints(A.G, rec.y));
        const bool ov = inq && exlo < tx1 && exhi > tx0 && ((rt >> tile_bit(tx0, exlo)) & 1ull) && paints;
        const unsigned long long om = __ballot(ov);
        if (!om) continue;
        const int nov = (int)__popcll(om);
        if (nlist + nov > kCovList) {
            flush(nlist);
            nlist = 0;
        }
        if (ov) {
            const int sl = nlist + (int)__popcll(om & lt);
            const uint32_t c = rec.y;
            if (GRAD && !(c & 0xFF000000u))
                l_src[w][sl] = make_float4(__uint_as_float(c), 0.f, 0.f, -1.f);
            else
                l_src[w][sl] = make_float4(s_dec[c & 0xFFu], s_dec[(c >> 8) & 0xFFu], s_dec[(c >> 16) & 0xFFu],
                                           unorm_dec((int)(c >> 24)));
            l_ref[w][sl] = make_uint2(rec.w, rec.z);
        }
        nlist += nov;
    }
    if (nlist) flush(nlist);

    if (x >= W) return;
    const int y0 = 2 * r;
    *(uint32_t *)(image + (size_t)y0 * pitch + (size_t)x * 4) = pack_dec<SRGB>(s_step, s_lut, d0);
    if (y0 + 1 < H) *(uint32_t *)(image + (size_t)(y0 + 1) * pitch + (size_t)x * 4) = pack_dec<SRGB>(s_step, s_lut, d1);
}

// ---------------------------------------------------------------------------------
// multisample compositor (ms_output_{8,32}.frag.glsl: every covered sample of an N-sample
// 8-bit framebuffer is blended src-over and re-quantised; ras_base.cpp:1148-1163 then
// resolves each pixel to the average of its samples' decoded values, summed in sample
// order).  The bucket lookup is the coverage compositor's (lane = pixel column); the
// blending runs with lane = sample: lane q * 2N + dy * N + s holds sample s of pixel
// (column p * CPR + q, row 2r + dy) of the 64x2 tile in state register p, CPR = 64 / 2N
// columns per register.  Each register that holds a pixel covered by the bucket costs one
// wave-wide blend (its lanes select by their sample bits), so a 2x2 fragment or a span
// column costs the same whatever its sample count, and no lane loops over samples.  Solid
// paints run over the state registers in static blocks of four registers at 32 samples,
// two at 16 and 8 (apply_blocks): a register's blend is four byte lookups in the path's
// table and one select whose lane mask is read straight from the column lanes' sample
// bits, with no register indexing.
// ---------------------------------------------------------------------------------
typedef uint32_t u32x32 __attribute__((ext_vector_type(32)));
typedef uint32_t u32x16 __attribute__((ext_vector_type(16)));

// Runs f(state, p) -> new state on every register of the BLK-register blocks of vec that
// hold a set bit of m (p = pbase + register), with static register numbers: the state
// stays in plain VGPRs (no indexing, no per-register loop), and a block's LDS round trips
// overlap, at the price of the block's uncovered registers (f leaves a register it does not
// cover unchanged).
template <int N, int BLK, class V, class F>
__device__ __forceinline__ void apply_blocks(V &vec, uint32_t m, int pbase, const F &f) {
#pragma unroll
    for (int b0 = 0; b0 < N; b0 += BLK) {
        if ((m >> b0) & ((1u << BLK) - 1u)) {
#pragma unroll
            for (int i = 0; i < BLK; ++i) vec[b0 + i] = f(vec[b0 + i], pbase + b0 + i);
        }
    }
}

// Runs f(state, p) -> new state on every register p of the set bits of m (p = pbase + bit)
// by index: the register number is wave-uniform, so the indexed state stays in VGPRs
// (s_set_gpr_idx).  The gradient variants use these, since their paint arithmetic per
// register is large (the static blocks would spend it on uncovered registers too) and a
// function that mixes indexed and static access to the state spills it.  The 4-way form
// takes four registers an iteration so that their LDS round trips overlap; missing ones
// repeat the first register, whose duplicate writes store the same value, so every write
// is unconditional.  A write that depends on control flow inside the loop makes the
// compiler copy the whole 32-register vector at every iteration, and so does a loop that
// writes two vectors: each call covers one.
template <class V, class F>
__device__ __forceinline__ void apply_regs4(V &vec, uint32_t m, int pbase, const F &f) {
    while (m) {
        const int c0 = __builtin_amdgcn_readfirstlane((int)__builtin_ctz(m));
        m &= m - 1u;
        const int c1 = m ? __builtin_amdgcn_readfirstlane((int)__builtin_ctz(m)) : c0;
        m &= m - 1u;
        const int c2 = m ? __builtin_amdgcn_readfirstlane((int)__builtin_ctz(m)) : c0;
        m &= m - 1u;
        const int c3 = m ? __builtin_amdgcn_readfirstlane((int)__builtin_ctz(m)) : c0;
        m &= m - 1u;
        const uint32_t o0 = vec[c0], o1 = vec[c1], o2 = vec[c2], o3 = vec[c3];
        const uint32_t n0 = f(o0, pbase + c0), n1 = f(o1, pbase + c1), n2 = f(o2, pbase + c2),
                       n3 = f(o3, pbase + c3);
        vec[c3] = n3;
        vec[c2] = n2;
        vec[c1] = n1;
        vec[c0] = n0;
    }
}

template <class V, class F>
__device__ __forceinline__ void apply_regs1(V &vec, uint32_t m, int pbase, const F &f) {
    while (m) {
        const int c = __builtin_amdgcn_readfirstlane((int)__builtin_ctz(m));
        m &= m - 1u;
        vec[c] = f(vec[c], pbase + c);
    }
}

template <int NREG> struct SampleRegs;
template <> struct SampleRegs<64> {
    u32x32 a, b;
    __device__ __forceinline__ void init(uint32_t v) {
#pragma unroll
        for (int i = 0; i < 32; ++i) { a[i] = v; b[i] = v; }
    }
    __device__ __forceinline__ uint32_t get(int p) const { return p < 32 ? a[p] : b[p - 32]; }
    template <int BLK, class F> __device__ __forceinline__ void apply(uint64_t m, const F &f) {
        apply_blocks<32, BLK>(a, (uint32_t)m, 0, f);
        apply_blocks<32, BLK>(b, (uint32_t)(m >> 32), 32, f);
    }
    template <class F> __device__ __forceinline__ void apply4(uint64_t m, const F &f) {
        apply_regs4(a, (uint32_t)m, 0, f);
        apply_regs4(b, (uint32_t)(m >> 32), 32, f);
    }
    template <class F> __device__ __forceinline__ void apply1(uint64_t m, const F &f) {
        apply_regs1(a, (uint32_t)m, 0, f);
        apply_regs1(b, (uint32_t)(m >> 32), 32, f);
    }
};
template <> struct SampleRegs<32> {
    u32x32 a;
    __device__ __forceinline__ void init(uint32_t v) {
#pragma unroll
        for (int i = 0; i < 32; ++i) a[i] = v;
    }
    __device__ __forceinline__ uint32_t get(int p) const { return a[p]; }
    template <int BLK, class F> __device__ __forceinline__ void apply(uint64_t m, const F &f) {
        apply_blocks<32, BLK>(a, (uint32_t)m, 0, f);
    }
    template <class F> __device__ __forceinline__ void apply4(uint64_t m, const F &f) { apply_regs4(a, (uint32_t)m, 0, f); }
    template <class F> __device__ __forceinline__ void apply1(uint64_t m, const F &f) { apply_regs1(a, (uint32_t)m, 0, f); }
};
template <> struct SampleRegs<16> {
    u32x16 a;
    __device__ __forceinline__ void init(uint32_t v) {
#pragma unroll
        for (int i = 0; i < 16; ++i) a[i] = v;
    }
    __device__ __forceinline__ uint32_t get(int p) const { return a[p]; }
    template <int BLK, class F> __device__ __forceinline__ void apply(uint64_t m, const F &f) {
        apply_blocks<16, BLK>(a, (uint32_t)m, 0, f);
    }
    template <class F> __device__ __forceinline__ void apply4(uint64_t m, const F &f) { apply_regs4(a, (uint32_t)m, 0, f); }
    template <class F> __device__ __forceinline__ void apply1(uint64_t m, const F &f) { apply_regs1(a, (uint32_t)m, 0, f); }
};

// Blend tables of the solid paints, one per path, built once per frame: byte ch * 256 + k
// of a path's 1 KiB table holds the code that channel ch takes when the path's paint
// (decoded colour, alpha a) is blended over code k -- blend_code's arithmetic per channel.
// Lane l of a path's wave computes codes 4l .. 4l + 3 of every channel.  Words without
// alpha (gradients, no paint) get no table.
template <bool SRGB>
__global__ void __launch_bounds__(256) k_paint_tables(uint32_t P, const uint32_t *__restrict__ finfo,
                                                      const float *__restrict__ dec, const float4 *__restrict__ step,
                                                      const uint8_t *__restrict__ lut, uint32_t *__restrict__ ptab) {
    const uint32_t path = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (path >= P) return;
    const uint32_t c = finfo[path];
    if (!(c & 0xFF000000u)) return;
    const int lane = threadIdx.x & 63;
    const float4 src = make_float4(dec[c & 0xFFu], dec[(c >> 8) & 0xFFu], dec[(c >> 16) & 0xFFu], unorm_dec((int)(c >> 24)));
    const float a = src.w, oma = 1.f - a;
    uint32_t e[4];  // e[ch]: channel ch's codes for k = 4l .. 4l + 3, one byte each
    e[0] = e[1] = e[2] = e[3] = 0u;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int k = 4 * lane + j;
        const float dk = SRGB ? dec[k] : unorm_dec(k);
        const float v0 = src.x * a + dk * oma;
        const float v1 = src.y * a + dk * oma;
        const float v2 = src.z * a + dk * oma;
        const float va = a * a + unorm_dec(k) * oma;
        float t;
        const uint32_t c0 = SRGB ? (uint32_t)srgb_code(step, lut, v0, t) : (uint32_t)unorm_code(v0);
        const uint32_t c1 = SRGB ? (uint32_t)srgb_code(step, lut, v1, t) : (uint32_t)unorm_code(v1);
        const uint32_t c2 = SRGB ? (uint32_t)srgb_code(step, lut, v2, t) : (uint32_t)unorm_code(v2);
        e[0] |= c0 << (8 * j);
        e[1] |= c1 << (8 * j);
        e[2] |= c2 << (8 * j);
        e[3] |= (uint32_t)unorm_code(va) << (8 * j);
    }
    uint32_t *t = ptab + (size_t)path * 256;
#pragma unroll
    for (int ch = 0; ch < 4; ++ch) t[ch * 64 + lane] = e[ch];
}

// A code's blend through a planar table: channel ch's byte indexes bytes ch * 256 .. +255.
__device__ __forceinline__ uint32_t table_code(const uint32_t *T, uint32_t st) {
    const uint8_t *b = (const uint8_t *)T;
    const uint32_t lo = (uint32_t)b[st & 0xFFu] | ((uint32_t)b[256 + ((st >> 8) & 0xFFu)] << 8);
    const uint32_t hi = (uint32_t)b[512 + ((st >> 16) & 0xFFu)] | ((uint32_t)b[768 + (st >> 24)] << 8);
    return __builtin_amdgcn_perm(hi, lo, 0x05040100u);  // the low half-words of lo and hi
}

// One sample's framebuffer code blended with paint src (decoded colour, alpha a) and
// re-quantised: the arithmetic of blend_dec on the decoded code.
template <bool SRGB>
__device__ __forceinline__ uint32_t blend_code(const float4 *step, const uint8_t *lut, const float *dec, uint32_t code,
                                               const float4 &src, float a) {
    const float oma = 1.f - a;
    const uint32_t k0 = code & 0xFFu, k1 = (code >> 8) & 0xFFu, k2 = (code >> 16) & 0xFFu;
    const float d0 = SRGB ? dec[k0] : unorm_dec((int)k0);
    const float d1 = SRGB ? dec[k1] : unorm_dec((int)k1);
    const float d2 = SRGB ? dec[k2] : unorm_dec((int)k2);
    const float da = unorm_dec((int)(code >> 24));
    const float v0 = src.x * a + d0 * oma;
    const float v1 = src.y * a + d1 * oma;
    const float v2 = src.z * a + d2 * oma;
    const float va = a * a + da * oma;
    float t;
    const uint32_t c0 = SRGB ? (uint32_t)srgb_code(step, lut, v0, t) : (uint32_t)unorm_code(v0);
    const uint32_t c1 = SRGB ? (uint32_t)srgb_code(step, lut, v1, t) : (uint32_t)unorm_code(v1);
    const uint32_t c2 = SRGB ? (uint32_t)srgb_code(step, lut, v2, t) : (uint32_t)unorm_code(v2);
    return c0 | (c1 << 8) | (c2 << 16) | ((uint32_t)unorm_code(va) << 24);
}

// Bit i of the result is set when any of bits [i * CPR, (i + 1) * CPR) of e is.
template <int CPR>
__device__ __forceinline__ uint64_t group_bits(uint64_t e) {
    if (CPR == 1) return e;
    if (CPR == 2) {
        uint64_t x = (e | (e >> 1)) & 0x5555555555555555ull;
        x = (x | (x >> 1)) & 0x3333333333333333ull;
        x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
        x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
        x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
        return (x | (x >> 16)) & 0xFFFFFFFFull;
    }
    uint64_t x = (e | (e >> 1) | (e >> 2) | (e >> 3)) & 0x1111111111111111ull;
    x = (x | (x >> 3)) & 0x0303030303030303ull;
    x = (x | (x >> 6)) & 0x000F000F000F000Full;
    x = (x | (x >> 12)) & 0x000000FF000000FFull;
    return (x | (x >> 24)) & 0xFFFFull;
}

// The sample bits of the two pixels of column offset dx in a merged mask (the reference's
// layout: N = 8 one byte per pixel (dx * 2 + dy), N = 16 one half-word per pixel in word dx,
// N = 32 one word per pixel dx * 2 + dy): row 0 in the low N bits, row 1 above (N = 32:
// row 1 in `hi`).
template <int SPP>
__device__ __forceinline__ uint32_t column_bits(const uint32_t *m, int dx, uint32_t &hi) {
    if (SPP == 32) {
        hi = dx ? m[3] : m[1];
        return dx ? m[2] : m[0];
    }
    hi = 0u;
    if (SPP == 16) return dx ? m[1] : m[0];
    return (m[0] >> (16 * dx)) & 0xFFFFu;
}

// Measured on the headline scene's multisample composite at 32 / 16 / 8 samples: the
// static blocks take 10.0 / 5.6 / 4.15 ms where the indexed four-register loop took
// 13.9 / 7.4 / 4.6 ms (per register it spent 14 VALU, 9 scalar and 5 LDS instructions
// against 10 VALU and 4 LDS, the blocks' idle registers included).  At 32 samples, blocks
// of two or eight registers, one- or four-bucket groups, and three or five waves per SIMD
// were all slower (10.2 / 10.3 / 10.6 / 10.5 / 11.7 / 15.1 ms: five waves spill 176 bytes
// a lane); blocks of four cost 5.7 / 4.3 ms at 16 / 8 samples.  Two buckets a group with
// registers capped for 4 / 5 / 7 waves per SIMD at 32 / 16 / 8 samples (16 bytes a lane
// spilled at 32; six waves at 16 samples: 6.0 ms, eight at 8: 4.6 ms).  The gradient
// variants, whose paint arithmetic would spill under those caps, keep four-bucket groups,
// uncapped registers and the indexed loops.
template <int SPP, bool SRGB, bool GRAD>
__global__ void __launch_bounds__(256, GRAD ? 1 : (SPP == 32 ? 4 : (SPP == 16 ? 5 : 7))) k_composite_ms(CovArgs A) {
    constexpr int NW = Words<SPP>::n;
    constexpr int CPR = 64 / (2 * SPP);
    constexpr int NREG = 64 / CPR;
    constexpr int G = GRAD ? 4 : 2;
    constexpr uint32_t FULL = SPP == 8 ? 0xFFFFu : 0xFFFFFFFFu;  // both rows' bits (N = 32: each row)
    __shared__ float4 s_step[SRGB ? 256 : 1];
    __shared__ uint8_t s_lut[SRGB ? 4096 : 4];
    __shared__ float s_dec[256];
    __shared__ float s_deca[GRAD ? 256 : 1];
    __shared__ float4 l_src[4][kCovList];
    __shared__ uint2 l_ref[4][kCovList];
    __shared__ uint32_t l_pid[4][kCovList];
    __shared__ uint32_t slot[4][G][64];
    __shared__ uint32_t tbl[4][G][256];
    for (int i = threadIdx.x; i < 256; i += 256) {
        if (SRGB) s_step[i] = A.step_c[i];
        s_dec[i] = A.dec_c[i];
        if (GRAD) s_deca[i] = A.dec_a[i];
    }
    for (int i = threadIdx.x; i < (SRGB ? 1024 : 0); i += 256) ((uint32_t *)s_lut)[i] = ((const uint32_t *)A.lut_c)[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int W = A.W, H = A.H, tiles_x = A.tiles_x;
    // XCD-aware tile order of the coverage compositor: XCD b % 8 takes whole cell rows
    const int xcd = (int)(blockIdx.x & 7u), t = (int)(blockIdx.x >> 3) * 4 + w;
    const int rr = t / tiles_x, r = xcd + 8 * rr;
    if (r >= (H + 1) / 2) return;
    const int tx0 = (t - rr * tiles_x) * 64;
    const int tx1 = tmin(tx0 + 64, W);
    const int x = tx0 + lane;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // blend-phase coordinates of this lane
    const int q = lane / (2 * SPP), dy = (lane / SPP) & 1;
    SampleRegs<NREG> S;
    S.init(0xFFFFFFFFu);

    auto flush = [&](int nlist) {
        wave_sync_lds();
        for (int g0 = 0; g0 < nlist; g0 += G) {
            const int ng = tmin(G, nlist - g0);
            uint32_t it[G], mk[G][NW];
            uint4 tv[G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                slot[w][g][lane] = 0u;
                const uint2 ref = g < ng ? l_ref[w][g0 + g] : make_uint2(0u, 0u);
                const bool ld = lane < (int)ref.y;
                it[g] = ld ? A.items[ref.x + lane] : 0u;
#pragma unroll
                for (int k = 0; k < NW; ++k) mk[g][k] = 0u;
                if (ld) load_mask<NW>(A.imask, (size_t)ref.x + lane, mk[g]);
                // the path's blend table (solid paints; a gradient's l_src.w is negative)
                const bool solid = g < ng && !(GRAD && l_src[w][g0 + g].w < 0.f);
                tv[g] = solid ? *(const uint4 *)(A.ptab + (size_t)l_pid[w][g0 + g] * 256 + 4 * lane)
                              : make_uint4(0u, 0u, 0u, 0u);
            }
#pragma unroll
            for (int g = 0; g < G; ++g) *(uint4 *)&tbl[w][g][4 * lane] = tv[g];
            wave_sync_lds();
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (g >= ng) break;
                const uint2 ref = l_ref[w][g0 + g];
                if (ref.y > 64u) {
                    scatter_big(A.items, ref.x, (int)ref.y, tx0, slot[w][g]);
                } else if (lane < (int)ref.y) {
                    const int x0 = (int)(it[g] & 0xFFFFu), iw = (int)(it[g] >> 16);
                    const int end = x0 + (iw ? iw : 2);
                    if (end > tx0 && x0 < tx0 + 64) atomicMax(&slot[w][g][tmax(x0 - tx0, 0)], (uint32_t)(lane + 1));
                }
            }
            wave_sync_lds();
            // per column (lane): the sample bits of both rows under each bucket's covering
            // item, 0 where no item covers it; the covering item's extent for gradients
            uint32_t cb[G], cbh[G], ce[GRAD ? G : 1];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                cb[g] = 0u;
                cbh[g] = 0u;
                if (GRAD) ce[g] = 0u;
                if (g >= ng) break;
                const uint2 ref = l_ref[w][g0 + g];
                const int p = (int)(uint32_t)dpp_scan((int)slot[w][g][lane], 0, OpMax()) - 1;
                uint32_t pit, pm[NW];
                if (ref.y > 64u) {
                    pit = p >= 0 ? A.items[ref.x + p] : 0u;
#pragma unroll
                    for (int k = 0; k < NW; ++k) pm[k] = 0u;
                    if (p >= 0) load_mask<NW>(A.imask, (size_t)ref.x + p, pm);
                } else {
                    pit = (uint32_t)__shfl((int)it[g], tmax(p, 0), 64);
#pragma unroll
                    for (int k = 0; k < NW; ++k) pm[k] = (uint32_t)__shfl((int)mk[g][k], tmax(p, 0), 64);
                }
                const int px0 = (int)(pit & 0xFFFFu), pw = (int)(pit >> 16);
                const int pend = px0 + (pw ? pw : 2);
                if (p >= 0 && x < pend && x < W) {
                    uint32_t hi = FULL, lo = FULL;
                    if (pw == 0) lo = column_bits<SPP>(pm, x - px0, hi);
                    cb[g] = lo;
                    cbh[g] = SPP == 32 ? hi : 0u;
                    if (GRAD) ce[g] = (uint32_t)px0 | ((uint32_t)(pw ? pw : 2) << 16);
                }
            }
            // blend, bucket by bucket in path order, one wave-wide pass per state register
            // holding a covered column: four LDS reads of the path's blend table for a
            // solid paint, the blend arithmetic on the pixel's own paint for a gradient.
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (g >= ng) break;
                const uint64_t regs = group_bits<CPR>(__ballot((cb[g] | cbh[g]) != 0u));
                if (!regs) continue;
                const float4 src = l_src[w][g0 + g];
                const bool gpaint = GRAD && src.w < 0.f;
                const uint32_t ceg = GRAD ? ce[g] : 0u;
                // register p's lanes are its columns' sample bits in column order (lane
                // q * 2N + dy * N + s is bit dy * N + s of column p * CPR + q; N = 32: the
                // rows' words cb, cbh), so the covered lanes form one wave-wide mask read
                // from the columns' lanes -- the select takes it as a lane mask directly
                const uint32_t cbg = cb[g], cbhg = cbh[g];
                auto covers = [&](int p) -> bool {
                    uint64_t m;
                    if (SPP == 32) {
                        m = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)cbg, p) |
                            ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)cbhg, p) << 32);
                    } else {
                        m = 0ull;
#pragma unroll
                        for (int k = 0; k < CPR; ++k)
                            m |= (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)cbg, p * CPR + k) << (2 * SPP * k);
                    }
                    return __builtin_amdgcn_inverse_ballot_w64(m);
                };
                if (!gpaint) {
                    auto blend = [&](uint32_t st, int p) -> uint32_t {
                        const uint32_t nst = table_code(tbl[w][g], st);
                        return covers(p) ? nst : st;
                    };
                    if (GRAD) S.apply4(regs, blend);
                    else S.template apply<SPP == 32 ? 4 : 2>(regs, blend);
                } else {
                    S.apply1(regs, [&](uint32_t st, int p) -> uint32_t {
                        const int c = p * CPR + q;
                        const uint32_t e = (uint32_t)__shfl((int)ceg, c, 64);
                        const float4 pa = grad_color(A.G, __float_as_uint(src.x), (int)(e & 0xFFFFu), (int)(e >> 16),
                                                     tx0 + c, 2 * r + dy, s_dec, s_deca);
                        const uint32_t nst = blend_code<SRGB>(s_step, s_lut, s_dec, st, pa, pa.w);
                        return covers(p) ? nst : st;
                    });
                }
            }
            wave_sync_lds();
        }
    };

    const uint32_t q0 = A.rowbeg[r], q1 = A.rowend[r];
    int nlist = 0;
    for (uint32_t qc = q0; qc < q1; qc += 64) {
        const bool inq = qc + lane < q1;
        const uint4 rec = inq ? A.rrec[qc + lane] : make_uint4(0xFFFFu, 0u, 0u, 0u);
        const uint64_t rt = inq ? A.rtile[qc + lane] : 0ull;
        const int exlo = (int)(rec.x & 0xFFFFu), exhi = (int)(rec.x >> 16);
        const bool paints = (rec.y & 0xFF000000u) || (GRAD && rec.y != 0u && grad_paints(A.G, rec.y));
        const bool ov = inq && exlo < tx1 && exhi > tx0 && ((rt >> tile_bit(tx0, exlo)) & 1ull) && paints;
        const unsigned long long om = __ballot(ov);
        if (!om) continue;
        const int nov = (int)__popcll(om);
        if (nlist + nov > kCovList) {
            flush(nlist);
            nlist = 0;
        }
        if (ov) {
            const int sl = nlist + (int)__popcll(om & lt);
            const uint32_t c = rec.y;
            if (GRAD && !(c & 0xFF000000u))
                l_src[w][sl] = make_float4(__uint_as_float(c), 0.f, 0.f, -1.f);
            else
                l_src[w][sl] = make_float4(s_dec[c & 0xFFu], s_dec[(c >> 8) & 0xFFu], s_dec[(c >> 16) & 0xFFu],
                                           unorm_dec((int)(c >> 24)));
            l_ref[w][sl] = make_uint2(rec.w, rec.z);
            l_pid[w][sl] = A.rpath[qc + lane];
        }
        nlist += nov;
    }
    if (nlist) flush(nlist);

    // resolve: G registers at a time are staged in this wave's slot rows (G x 64 words);
    // lane (pixel, channel) sums the pixel's decoded samples in sample order, divides by N
    // and re-quantises, and the four channel lanes of a pixel assemble its word
    static_assert(NREG % G == 0, "state registers are staged G at a time");
    uint32_t *stg = &slot[w][0][0];
    wave_sync_lds();
#pragma unroll
    for (int p0 = 0; p0 < NREG; p0 += G) {
#pragma unroll
        for (int k = 0; k < G; ++k) stg[k * 64 + lane] = S.get(p0 + k);
        wave_sync_lds();
#pragma unroll
        for (int t0 = 0; t0 < 8 * G * CPR; t0 += 64) {
            const int tt = t0 + lane;
            const int ch = tt & 3, pix = tt >> 2;           // pix = (k * CPR + qq) * 2 + row, k < G
            const int row = pix & 1, kq = pix >> 1, k = kq / CPR, qq = kq - k * CPR;
            const bool live = tt < 8 * G * CPR;
            const uint32_t *smps = stg + k * 64 + qq * 2 * SPP + row * SPP;
            float sum = 0.f;
            if (live) {
#pragma unroll 8
                for (int s2 = 0; s2 < SPP; ++s2) {
                    const uint32_t b = (smps[s2] >> (8 * ch)) & 0xFFu;
                    sum += (SRGB && ch < 3) ? s_dec[b] : unorm_dec((int)b);
                }
            }
            const float avg = sum / (float)SPP;
            float tmp;
            const uint32_t code = (SRGB && ch < 3) ? (uint32_t)srgb_code(s_step, s_lut, avg, tmp) : (uint32_t)unorm_code(avg);
            uint32_t word = code << (8 * ch);
            word |= (uint32_t)__shfl_xor((int)word, 1, 64);
            word |= (uint32_t)__shfl_xor((int)word, 2, 64);
            const int col = tx0 + (p0 + k) * CPR + qq, y = 2 * r + row;
            if (live && ch == 0 && col < W && y < H) *(uint32_t *)(A.image + (size_t)y * A.pitch + (size_t)col * 4) = word;
        }
        wave_sync_lds();
    }
}

// Row-list records for the compositors: one 16-byte load per listed bucket (and its path
// id for the multisample compositor's blend tables).
__global__ void k_row_records(int64_t NB, const uint32_t *__restrict__ n_ptr, const uint32_t *__restrict__ rowvals,
                              const uint32_t *__restrict__ bext, const int2 *__restrict__ binfo,
                              const uint32_t *__restrict__ finfo, const uint32_t *__restrict__ bicnt,
                              const uint32_t *__restrict__ boff, const uint64_t *__restrict__ btile,
                              uint4 *__restrict__ rrec, uint64_t *__restrict__ rtile, uint32_t *__restrict__ rpath) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NB || q >= (int64_t)*n_ptr) return;
    const uint32_t bb = rowvals[q];
    const uint32_t path = (uint32_t)binfo[bb].x;
    rrec[q] = make_uint4(bext[bb], finfo[path], bicnt[bb], 2u * boff[bb]);
    rtile[q] = btile[bb];
    if (rpath) rpath[q] = path;
}

// Device requantisation of arbitrary values (colour-table self-check): codes and the
// decoded values of the codes.
template <bool SRGB>
__global__ void k_encode_values(int64_t n, const float *__restrict__ v, const float4 *__restrict__ step,
                                const uint8_t *__restrict__ lut, uint8_t *__restrict__ code, float *__restrict__ dec) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float d;
    int k;
    if (SRGB) {
        k = srgb_code(step, lut, v[i], d);
    } else {
        k = unorm_code(v[i]);
        d = unorm_dec(k);
    }
    code[i] = (uint8_t)k;
    dec[i] = d;
}

// Exhaustive comparison of the device requantisation with the threshold tables over every
// float bit pattern in [lo, hi): counts disagreements of code or decoded value.
template <bool SRGB>
__global__ void k_encode_sweep(uint32_t lo, uint32_t hi, const float4 *__restrict__ step, const uint8_t *__restrict__ lut,
                               const float *__restrict__ thr, const float *__restrict__ dec, unsigned long long *bad) {
    unsigned long long nbad = 0;
    for (uint64_t b = (uint64_t)lo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < hi;
         b += (uint64_t)gridDim.x * blockDim.x) {
        const float v = __uint_as_float((uint32_t)b);
        int ref = 0;
#pragma unroll 1
        for (int s = 128; s >= 1; s >>= 1)
            if (ref + s <= 255 && v >= thr[ref + s]) ref += s;
        float d;
        int k;
        if (SRGB) {
            k = srgb_code(step, lut, v, d);
        } else {
            k = unorm_code(v);
            d = unorm_dec(k);
        }
        nbad += (k != ref || __float_as_uint(d) != __float_as_uint(dec[ref])) ? 1ull : 0ull;
    }
#pragma unroll
    for (int dd = 32; dd >= 1; dd >>= 1) nbad += __shfl_xor(nbad, dd, 64);
    if ((threadIdx.x & 63) == 0 && nbad) atomicAdd(bad, nbad);
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

__global__ void __launch_bounds__(256) k_count_reduce(const uint32_t *__restrict__ a, const uint32_t *__restrict__ c,
                                                      size_t n, unsigned long long *out) {
    __shared__ unsigned long long part[2][4];
    unsigned long long sa = 0, sc = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        sa += a[i];
        sc += c[i];
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        sa += __shfl_xor(sa, d, 64);
        sc += __shfl_xor(sc, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        part[0][threadIdx.x >> 6] = sa;
        part[1][threadIdx.x >> 6] = sc;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const unsigned long long t = part[threadIdx.x][0] + part[threadIdx.x][1] + part[threadIdx.x][2] + part[threadIdx.x][3];
        atomicAdd(&out[threadIdx.x], t);
    }
}

inline unsigned blocks_for(int64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

void launch_resolve(const SceneDev &sc, const Tables &t, Frame &f, hipStream_t s) {
    if (!f.NB) return;
    ResolveArgs A;
    A.NB = f.NB; A.binfo = f.binfo; A.boff = f.boff; A.bcarry = f.bcarry; A.frule = sc.frule;
    A.rec_meta = f.rec_meta; A.rec_mask = f.rec_mask; A.scratch = f.scratch_keys; A.W = f.W;
    A.pcanvas = sc.pcanvas; A.scarry = f.scarry;
    // the merged masks feed the multisample compositor and the debug draw list; the
    // coverage compositor reads only the per-pixel sample counts
    A.items = f.items; A.imask = f.keep_masks ? f.imask : nullptr; A.icov = f.icov; A.bicnt = f.bicnt;
    A.bmcnt = f.bmcnt; A.bext = f.bext;
    A.btile = f.btile;
    A.big_list = f.big_list; A.big_count = f.big_count; A.err = f.err_flag;  // cleared at the frame's start
    const unsigned g1 = (unsigned)std::min<int64_t>(f.NB, f.resolve_grid > 0 ? f.resolve_grid : kResolveBlocks);
    const unsigned gc = f.class_grid > 0 ? (unsigned)f.class_grid : 4096u;
    dim3 b64(64);
#define SL_LAUNCH(N)                                                                  \
    do {                                                                              \
        hipLaunchKernelGGL(k_resolve_first<N>, dim3(g1), b64, 0, s, A);               \
        hipLaunchKernelGGL(k_resolve_big<N>, dim3(gc), b64, 0, s, A);                  \
        hipLaunchKernelGGL(k_resolve_huge<N>, dim3(1024), b64, 0, s, A);              \
    } while (0)
    if (t.spp == 8) SL_LAUNCH(8);
    else if (t.spp == 16) SL_LAUNCH(16);
    else SL_LAUNCH(32);
#undef SL_LAUNCH
}

void launch_scene_carry(const SceneDev &sc, Frame &f, hipStream_t s) {
    (void)sc;
    if (!f.nscenes) return;
    hipLaunchKernelGGL(k_scene_carry, dim3(blocks_for(f.nscenes, 256)), dim3(256), 0, s, f.nscenes, f.sfirst,
                       (const uint32_t *)f.pbase, (const uint32_t *)f.bcarry, f.scarry);
}

void launch_order_rows(const SceneDev &sc, Frame &f, hipStream_t s) {
    const int64_t R = f.R;
    launch_fill2_u32(f.rowbeg, f.rowend, 0u, (size_t)R, s);
    if (!f.NB) return;
    hipLaunchKernelGGL(k_nz_flags, dim3(blocks_for(f.NB, 256)), dim3(256), 0, s, f.NB, (const uint32_t *)f.bicnt, f.nz_flag);
    scan_exclusive_u32(f.nz_flag, f.nz_off, (size_t)f.NB, f.scan_tmp, s);
    hipLaunchKernelGGL(k_nz_compact, dim3(blocks_for(f.NB, 256)), dim3(256), 0, s, f.NB, (const uint32_t *)f.nz_flag,
                       (const uint32_t *)f.nz_off, (const int2 *)f.binfo, sc.pcanvas, f.nz_keys[0], f.nz_vals[0]);
    const uint32_t *n_ptr = f.nz_off + f.NB;
    const uint32_t ntiles = (uint32_t)((f.NB + kRsTile - 1) / kRsTile);
    int bits = 0;
    while ((int64_t)1 << bits < R) ++bits;
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
    hipLaunchKernelGGL(k_row_records, dim3(blocks_for(f.NB, 256)), dim3(256), 0, s, f.NB, n_ptr,
                           (const uint32_t *)f.nz_vals[0], (const uint32_t *)f.bext, (const int2 *)f.binfo, sc.finfo,
                           (const uint32_t *)f.bicnt, (const uint32_t *)f.boff, (const uint64_t *)f.btile, f.rrec,
                           f.rtile, f.rpath);
}

void launch_composite(const SceneDev &sc, const Tables &t, Frame &f, hipStream_t s) {
    const int R = (f.H + 1) / 2;
    const int tiles_x = (f.W + 63) / 64;
    const int64_t ntiles = f.nscenes ? f.ntiles : (int64_t)R * tiles_x;
    if (!f.ms) {
        CovArgs C;
        C.W = f.W; C.H = f.H; C.tiles_x = tiles_x;
        C.rowbeg = f.rowbeg; C.rowend = f.rowend; C.rrec = f.rrec; C.rtile = f.rtile;
        C.items = f.items; C.icov = f.icov; C.imask = nullptr; C.rpath = nullptr; C.ptab = nullptr; C.dec_c = t.dec_c; C.dec_a = t.dec_a; C.step_c = t.step_c; C.lut_c = t.lut_c;
        C.image = f.image; C.pitch = f.pitch;
        C.G = f.grad;
        C.nscenes = f.nscenes; C.ntiles = ntiles; C.stile = f.stile; C.simg = f.simg;
        // one scene: 8 XCD row sets of ceil(R / 8) rows each; a batch: tiles in order
        dim3 g(f.nscenes ? blocks_for(ntiles, 4) : 8u * blocks_for((int64_t)tiles_x * ((R + 7) / 8), 4)), b(256);
#define SL_LAUNCH(N, SR)                                                                   \
    do {                                                                                   \
        if (f.nscenes) {                                                                   \
            if (f.comp_group == 4) hipLaunchKernelGGL((k_composite_cov<N, SR, 4, true, false>), g, b, 0, s, C); \
            else hipLaunchKernelGGL((k_composite_cov<N, SR, 8, true, false>), g, b, 0, s, C);               \
        } else if (f.has_grad) {                                                           \
            hipLaunchKernelGGL((k_composite_cov<N, SR, 4, false, true>), g, b, 0, s, C);   \
        } else {                                                                           \
            if (f.comp_group == 4) hipLaunchKernelGGL((k_composite_cov<N, SR, 4, false, false>), g, b, 0, s, C); \
            else hipLaunchKernelGGL((k_composite_cov<N, SR, 8, false, false>), g, b, 0, s, C);               \
        }                                                                                  \
    } while (0)
        if (f.srgb) {
            if (t.spp == 8) SL_LAUNCH(8, true);
            else if (t.spp == 16) SL_LAUNCH(16, true);
            else SL_LAUNCH(32, true);
        } else {
            if (t.spp == 8) SL_LAUNCH(8, false);
            else if (t.spp == 16) SL_LAUNCH(16, false);
            else SL_LAUNCH(32, false);
        }
#undef SL_LAUNCH
        return;
    }
    CovArgs C;
    C.W = f.W; C.H = f.H; C.tiles_x = tiles_x;
    C.rowbeg = f.rowbeg; C.rowend = f.rowend; C.rrec = f.rrec; C.rtile = f.rtile;
    C.items = f.items; C.icov = f.icov; C.imask = f.imask; C.rpath = f.rpath; C.ptab = f.ptab;
    C.dec_c = t.dec_c; C.dec_a = t.dec_a; C.step_c = t.step_c; C.lut_c = t.lut_c;
    C.image = f.image; C.pitch = f.pitch;
    C.G = f.grad;
    C.nscenes = 0; C.ntiles = ntiles; C.stile = nullptr; C.simg = nullptr;
    if (sc.P) {
        if (f.srgb)
            hipLaunchKernelGGL(k_paint_tables<true>, dim3(blocks_for(sc.P, 4)), dim3(256), 0, s, (uint32_t)sc.P, sc.finfo,
                               t.dec_c, t.step_c, t.lut_c, f.ptab);
        else
            hipLaunchKernelGGL(k_paint_tables<false>, dim3(blocks_for(sc.P, 4)), dim3(256), 0, s, (uint32_t)sc.P, sc.finfo,
                               t.dec_c, t.step_c, t.lut_c, f.ptab);
    }
    dim3 g(8u * blocks_for((int64_t)tiles_x * ((R + 7) / 8), 4)), b(256);
#define SL_LAUNCH(N)                                                                                    \
    do {                                                                                                \
        if (f.has_grad) {                                                                               \
            if (f.srgb) hipLaunchKernelGGL((k_composite_ms<N, true, true>), g, b, 0, s, C);             \
            else hipLaunchKernelGGL((k_composite_ms<N, false, true>), g, b, 0, s, C);                   \
        } else {                                                                                        \
            if (f.srgb) hipLaunchKernelGGL((k_composite_ms<N, true, false>), g, b, 0, s, C);            \
            else hipLaunchKernelGGL((k_composite_ms<N, false, false>), g, b, 0, s, C);                  \
        }                                                                                               \
    } while (0)
    if (t.spp == 8) SL_LAUNCH(8);
    else if (t.spp == 16) SL_LAUNCH(16);
    else SL_LAUNCH(32);
#undef SL_LAUNCH
}

void launch_encode_values(const Tables &t, int channel, const float *v, int64_t n, uint8_t *code, float *dec,
                          hipStream_t s) {
    if (n <= 0) return;
    dim3 g(blocks_for(n, 256)), b(256);
    if (channel == 0 && t.srgb)
        hipLaunchKernelGGL(k_encode_values<true>, g, b, 0, s, n, v, t.step_c, t.lut_c, code, dec);
    else
        hipLaunchKernelGGL(k_encode_values<false>, g, b, 0, s, n, v, t.step_c, t.lut_c, code, dec);
}

void launch_encode_sweep(const Tables &t, int channel, uint32_t lo, uint32_t hi, unsigned long long *bad,
                         hipStream_t s) {
    if (hi <= lo) return;
    const bool srgb = channel == 0 && t.srgb;
    const float *thr = channel == 0 ? t.thr_c : t.thr_a;
    const float *dec = channel == 0 ? t.dec_c : t.dec_a;
    if (srgb)
        hipLaunchKernelGGL(k_encode_sweep<true>, dim3(8192), dim3(256), 0, s, lo, hi, t.step_c, t.lut_c, thr, dec, bad);
    else
        hipLaunchKernelGGL(k_encode_sweep<false>, dim3(8192), dim3(256), 0, s, lo, hi, t.step_c, t.lut_c, thr, dec, bad);
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

void launch_count_reduce(const uint32_t *a, const uint32_t *c, size_t n, unsigned long long *out, hipStream_t s) {
    hipLaunchKernelGGL(k_count_reduce, dim3(256), dim3(256), 0, s, a, c, n, out);
}

}  // namespace sl
