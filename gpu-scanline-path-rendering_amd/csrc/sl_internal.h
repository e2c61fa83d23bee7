// sl_internal.h -- declarations shared by the host orchestration (sl_api.cpp) and the
// HIP kernels (sl_kernels.hip).  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sl {

// Per-frame device views handed to the kernel launchers.  Every pointer is a device
// pointer owned by the handle's arena.
// Block order for the 8 XCDs: hardware hands consecutive workgroups to the XCDs in turn,
// so workgroup b runs on XCD b % 8; mapping it to chunk start(b % 8) + b / 8 gives each
// XCD one contiguous range of a uniform-cost grid, so neighbouring output records are
// written through one XCD's L2 (the mapping is a bijection whatever the deal really is).
__device__ __forceinline__ uint32_t xcd_chunk(uint32_t b, uint32_t n) {
    const uint32_t x = b & 7u, i = b >> 3, q = n >> 3, r = n & 7u;
    return x * q + (x < r ? x : r) + i;
}

struct SceneDev {
    uint32_t P, C, V;
    const float2 *vtx;       // [V] vertices in unified (orientation-fixed) order
    const uint32_t *vpath;   // [V] owning path of each vertex
    const uint32_t *cvpos;   // [C]
    const uint8_t *ctype;    // [C]
    const uint32_t *cpath;   // [C]
    const float *cw;         // [C] rational weight (else -1)
    const uint8_t *crev;     // [C] 1 = vertex order was reversed at upload
    const uint4 *cinfo;      // [C] the same per curve in one record: {vertex pos, path,
                             //     type | reversed << 8, rational weight bits}
    const uint8_t *frule;    // [P]
    const uint32_t *finfo;   // [P] fill_info (RGBA8, alpha 0 -> 0)
    const int4 *pcanvas;     // [P] batched scenes: {W, H, first cell row, scene} of the
                             //     path's own canvas; null for a single scene
};

// Canvas of a path: the frame's, or the path's own scene's in a batch.
__device__ __forceinline__ void path_canvas(const SceneDev &sc, uint32_t path, int &W, int &H) {
    if (sc.pcanvas) {
        const int4 c = sc.pcanvas[path];
        W = c.x;
        H = c.y;
    }
}

struct Tables {
    int spp;                 // 8, 16 or 32
    int srgb;                // colour channels sRGB-encoded
    int words;               // mask words per fragment: 1, 2, 4
    const uint32_t *pixel;   // [65536 * words] line -> sample mask table
    const float *dec_c;      // [256] colour decode (sRGB or UNORM)
    const float *thr_c;      // [256] colour encode thresholds
    const uint8_t *lut_c;    // [4096] encode LUT
    const float *dec_a;      // [256] alpha decode (UNORM)
    const float *thr_a;      // [256]
    const uint8_t *lut_a;    // [4096]
    const float4 *step_c;    // [256] sRGB steps {thr[k+1], dec[k], dec[k+1], 0}
};

// Gradient paint of a frame (output_32.vert.glsl:55-97, output_32.frag.glsl:45-51).  A
// fill word g in [1, 0x01000000] names table slot g - 1; the slot paints when its ramp
// scale (word 7) is positive, otherwise the path draws nothing.
struct GradDev {
    const float4 *table;     // [3 * nslots] per slot {tr0, tr1, tr2, tr3}, {tr4, tr5, ramp
                             //   position (int bits), ramp scale}, {focal x, y, z, 0}
    const uint32_t *ramp;    // [rows * 1024] RGBA8 ramp texels
    uint32_t nslots;
    int rows;                // ramp rows (0: no ramp texture: the paint reads (0, 0, 0, 1))
    float irh;               // 1 / rows (0 without rows)
    float xoff;              // x offset of the line ends: 1 in coverage output, 0 in
                             //   multisample output (output_32.vert.glsl:68, ms_output_32:69)
    float ctx[13];           // inverse projection rx, ry, rw, rp (x, y, z each), a
                             //   (vg_main.cpp:117-143)
};

// bytes of the intersection work lists' class counts and cursors (Frame::lcount)
constexpr size_t kListCountBytes = 1024;

struct Frame {
    int W, H;
    float m[16];             // rows of the input transform, row-major
    // path / curve level
    float2 *tvtx;            // [V]
    uint32_t *pflags;        // [P] 8 region bits
    int2 *prow;              // [P] min / max valid cell row
    float4 *crec;            // [4C] per curve, one 64-byte record read by the intersection
                             //   passes: {x0, y0, x1, y1}, {x2, y2, x3, y3} (transformed),
                             //   the monotone cuts {t0..t3}, {type | cut count << 8, weight
                             //   bits, path, 0}
    uint32_t *cfrag;         // [C+1]  -> exclusive offsets after scan
    uint32_t *coff;          // [C+1]
    uint32_t *pcnt;          // [5C] per monotone piece: x cuts | y cuts << 16
    uint32_t *nchain;        // [C+1] solve chains per curve (2 per piece)
    uint32_t *choff;         // [C+1] exclusive scan of nchain
    uint4 *chains;           // [NCH] solve chains of long pieces, grouped by class, and
    uint4 *plist;            // [NCH/2] pieces: long ones, then short ones grouped by class;
                             //   entries {curve << 4 | piece << 1 | side (chains) or curve << 3 |
                             //   piece, ts offset of the piece, nx | ny << 16, 0}
    uint32_t *lcount;        // [2 * 101] class counts (50 chain, 1 long-piece, 50 short-piece
                             //   classes), then the place pass's cursors
    int64_t NCH;
    float *chain_t;          // [F] solved t | side per chain
    // fragments
    int64_t F;
    int2 *ts;                // [F]
    uint32_t *pnb;           // [P+1] buckets per path -> scanned to pbase
    uint32_t *pbase;         // [P+1]
    int64_t NB;
    int2 *binfo;             // [NB] (path, row or -1)
    uint32_t *bcount;        // [NB+1] -> boff
    uint32_t *boff;          // [NB+1]
    uint32_t *bwsum;         // [NB+1] per-bucket sum of scan windings -> bcarry
    uint32_t *bcarry;        // [NB+1]
    uint32_t *frank;         // [F] rank of each fragment in its bucket (~0: none)
    uint32_t *rec_meta;      // [F] (valid fragments only are written)
    uint32_t *rec_mask;      // [F * words]
    uint32_t *scratch_keys;  // [F] sort keys of large buckets
    // resolve outputs, per bucket and per item slot (2 slots per valid fragment)
    uint32_t *items;         // [2F]
    uint32_t *imask;         // [2F * words]
    uint32_t *icov;          // [2F] per merged item: sample counts of its 4 pixels
    uint32_t *bicnt;         // [NB+1]
    uint32_t *bmcnt;         // [NB+1]
    uint32_t *bext;          // [NB]
    uint64_t *btile;         // [NB] 64-pixel tiles of the row holding items
    uint32_t *big_list;      // [NB] buckets with more than 64 fragments
    uint32_t *big_count;     // [1]
    uint32_t *err_flag;      // [1]
    // compositing order
    uint32_t *nz_flag;       // [NB+1]
    uint32_t *nz_off;        // [NB+1]
    uint32_t *nz_keys[2];    // [NB] row keys (ping-pong)
    uint32_t *nz_vals[2];    // [NB] bucket ids (ping-pong)
    uint32_t *rs_hist;       // radix histogram scratch
    uint32_t *rs_hoff;
    int64_t R;               // cell rows (of all scenes of a batch)
    // batched scenes (nscenes > 0): per scene {W, H, first tile, first cell row}, image
    // byte offsets, first path, and the winding-scan carry at its first bucket
    int nscenes;
    int64_t ntiles;
    const int4 *stile;
    const uint64_t *simg;
    const uint32_t *sfirst;
    uint32_t *scarry;
    // launch shape (autotuned for batches; results do not depend on it)
    int resolve_grid, class_grid, comp_group, frag_block;
    uint32_t *rowbeg;        // [R]
    uint32_t *rowend;        // [R]
    uint4 *rrec;             // [NB] row-list records {extent, colour, items, item base}
    uint64_t *rtile;         // [NB] row-list tile masks
    uint32_t *rpath;         // [NB] row-list path ids (multisample output)
    uint32_t *ptab;          // [P][256] per-path blend tables (multisample output)
    uint8_t *image;          // [H * pitch]
    int pitch;
    int srgb, ms;
    int keep_masks;          // resolve stores the merged masks (multisample output, debug dumps)
    int has_grad;            // some fill word names a gradient: the compositor's gradient variant
    GradDev grad;
    // scan scratch
    uint32_t *scan_tmp;
    size_t scan_tmp_elems;
};

// ---- launchers (sl_kernels.hip) ----
size_t scan_tmp_elems_needed(size_t n);
void scan_exclusive_u32(const uint32_t *in, uint32_t *out, size_t n, uint32_t *tmp, hipStream_t s);

void launch_path_init(const SceneDev &sc, Frame &f, hipStream_t s);
void launch_transform(const SceneDev &sc, Frame &f, hipStream_t s);
void launch_monotone(const SceneDev &sc, Frame &f, hipStream_t s);
void launch_intersect(const SceneDev &sc, Frame &f, hipStream_t s);
void launch_path_buckets(const SceneDev &sc, Frame &f, hipStream_t s);
void launch_bucket_info(const SceneDev &sc, Frame &f, hipStream_t s);
void launch_frag_count(const SceneDev &sc, const Tables &t, Frame &f, hipStream_t s);
void launch_frag_scatter(const SceneDev &sc, const Tables &t, Frame &f, hipStream_t s);
void launch_bucket_carry(Frame &f, hipStream_t s);
void launch_resolve(const SceneDev &sc, const Tables &t, Frame &f, hipStream_t s);
void launch_scene_carry(const SceneDev &sc, Frame &f, hipStream_t s);
void launch_order_rows(const SceneDev &sc, Frame &f, hipStream_t s);
void launch_composite(const SceneDev &sc, const Tables &t, Frame &f, hipStream_t s);
void launch_encode_values(const Tables &t, int channel, const float *v, int64_t n, uint8_t *code, float *dec,
                          hipStream_t s);
void launch_encode_sweep(const Tables &t, int channel, uint32_t lo, uint32_t hi, unsigned long long *bad,
                         hipStream_t s);
void launch_fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s);
void launch_valu_probe(int blocks, int iters, float *out, hipStream_t s);
void launch_fill_rows_u8(uint8_t *p, uint8_t v, size_t row_bytes, size_t pitch, size_t rows, hipStream_t s);

// GPU-side scene generation (sl_anim.hip): one diagram's interval between keyframes
// k - 1 (_a) and k (_b), device pointers, and the eased weight t of the frame
constexpr int kMaxChordDiagrams = 8;
constexpr int kChordKeyMs = 6000;  // time from one keyframe to the next (animation.cu:1708)
struct ChordInterval {
    int n_nodes, n_links;
    float t;
    float2 center;
    const float *node_t0_a, *node_t0_b, *node_t1_a, *node_t1_b, *node_size_a, *node_size_b;  // [n_nodes]
    const uint32_t *node_color;                                                                // [n_nodes]
    const uint2 *link_node;                  // [n_links] the two nodes
    const float4 *link_from, *link_to;       // [n_links] {pos0, pos1, size0, size1} at k - 1, k
};
// the scene arrays a frame writes (6 curves and 24 vertex slots per path)
struct ChordScene {
    float2 *vtx;
    uint32_t *vpath, *cvpos, *cpath, *finfo;
    uint8_t *ctype, *frule, *crev;
    float *cw;
    uint4 *cinfo;
};
float chord_ease(int time_ms);
void launch_chord_frame(const ChordInterval *d, int n_diagrams, uint32_t P, const ChordScene &out, hipStream_t s);
void launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s);
void launch_fill2_u32(uint32_t *p, uint32_t *q, uint32_t v, size_t n, hipStream_t s);

// static scratch of the allocation-free frame (sl_noalloc.hip): its size, and its device
// address on `device` (which must be current; looked up once per device)
size_t noalloc_scratch_bytes();
hipError_t noalloc_scratch(int device, void **out);

// debug (parity dumps)
void launch_debug_fragments(const SceneDev &sc, const Tables &t, Frame &f, int32_t *yx, int32_t *pid,
                            int8_t *wn, uint32_t *mask4, hipStream_t s);
void launch_debug_draw_list(const SceneDev &sc, const Tables &t, Frame &f, const uint32_t *moff,
                            const uint32_t *ioff, int4 *draw, uint32_t *mmask4, hipStream_t s);
// out[0] += sum of a, out[1] += sum of c
void launch_count_reduce(const uint32_t *a, const uint32_t *c, size_t n, unsigned long long *out, hipStream_t s);

}  // namespace sl
