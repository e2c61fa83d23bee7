// sl_api.cpp -- the C ABI (include/scanline.h): handle state, scene upload, per-frame
// orchestration of the HIP kernels, timing and parity dumps.
//
// The frame follows rasterizeImpl<2> (ras_cut_mask_comb_scanline.cu:879-1453):
//   transform+cull -> monotone split + fragment count -> scan -> intersections ->
//   (path,row) buckets -> fragment count / scatter -> per-bucket resolve (sort by x,
//   winding scan, stencil, spans) -> row order -> compute compositor.
// Like the reference it reads two counts back to the host per frame (F after the curve
// scan, the bucket total after the path scan) to size its buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/scanline.h"
#include "sl_internal.h"
#include "sl_tables.h"

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

// The per-frame device arrays of one frame.  A handle keeps one arena that grows by
// hipMalloc and is reused frame to frame; the allocation-free frame builds a temporary
// arena whose arrays are carved out of the library's static scratch instead.
struct Arena {
    DevBuf b_pcnt, b_nchain, b_choff, b_chains, b_chain_t;
    DevBuf b_tvtx, b_pflags, b_prow, b_crec, b_cfrag, b_coff, b_ts, b_pnb, b_pbase, b_binfo, b_bcount,
        b_boff, b_bwsum, b_bcarry, b_rec_meta, b_rec_mask, b_scratch, b_items, b_imask, b_icov, b_bicnt, b_bmcnt,
        b_bext, b_big_list, b_small, b_nz_flag, b_nz_off, b_nzk0, b_nzk1, b_nzv0, b_nzv1, b_rs_hist, b_rs_hoff,
        b_rowbeg, b_rowend, b_rrec, b_btile, b_rtile, b_rpath, b_ptab, b_image, b_scan_tmp, b_lcount;
    // carving state: non-null base = every array comes from [base, base + cap)
    uint8_t *carve_base = nullptr;
    size_t carve_cap = 0, carve_off = 0;
};

const char *kStageNames[] = {"transform_cull", "monotone_count", "intersect", "bucket_setup",
                             "fragments",      "resolve",        "order_rows", "composite"};
constexpr int kStages = 8;

}  // namespace

struct sl_handle {
    int device = 0;
    int spp = 32;
    int ms = 1;
    int srgb = 1;
    float m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};  // column-major
    int W = 0, H = 0;
    int vp_x = 0, vp_y = 0, vp_w = 0, vp_h = 0;  // presentation rectangle (0 size: the image's)
    bool has_scene = false;
    uint32_t P = 0, C = 0, V = 0;
    std::string err;
    sl_counts counts{};
    int64_t pixel_count = 0;
    bool timing = false;
    double stage_ms[kStages] = {0};
    bool debug = false;

    // batched scenes (the internal handles of an sl_batch): per-path canvases and the
    // per-scene tables of the compositor
    int nscenes = 0;
    int64_t batch_rows = 0, batch_tiles = 0, batch_image_bytes = 0;
    DevBuf pcanvas, stile, simg, sfirst, scarry;
    // launch shape (results do not depend on it)
    int resolve_grid = 0, class_grid = 0, comp_group = 4, frag_block = 256;
    // sizes of the last eager render; a replay of the same scene reuses them
    bool sized = false;
    uint32_t last_F = 0, last_NCH = 0, last_NB = 0;

    // scene (device)
    DevBuf vtx, vpath, cvpos, ctype, cpath, cw, crev, cinfo, frule, finfo;
    // gradient textures of the scene: ramp texels (rows of 1024) and 12-word table slots
    DevBuf gramp, gtable;
    uint32_t gslots = 0, gramp_rows = 0;
    bool has_grad = false;  // some path's fill word names a gradient slot
    // pinned host staging of the scene upload, kept and grown across sl_add_vg calls
    void *stage = nullptr;
    size_t stage_cap = 0;
    // tables (device)
    DevBuf ptable[3];
    DevBuf ctab[2];  // [srgb] : dec_c, thr_c, dec_a, thr_a, lut_c, lut_a packed
    // per-frame arena
    Arena A;
    // debug host copies
    std::vector<int32_t> d_ts, d_yx, d_pid, d_draw;
    std::vector<int8_t> d_wn;
    std::vector<uint32_t> d_mask, d_mmask, d_ptable;
    std::vector<uint8_t> d_image;
    std::vector<uint32_t> d_binfo, d_bicnt, d_bext, d_boff, d_rowkeys, d_rowvals, d_rowbeg, d_rowend;
    hipEvent_t ev[kStages + 1] = {};
    bool events_ready = false;
    // a replayed frame (see sl_rasterize) is still running until done_ev has passed
    hipEvent_t done_ev = nullptr;
    bool pending = false;
    hipStream_t pending_stream = nullptr;  // the stream done_ev was recorded on
    // recorded on pending_stream when a frame moves to another stream: everything queued
    // there so far (the frame and the caller's reads of its image) precedes the new frame
    hipEvent_t switch_ev = nullptr;
    // size of the image the last frame wrote (the output size may change since)
    int img_W = 0, img_H = 0;
    // device address of this GPU's static scratch (the allocation-free frame)
    void *noalloc_scratch = nullptr;
    int64_t NB = 0;
    uint32_t *nz_vals_sorted = nullptr;
    // Status of the last single-scene frame, copied back without waiting at the end of the
    // frame into pinned memory (words 0..7 of the frame's small block: big-bucket count,
    // error flags, then the merged-fragment and item totals) and read when the frame is
    // settled: the element counts of an eager frame, and every frame's error flags.
    uint32_t *fstat = nullptr;
    bool stat_pending = false;   // a queued frame writes fstat
    bool stat_counts = false;    // ...including the element counts (an eager frame)
    bool timing_pending = false; // ...and its stage events are to be read
    int frame_rc = SL_OK;        // error of the last settled frame
    std::string frame_err;
    // keyframes of the scene generated on the device (sl_chord_load), one block on the
    // device; per diagram the byte offsets of its arrays in that block
    struct Chord {
        int n_keys = 0, n_nodes = 0;
        std::vector<uint32_t> link_offset;
        float cx = 0.f, cy = 0.f;
        size_t o_t0 = 0, o_t1 = 0, o_size = 0, o_color = 0, o_node = 0, o_from = 0, o_to = 0;
    };
    std::vector<Chord> chord;
    DevBuf chord_keys;
};

namespace {

int fail(sl_handle *h, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    return code;
}

#define HIPCHK(h, call)                                                                                  \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess) return fail((h), SL_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

int ensure(sl_handle *h, DevBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return SL_OK;
    if (b.p) {
        hipError_t e = hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        if (e != hipSuccess) return fail(h, SL_E_HIP, "hipFree failed: %s", hipGetErrorString(e));
    }
    size_t want = bytes + bytes / 4;  // headroom for frame-to-frame growth
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        (void)hipGetLastError();
        return fail(h, SL_E_OOM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    b.cap = want;
    return SL_OK;
}

void release(DevBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

// ensure() for a frame's arena: a carved arena hands out 256-byte aligned pieces of its
// block (a piece that is too small is replaced by a new one; nothing is ever freed).
int ensure_in(sl_handle *h, Arena &A, DevBuf &b, size_t bytes) {
    if (!A.carve_base) return ensure(h, b, bytes);
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return SL_OK;
    const size_t at = (A.carve_off + 255) & ~(size_t)255;
    if (at > A.carve_cap || bytes > A.carve_cap - at)
        return fail(h, SL_E_UNSUPPORTED,
                    "the allocation-free frame needs more than its %zu MiB of static scratch for this scene "
                    "(%zu bytes in use, %zu more asked); render it with sl_rasterize",
                    sl::noalloc_scratch_bytes() >> 20, at, bytes);
    b.p = A.carve_base + at;
    b.cap = bytes;
    A.carve_off = at + bytes;
    return SL_OK;
}

template <class T> T *P_(DevBuf &b) { return (T *)b.p; }

struct I4 {
    int x, y, z, w;
};

// saveOutputToFile's pixel order (ras_base.cpp:1249-1262): rows flipped, alpha 255
void flip_rows_alpha(uint8_t *rgba, int W, int H) {
    const size_t row = (size_t)W * 4;
    std::vector<uint8_t> tmp(row);
    for (int y = 0; y < H / 2; ++y) {
        uint8_t *a = rgba + row * y, *b = rgba + row * (H - 1 - y);
        std::memcpy(tmp.data(), a, row);
        std::memcpy(a, b, row);
        std::memcpy(b, tmp.data(), row);
    }
    for (size_t i = 3; i < row * H; i += 4) rgba[i] = 255;
}

int upload(sl_handle *h, DevBuf &b, const void *src, size_t bytes) {
    int rc = ensure(h, b, bytes);
    if (rc) return rc;
    if (bytes) HIPCHK(h, hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    return SL_OK;
}

int words_for(int spp) { return spp == 8 ? 1 : (spp == 16 ? 2 : 4); }
int spp_slot(int spp) { return spp == 8 ? 0 : (spp == 16 ? 1 : 2); }

// colour table block layout in one device buffer
struct ColorBlock {
    float step_c[256][4];  // {thr[k+1], dec[k], dec[k+1], 0} of the colour encoding
    float dec_c[256], thr_c[256], dec_a[256], thr_a[256];
    uint8_t lut_c[4096], lut_a[4096];
};

int ensure_tables(sl_handle *h) {
    DevBuf &pt = h->ptable[spp_slot(h->spp)];
    if (!pt.p) {
        std::vector<uint32_t> tab;
        sl::build_pixel_table(h->spp, tab);
        int rc = upload(h, pt, tab.data(), tab.size() * 4);
        if (rc) return rc;
    }
    DevBuf &ct = h->ctab[h->srgb ? 1 : 0];
    if (!ct.p) {
        sl::ColorTables c, a;
        if (!sl::build_color_tables(h->srgb, c) || !sl::build_color_tables(0, a))
            return fail(h, SL_E_UNSUPPORTED, "colour encode table has two thresholds in one LUT bucket");
        ColorBlock blk;
        std::memcpy(blk.dec_c, c.dec, sizeof blk.dec_c);
        std::memcpy(blk.thr_c, c.thr, sizeof blk.thr_c);
        std::memcpy(blk.lut_c, c.lut, sizeof blk.lut_c);
        std::memcpy(blk.dec_a, a.dec, sizeof blk.dec_a);
        std::memcpy(blk.thr_a, a.thr, sizeof blk.thr_a);
        std::memcpy(blk.lut_a, a.lut, sizeof blk.lut_a);
        for (int k = 0; k < 256; ++k) {
            blk.step_c[k][0] = k < 255 ? c.thr[k + 1] : INFINITY;
            blk.step_c[k][1] = c.dec[k];
            blk.step_c[k][2] = c.dec[k < 255 ? k + 1 : 255];
            blk.step_c[k][3] = 0.f;
        }
        int rc = upload(h, ct, &blk, sizeof blk);
        if (rc) return rc;
    }
    return SL_OK;
}

bool tables_ready(sl_handle *h) {
    return h->ptable[spp_slot(h->spp)].p && h->ctab[h->srgb ? 1 : 0].p;
}

sl::Tables tables_view(sl_handle *h) {
    sl::Tables t;
    t.spp = h->spp;
    t.words = words_for(h->spp);
    t.pixel = P_<uint32_t>(h->ptable[spp_slot(h->spp)]);
    ColorBlock *cb = P_<ColorBlock>(h->ctab[h->srgb ? 1 : 0]);
    t.dec_c = cb->dec_c; t.thr_c = cb->thr_c; t.lut_c = cb->lut_c;
    t.dec_a = cb->dec_a; t.thr_a = cb->thr_a; t.lut_a = cb->lut_a;
    t.step_c = (const float4 *)cb->step_c;
    t.srgb = h->srgb;
    return t;
}

sl::SceneDev scene_view(sl_handle *h) {
    sl::SceneDev s;
    s.P = h->P; s.C = h->C; s.V = h->V;
    s.vtx = P_<float2>(h->vtx); s.vpath = P_<uint32_t>(h->vpath);
    s.cvpos = P_<uint32_t>(h->cvpos); s.ctype = P_<uint8_t>(h->ctype); s.cpath = P_<uint32_t>(h->cpath);
    s.cw = P_<float>(h->cw); s.crev = P_<uint8_t>(h->crev); s.cinfo = P_<uint4>(h->cinfo);
    s.frule = P_<uint8_t>(h->frule); s.finfo = P_<uint32_t>(h->finfo);
    s.pcanvas = h->nscenes ? P_<int4>(h->pcanvas) : nullptr;
    return s;
}

int check_launch(sl_handle *h, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(h, SL_E_HIP, "kernel launch (%s) failed: %s", what, hipGetErrorString(e));
    return SL_OK;
}

// The status a settled frame left in fstat: its error flags and, for an eager frame, the
// element counts and stage times.
void collect_status(sl_handle *h) {
    if (!h->stat_pending) return;
    h->stat_pending = false;
    const uint32_t *st = h->fstat;
    h->frame_rc = SL_OK;
    h->frame_err.clear();
    if (st[1] & 1u) {
        h->frame_rc = SL_E_HIP;
        h->frame_err = "internal error: fragment row outside its path's bucket range";
    } else if (st[1] & 2u) {
        h->frame_rc = SL_E_UNSUPPORTED;
        h->frame_err = "a (path, row) bucket holds more than 65536 fragments";
    }
    if (h->stat_counts) {
        unsigned long long M = 0, MS = 0;
        memcpy(&M, st + 4, 8);
        memcpy(&MS, st + 6, 8);
        h->counts.merged_fragment = (int64_t)M;
        h->counts.span = (int64_t)(MS - M);
        h->stat_counts = false;
    }
    if (h->timing_pending) {
        for (int i = 0; i < kStages; ++i) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, h->ev[i], h->ev[i + 1]);
            h->stage_ms[i] = ms;
        }
        h->timing_pending = false;
    }
}

// One frame.  `replay` re-issues the work of the previous eager render of the same scene
// without any host synchronisation (buffer sizes and counts are taken from that render),
// so that the launch sequence can be captured into a HIP graph or timed back to back.
// The gradient paint's inverse projection, as the reference's driver sets it
// (updateRasterizerTransform, vg_main.cpp:117-143): with R the inverse of the input
// transform, a window point s maps to rp + a / rd.z * rd, rd = rx s.x + ry s.y + rw, where
// rp = R (W/2, H/2, -min(W, H), 1), rx, ry = R's first two columns, rw = R's last column -
// rp and a = -rp.z: the ray from the eye through s meets the scene plane z = 0.  Computed
// in double (Gauss-Jordan with partial pivoting) and rounded to float once; a singular
// transform gives zeros.
void inverse_projection(const float m[16], int W, int H, float ctx[13]) {
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            a[r][c] = (double)m[4 * c + r];  // column-major input
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

// A frame's arrays come from `A`: the handle's own arena, or a carved one (the
// allocation-free frame), which must then name the caller's image `out_image` / `out_pitch`.
int render(sl_handle *h, hipStream_t s, bool replay, Arena &A, uint8_t *out_image = nullptr, int out_pitch = 0) {
    int rc;
    const bool carved = A.carve_base != nullptr;
    if (replay && !h->sized) return fail(h, SL_E_STATE, "replay before an eager render of the scene");
    if (!h->has_scene) return fail(h, SL_E_STATE, "rasterize called before add_vg");
    if (h->W <= 0 || h->H <= 0) return fail(h, SL_E_STATE, "output size not set (resize_output_buffer)");
    if (carved) {
        if (replay || h->debug || h->nscenes || !out_image) return fail(h, SL_E_STATE, "carved frame misuse");
        if (!tables_ready(h))
            return fail(h, SL_E_STATE, "mask and colour tables for %d spp are not built (they are built by "
                        "sl_add_vg, sl_set_samples and sl_enable_srgb_correction)", h->spp);
    } else if ((rc = ensure_tables(h))) {
        return rc;
    }
    const sl::SceneDev sc = scene_view(h);
    const sl::Tables tb = tables_view(h);
    const int NW = tb.words;
    const uint32_t P = h->P, C = h->C, V = h->V;

    sl::Frame f{};
    f.W = h->W; f.H = h->H;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) f.m[4 * i + j] = h->m[4 * j + i];
    f.srgb = h->srgb; f.ms = h->ms;
    f.keep_masks = (h->ms || h->debug) ? 1 : 0;
    f.has_grad = (h->has_grad && !h->nscenes) ? 1 : 0;
    if (f.has_grad) {
        f.grad.table = P_<float4>(h->gtable);
        f.grad.ramp = P_<uint32_t>(h->gramp);
        f.grad.nslots = h->gslots;
        f.grad.rows = (int)h->gramp_rows;
        f.grad.irh = h->gramp_rows ? 1.f / (float)h->gramp_rows : 0.f;
        f.grad.xoff = h->ms ? 0.f : 1.f;
        inverse_projection(h->m, h->W, h->H, f.grad.ctx);
    }
    f.pitch = carved ? out_pitch : h->W * 4;
    f.R = h->nscenes ? h->batch_rows : (int64_t)(h->H + 1) / 2;
    f.nscenes = h->nscenes;
    f.ntiles = h->batch_tiles;
    f.stile = P_<int4>(h->stile);
    f.simg = P_<uint64_t>(h->simg);
    f.sfirst = P_<uint32_t>(h->sfirst);
    f.scarry = P_<uint32_t>(h->scarry);
    f.resolve_grid = h->resolve_grid; f.class_grid = h->class_grid;
    f.comp_group = h->comp_group; f.frag_block = h->frag_block;
    const size_t image_bytes = h->nscenes ? (size_t)h->batch_image_bytes : (size_t)h->H * f.pitch;

    size_t scan_need = sl::scan_tmp_elems_needed((size_t)std::max(C, P) + 1);
    int rc2 = SL_OK;
    (void)rc2;
    if ((rc = ensure_in(h, A, A.b_tvtx, (size_t)V * 8)) || (rc = ensure_in(h, A, A.b_pflags, (size_t)P * 4)) ||
        (rc = ensure_in(h, A, A.b_prow, (size_t)P * 8)) || (rc = ensure_in(h, A, A.b_crec, (size_t)C * 64)) ||
        (rc = ensure_in(h, A, A.b_cfrag, ((size_t)C + 1) * 4)) ||
        (rc = ensure_in(h, A, A.b_coff, ((size_t)C + 1) * 4)) || (rc = ensure_in(h, A, A.b_pnb, ((size_t)P + 1) * 4)) ||
        (rc = ensure_in(h, A, A.b_pcnt, (size_t)C * 20)) || (rc = ensure_in(h, A, A.b_nchain, ((size_t)C + 1) * 4)) ||
        (rc = ensure_in(h, A, A.b_choff, ((size_t)C + 1) * 4)) ||
        (rc = ensure_in(h, A, A.b_pbase, ((size_t)P + 1) * 4)) || (rc = ensure_in(h, A, A.b_small, 64)) ||
        (rc = ensure_in(h, A, A.b_lcount, sl::kListCountBytes)) ||
        (!carved && (rc = ensure_in(h, A, A.b_image, image_bytes))) ||
        (rc = ensure_in(h, A, A.b_scan_tmp, scan_need * 4)))
        return rc;
    f.tvtx = P_<float2>(A.b_tvtx); f.pflags = P_<uint32_t>(A.b_pflags); f.prow = P_<int2>(A.b_prow);
    f.crec = P_<float4>(A.b_crec); f.cfrag = P_<uint32_t>(A.b_cfrag);
    f.coff = P_<uint32_t>(A.b_coff); f.pnb = P_<uint32_t>(A.b_pnb); f.pbase = P_<uint32_t>(A.b_pbase);
    f.pcnt = P_<uint32_t>(A.b_pcnt); f.nchain = P_<uint32_t>(A.b_nchain); f.choff = P_<uint32_t>(A.b_choff);
    f.image = carved ? out_image : P_<uint8_t>(A.b_image);
    if (!carved) { h->img_W = h->W; h->img_H = h->H; }
    f.big_count = P_<uint32_t>(A.b_small);
    f.err_flag = P_<uint32_t>(A.b_small) + 1;
    f.scan_tmp = P_<uint32_t>(A.b_scan_tmp);

    auto mark = [&](int i) {
        if (h->timing && !replay) (void)hipEventRecord(h->ev[i], s);
    };
    mark(0);
    sl::launch_fill_u32(f.big_count, 0u, 2, s);  // the big-bucket count and the error flag
    sl::launch_path_init(sc, f, s);
    sl::launch_transform(sc, f, s);
    if ((rc = check_launch(h, "transform"))) return rc;
    mark(1);
    sl::launch_monotone(sc, f, s);
    sl::scan_exclusive_u32(f.cfrag, f.coff, C, f.scan_tmp, s);
    sl::scan_exclusive_u32(f.nchain, f.choff, C, f.scan_tmp, s);
    if ((rc = check_launch(h, "monotone"))) return rc;
    // the bucket ranges need only the paths' row ranges, final after the monotone pass, so
    // an eager frame reads its three sizes (fragments, chains, buckets) in one round trip
    sl::launch_path_buckets(sc, f, s);
    sl::scan_exclusive_u32(f.pnb, f.pbase, P, f.scan_tmp, s);
    uint32_t F = 0, NCH = 0, NB = 0;
    if (replay) {
        F = h->last_F;
        NCH = h->last_NCH;
        NB = h->last_NB;
    } else {
        if (!carved) h->sized = false;
        HIPCHK(h, hipMemcpyAsync(&F, f.coff + C, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(h, hipMemcpyAsync(&NCH, f.choff + C, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(h, hipMemcpyAsync(&NB, f.pbase + P, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(h, hipStreamSynchronize(s));
        h->counts.path = P; h->counts.curve = C; h->counts.vertex = V;
        h->counts.curve_fragment = F;
        h->counts.merged_fragment = 0;
        h->counts.span = 0;
        h->pixel_count = 0;
        h->NB = 0;
    }
    f.F = F;
    f.NCH = NCH;
    if (F == 0) {
        // nothing to draw: the white clear of the reference's output pass
        if (carved) sl::launch_fill_rows_u8(f.image, 255, (size_t)f.W * 4, (size_t)f.pitch, (size_t)f.H, s);
        else sl::launch_fill_u8(f.image, 255, image_bytes, s);
        for (int i = 2; i <= kStages; ++i) mark(i);
        if ((rc = check_launch(h, "clear"))) return rc;
        if (replay) return SL_OK;
        HIPCHK(h, hipStreamSynchronize(s));
        if (!carved) { h->last_F = 0; h->last_NCH = 0; h->last_NB = 0; h->sized = true; }
        return SL_OK;
    }
    if ((rc = ensure_in(h, A, A.b_ts, (size_t)F * 8)) || (rc = ensure_in(h, A, A.b_chains, ((size_t)NCH + (size_t)NCH / 2 + 1) * 16)) ||
        (rc = ensure_in(h, A, A.b_chain_t, (size_t)F * 4 + 64)))  // padded: read in aligned 16-byte groups
        return rc;
    f.ts = P_<int2>(A.b_ts);
    f.chains = P_<uint4>(A.b_chains);
    f.plist = f.chains + (size_t)NCH;
    f.lcount = P_<uint32_t>(A.b_lcount);
    f.chain_t = P_<float>(A.b_chain_t);
    mark(2);
    sl::launch_intersect(sc, f, s);
    if ((rc = check_launch(h, "intersect"))) return rc;
    mark(3);
    f.NB = NB;
    if (!replay) h->NB = NB;
    const size_t R = (size_t)f.R;
    const size_t rs_tiles = ((size_t)NB + 2047) / 2048;
    scan_need = sl::scan_tmp_elems_needed(
        std::max<size_t>({(size_t)C + 1, (size_t)P + 1, (size_t)NB + 1, 256 * rs_tiles + 1, (size_t)F + 1}));
    if ((rc = ensure_in(h, A, A.b_binfo, (size_t)NB * 8)) || (rc = ensure_in(h, A, A.b_bcount, ((size_t)NB + 1) * 4)) ||
        (rc = ensure_in(h, A, A.b_boff, ((size_t)NB + 1) * 4)) || (rc = ensure_in(h, A, A.b_bwsum, ((size_t)NB + 1) * 4)) ||
        (rc = ensure_in(h, A, A.b_bcarry, ((size_t)NB + 1) * 4)) ||
        (rc = ensure_in(h, A, A.b_rec_meta, (size_t)F * 4)) || (rc = ensure_in(h, A, A.b_rec_mask, (size_t)F * 4 * NW)) ||
        (rc = ensure_in(h, A, A.b_scratch, (size_t)F * 8)) || (rc = ensure_in(h, A, A.b_items, (size_t)F * 8)) ||
        (rc = ensure_in(h, A, A.b_imask, (carved && !f.keep_masks) ? 0 : (size_t)F * 8 * NW)) || (rc = ensure_in(h, A, A.b_icov, (size_t)F * 8)) || (rc = ensure_in(h, A, A.b_bicnt, ((size_t)NB + 1) * 4)) ||
        (rc = ensure_in(h, A, A.b_bmcnt, ((size_t)NB + 1) * 4)) || (rc = ensure_in(h, A, A.b_bext, (size_t)NB * 4)) ||
        (rc = ensure_in(h, A, A.b_big_list, (size_t)NB * 4)) || (rc = ensure_in(h, A, A.b_nz_flag, ((size_t)NB + 1) * 4)) ||
        (rc = ensure_in(h, A, A.b_nz_off, ((size_t)NB + 1) * 4)) || (rc = ensure_in(h, A, A.b_nzk0, (size_t)NB * 4)) ||
        (rc = ensure_in(h, A, A.b_nzk1, (size_t)NB * 4)) || (rc = ensure_in(h, A, A.b_nzv0, (size_t)NB * 4)) ||
        (rc = ensure_in(h, A, A.b_nzv1, (size_t)NB * 4)) || (rc = ensure_in(h, A, A.b_rs_hist, 256 * rs_tiles * 4 + 16)) ||
        (rc = ensure_in(h, A, A.b_rs_hoff, (256 * rs_tiles + 1) * 4 + 16)) || (rc = ensure_in(h, A, A.b_rowbeg, R * 4)) ||
        (rc = ensure_in(h, A, A.b_rowend, R * 4)) || (rc = ensure_in(h, A, A.b_rrec, (size_t)NB * 16)) ||
        (rc = ensure_in(h, A, A.b_btile, (size_t)NB * 8)) || (rc = ensure_in(h, A, A.b_rtile, (size_t)NB * 8)) ||
        (rc = ensure_in(h, A, A.b_rpath, f.ms ? (size_t)NB * 4 : 0)) || (rc = ensure_in(h, A, A.b_ptab, f.ms ? (size_t)P * 1024 : 0)) ||
        (rc = ensure_in(h, A, A.b_scan_tmp, scan_need * 4)))
        return rc;
    f.binfo = P_<int2>(A.b_binfo); f.bcount = P_<uint32_t>(A.b_bcount); f.boff = P_<uint32_t>(A.b_boff);
    f.bwsum = P_<uint32_t>(A.b_bwsum); f.bcarry = P_<uint32_t>(A.b_bcarry);
    f.rec_meta = P_<uint32_t>(A.b_rec_meta); f.rec_mask = P_<uint32_t>(A.b_rec_mask);
    f.scratch_keys = P_<uint32_t>(A.b_scratch); f.items = P_<uint32_t>(A.b_items); f.imask = P_<uint32_t>(A.b_imask);
    f.icov = P_<uint32_t>(A.b_icov);
    // the rank array lives in the large-bucket scratch (used only later, by the resolve);
    // the winding scan's output lives in the item masks the resolve overwrites
    f.frank = f.scratch_keys;
    f.bicnt = P_<uint32_t>(A.b_bicnt); f.bmcnt = P_<uint32_t>(A.b_bmcnt); f.bext = P_<uint32_t>(A.b_bext);
    f.big_list = P_<uint32_t>(A.b_big_list); f.nz_flag = P_<uint32_t>(A.b_nz_flag); f.nz_off = P_<uint32_t>(A.b_nz_off);
    f.nz_keys[0] = P_<uint32_t>(A.b_nzk0); f.nz_keys[1] = P_<uint32_t>(A.b_nzk1);
    f.nz_vals[0] = P_<uint32_t>(A.b_nzv0); f.nz_vals[1] = P_<uint32_t>(A.b_nzv1);
    f.rs_hist = P_<uint32_t>(A.b_rs_hist); f.rs_hoff = P_<uint32_t>(A.b_rs_hoff);
    f.rowbeg = P_<uint32_t>(A.b_rowbeg); f.rowend = P_<uint32_t>(A.b_rowend);
    f.rrec = P_<uint4>(A.b_rrec);
    f.btile = P_<uint64_t>(A.b_btile);
    f.rtile = P_<uint64_t>(A.b_rtile);
    f.rpath = f.ms ? P_<uint32_t>(A.b_rpath) : nullptr;
    f.ptab = f.ms ? P_<uint32_t>(A.b_ptab) : nullptr;
    f.scan_tmp = P_<uint32_t>(A.b_scan_tmp);

    sl::launch_bucket_info(sc, f, s);
    if ((rc = check_launch(h, "bucket_info"))) return rc;
    mark(4);
    sl::launch_frag_count(sc, tb, f, s);
    sl::scan_exclusive_u32(f.bcount, f.boff, NB, f.scan_tmp, s);
    sl::launch_frag_scatter(sc, tb, f, s);
    sl::launch_bucket_carry(f, s);
    sl::launch_scene_carry(sc, f, s);
    if ((rc = check_launch(h, "fragments"))) return rc;
    mark(5);
    sl::launch_resolve(sc, tb, f, s);
    if ((rc = check_launch(h, "resolve"))) return rc;
    mark(6);
    sl::launch_order_rows(sc, f, s);
    if ((rc = check_launch(h, "order_rows"))) return rc;
    if (!carved) h->nz_vals_sorted = f.nz_vals[0];
    mark(7);
    sl::launch_composite(sc, tb, f, s);
    if ((rc = check_launch(h, "composite"))) return rc;
    mark(8);
    // a single-scene frame (batches and parity dumps keep the synchronous read below) hands
    // its status back without waiting: a replay its error flags, an eager frame also the
    // element counts; the host reads them when it settles the frame
    const bool deferred = !h->nscenes && !h->debug;
    if (replay) {
        if (deferred) {
            // words 4..7 (the counts) keep what an eager frame still pending wrote: a replay's
            // counts are its eager frame's
            HIPCHK(h, hipMemcpyAsync(h->fstat, A.b_small.p, 8, hipMemcpyDeviceToHost, s));
            h->stat_counts = h->stat_pending && h->stat_counts;
            h->stat_pending = true;
        }
        return SL_OK;
    }

    // element counts (M, S) and the error flag
    unsigned long long *red = (unsigned long long *)(P_<uint32_t>(A.b_small) + 4);
    sl::launch_fill_u32((uint32_t *)red, 0u, 4, s);
    sl::launch_count_reduce(f.bmcnt, f.bicnt, NB, red, s);
    if (deferred) {
        HIPCHK(h, hipMemcpyAsync(h->fstat, A.b_small.p, 32, hipMemcpyDeviceToHost, s));
        h->stat_pending = true;
        h->stat_counts = true;
        h->timing_pending = h->timing;
        if (!carved) { h->last_F = F; h->last_NCH = NCH; h->last_NB = NB; h->sized = true; }
        return SL_OK;
    }
    uint32_t small[8];
    HIPCHK(h, hipMemcpyAsync(small, A.b_small.p, 32, hipMemcpyDeviceToHost, s));
    HIPCHK(h, hipStreamSynchronize(s));
    const unsigned long long M = *(unsigned long long *)&small[4], MS = *(unsigned long long *)&small[6];
    h->counts.merged_fragment = (int64_t)M;
    h->counts.span = (int64_t)(MS - M);
    if (small[1] & 1u) return fail(h, SL_E_HIP, "internal error: fragment row outside its path's bucket range");
    if (small[1] & 2u) return fail(h, SL_E_UNSUPPORTED, "a (path, row) bucket holds more than 65536 fragments");
    if (!carved) { h->last_F = F; h->last_NCH = NCH; h->last_NB = NB; h->sized = true; }

    if (h->timing) {
        for (int i = 0; i < kStages; ++i) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, h->ev[i], h->ev[i + 1]);
            h->stage_ms[i] = ms;
        }
    }

    if (h->debug) {
        // parity dumps in the reference's layouts
        h->d_ts.resize((size_t)F * 2);
        HIPCHK(h, hipMemcpy(h->d_ts.data(), f.ts, (size_t)F * 8, hipMemcpyDeviceToHost));
        DevBuf dyx, dpid, dwn, dmask, dmoff, dioff, ddraw, dmm;
        int r2 = ensure(h, dyx, (size_t)F * 4);
        if (!r2) r2 = ensure(h, dpid, (size_t)F * 4);
        if (!r2) r2 = ensure(h, dwn, (size_t)F);
        if (!r2) r2 = ensure(h, dmask, (size_t)F * 16);
        if (!r2) r2 = ensure(h, dmoff, ((size_t)NB + 1) * 4);
        if (!r2) r2 = ensure(h, dioff, ((size_t)NB + 1) * 4);
        if (!r2) r2 = ensure(h, ddraw, (size_t)std::max<unsigned long long>(MS, 1) * 16);
        if (!r2) r2 = ensure(h, dmm, (size_t)std::max<unsigned long long>(M, 1) * 16);
        if (!r2) {
            sl::launch_debug_fragments(sc, tb, f, P_<int32_t>(dyx), P_<int32_t>(dpid), P_<int8_t>(dwn),
                                       P_<uint32_t>(dmask), s);
            sl::scan_exclusive_u32(f.bmcnt, P_<uint32_t>(dmoff), NB, f.scan_tmp, s);
            sl::scan_exclusive_u32(f.bicnt, P_<uint32_t>(dioff), NB, f.scan_tmp, s);
            sl::launch_debug_draw_list(sc, tb, f, P_<uint32_t>(dmoff), P_<uint32_t>(dioff), P_<int4>(ddraw),
                                       P_<uint32_t>(dmm), s);
            r2 = check_launch(h, "debug");
        }
        if (!r2) {
            h->d_yx.resize(F); h->d_pid.resize(F); h->d_wn.resize(F); h->d_mask.resize((size_t)F * 4);
            h->d_draw.resize((size_t)MS * 4); h->d_mmask.resize((size_t)M * 4);
            hipError_t e = hipMemcpy(h->d_yx.data(), dyx.p, (size_t)F * 4, hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(h->d_pid.data(), dpid.p, (size_t)F * 4, hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(h->d_wn.data(), dwn.p, (size_t)F, hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(h->d_mask.data(), dmask.p, (size_t)F * 16, hipMemcpyDeviceToHost);
            if (e == hipSuccess && MS) e = hipMemcpy(h->d_draw.data(), ddraw.p, (size_t)MS * 16, hipMemcpyDeviceToHost);
            if (e == hipSuccess && M) e = hipMemcpy(h->d_mmask.data(), dmm.p, (size_t)M * 16, hipMemcpyDeviceToHost);
            if (e != hipSuccess) r2 = fail(h, SL_E_HIP, "debug copy failed: %s", hipGetErrorString(e));
        }
        if (!r2) {
            // internal bucket / row-order arrays for diagnosing the compositor
            uint32_t NZ = 0;
            hipError_t e = hipMemcpy(&NZ, f.nz_off + NB, 4, hipMemcpyDeviceToHost);
            h->d_binfo.resize((size_t)NB * 2); h->d_bicnt.resize(NB); h->d_bext.resize(NB);
            h->d_boff.resize((size_t)NB + 1); h->d_rowkeys.resize(NZ); h->d_rowvals.resize(NZ);
            h->d_rowbeg.resize(R); h->d_rowend.resize(R);
            if (e == hipSuccess) e = hipMemcpy(h->d_binfo.data(), f.binfo, (size_t)NB * 8, hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(h->d_bicnt.data(), f.bicnt, (size_t)NB * 4, hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(h->d_bext.data(), f.bext, (size_t)NB * 4, hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(h->d_boff.data(), f.boff, ((size_t)NB + 1) * 4, hipMemcpyDeviceToHost);
            if (e == hipSuccess && NZ) e = hipMemcpy(h->d_rowkeys.data(), f.nz_keys[0], (size_t)NZ * 4, hipMemcpyDeviceToHost);
            if (e == hipSuccess && NZ) e = hipMemcpy(h->d_rowvals.data(), f.nz_vals[0], (size_t)NZ * 4, hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(h->d_rowbeg.data(), f.rowbeg, R * 4, hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(h->d_rowend.data(), f.rowend, R * 4, hipMemcpyDeviceToHost);
            if (e != hipSuccess) r2 = fail(h, SL_E_HIP, "debug copy failed: %s", hipGetErrorString(e));
        }
        int64_t pc = 0;
        for (size_t k = 0; k < (size_t)MS; ++k) pc += 2 * (int64_t)h->d_draw[4 * k + 1];
        h->pixel_count = pc;
        release(dyx); release(dpid); release(dwn); release(dmask); release(dmoff); release(dioff);
        release(ddraw); release(dmm);
        if (r2) return r2;
    }
    return SL_OK;
}

}  // namespace

extern "C" {

const char *sl_version(void) { return "scanline-mi355x 0.1 (gfx950)"; }

// Wait for a queued frame to finish: it reads the scene and writes the frame buffers, so
// the host must not overwrite or read them before it has passed.
static int settle(sl_handle *h) {
    if (h->pending) {
        HIPCHK(h, hipEventSynchronize(h->done_ev));
        h->pending = false;
    }
    collect_status(h);
    return SL_OK;
}

int sl_create(int device, sl_handle **out) {
    if (!out) return SL_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return SL_E_NODEVICE;
    }
    if (device < 0) device = 0;
    if (device >= n) return SL_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return SL_E_HIP;
    sl_handle *h = new (std::nothrow) sl_handle;
    if (!h) return SL_E_OOM;
    h->device = device;
    for (int i = 0; i <= kStages; ++i) {
        if (hipEventCreate(&h->ev[i]) != hipSuccess) {
            for (int k = 0; k < i; ++k) (void)hipEventDestroy(h->ev[k]);
            delete h;
            return SL_E_HIP;
        }
    }
    if (hipEventCreateWithFlags(&h->done_ev, hipEventDisableTiming) != hipSuccess) {
        for (int k = 0; k <= kStages; ++k) (void)hipEventDestroy(h->ev[k]);
        delete h;
        return SL_E_HIP;
    }
    if (hipEventCreateWithFlags(&h->switch_ev, hipEventDisableTiming) != hipSuccess) {
        for (int k = 0; k <= kStages; ++k) (void)hipEventDestroy(h->ev[k]);
        (void)hipEventDestroy(h->done_ev);
        delete h;
        return SL_E_HIP;
    }
    h->events_ready = true;
    if (hipHostMalloc((void **)&h->fstat, 64, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        for (int k = 0; k <= kStages; ++k) (void)hipEventDestroy(h->ev[k]);
        (void)hipEventDestroy(h->done_ev);
        (void)hipEventDestroy(h->switch_ev);
        delete h;
        return SL_E_OOM;
    }
    memset(h->fstat, 0, 64);
    // set-up of the allocation-free frame: this GPU's address of the static scratch
    if (sl::noalloc_scratch(device, &h->noalloc_scratch) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipHostFree(h->fstat);
        for (int k = 0; k <= kStages; ++k) (void)hipEventDestroy(h->ev[k]);
        (void)hipEventDestroy(h->done_ev);
        (void)hipEventDestroy(h->switch_ev);
        delete h;
        return SL_E_HIP;
    }
    *out = h;
    return SL_OK;
}

int sl_destroy(sl_handle *h) {
    if (!h) return SL_OK;
    (void)hipSetDevice(h->device);
    (void)hipDeviceSynchronize();
    DevBuf *bufs[] = {&h->A.b_pcnt, &h->A.b_nchain, &h->A.b_choff, &h->A.b_chains, &h->A.b_chain_t, &h->vtx, &h->vpath, &h->cvpos, &h->ctype, &h->cpath, &h->cw, &h->crev, &h->cinfo, &h->frule, &h->finfo,
                      &h->ptable[0], &h->ptable[1], &h->ptable[2], &h->ctab[0], &h->ctab[1], &h->A.b_tvtx, &h->A.b_pflags,
                      &h->A.b_prow, &h->A.b_crec, &h->A.b_cfrag, &h->A.b_coff, &h->A.b_ts, &h->A.b_pnb, &h->A.b_pbase,
                      &h->A.b_binfo, &h->A.b_bcount, &h->A.b_boff, &h->A.b_bwsum, &h->A.b_bcarry, &h->A.b_rec_meta,
                      &h->A.b_rec_mask, &h->A.b_scratch, &h->A.b_items, &h->A.b_imask, &h->A.b_icov, &h->A.b_bicnt, &h->A.b_bmcnt, &h->A.b_bext,
                      &h->A.b_big_list, &h->A.b_small, &h->A.b_nz_flag, &h->A.b_nz_off, &h->A.b_nzk0, &h->A.b_nzk1, &h->A.b_nzv0,
                      &h->A.b_nzv1, &h->A.b_rs_hist, &h->A.b_rs_hoff, &h->A.b_rowbeg, &h->A.b_rowend, &h->A.b_image, &h->A.b_scan_tmp,
                      &h->A.b_rrec, &h->A.b_btile, &h->A.b_rtile, &h->A.b_rpath, &h->A.b_ptab, &h->A.b_lcount, &h->pcanvas, &h->stile, &h->simg, &h->sfirst, &h->scarry,
                      &h->gramp, &h->gtable, &h->chord_keys};
    for (DevBuf *b : bufs) release(*b);
    if (h->fstat) (void)hipHostFree(h->fstat);
    if (h->stage) (void)hipHostFree(h->stage);
    if (h->events_ready) {
        for (int i = 0; i <= kStages; ++i) (void)hipEventDestroy(h->ev[i]);
        (void)hipEventDestroy(h->done_ev);
        (void)hipEventDestroy(h->switch_ev);
    }
    delete h;
    return SL_OK;
}

int sl_set_samples(sl_handle *h, int samples) {
    if (!h) return SL_E_ARG;
    if (samples != 8 && samples != 16 && samples != 32)
        return fail(h, SL_E_UNSUPPORTED, "samples must be 8, 16 or 32 (got %d)", samples);
    h->spp = samples;
    h->sized = false;
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, SL_E_HIP, "hipSetDevice failed");
    return ensure_tables(h);  // built here, so a later allocation-free frame finds them
}

int sl_set_multisample_output(sl_handle *h, int enable) {
    if (!h) return SL_E_ARG;
    h->ms = enable ? 1 : 0;
    h->sized = false;
    return SL_OK;
}

int sl_enable_srgb_correction(sl_handle *h, int enable) {
    if (!h) return SL_E_ARG;
    h->srgb = enable ? 1 : 0;
    h->sized = false;
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, SL_E_HIP, "hipSetDevice failed");
    return ensure_tables(h);
}

int sl_set_input_transform(sl_handle *h, const float m[16]) {
    if (!h || !m) return SL_E_ARG;
    std::memcpy(h->m, m, sizeof h->m);
    h->sized = false;
    return SL_OK;
}

int sl_resize_output_buffer(sl_handle *h, int width, int height) {
    if (!h) return SL_E_ARG;
    if (width <= 0 || height <= 0 || width > 32766 || height > 32766)
        return fail(h, SL_E_ARG, "output size %dx%d outside 1..32766", width, height);
    h->W = width;
    h->H = height;
    h->sized = false;
    return SL_OK;
}

// The image of the last frame, if the output size is still the one it was rendered at.
static int check_image(sl_handle *h) {
    if (!h->A.b_image.p || h->img_W <= 0) return fail(h, SL_E_STATE, "no image rendered yet");
    if (h->img_W != h->W || h->img_H != h->H)
        return fail(h, SL_E_STATE, "output size changed to %dx%d since the last frame (%dx%d): rasterize first",
                    h->W, h->H, h->img_W, h->img_H);
    return SL_OK;
}

int sl_viewport(sl_handle *h, int x, int y, int width, int height) {
    if (!h) return SL_E_ARG;
    if (width < 0 || height < 0) return fail(h, SL_E_ARG, "viewport size %dx%d is negative", width, height);
    h->vp_x = x;
    h->vp_y = y;
    h->vp_w = width;
    h->vp_h = height;
    return SL_OK;
}

int sl_resize_viewport(sl_handle *h, int width, int height) {
    if (!h) return SL_E_ARG;
    if (width < 0 || height < 0) return fail(h, SL_E_ARG, "viewport size %dx%d is negative", width, height);
    h->vp_w = width;
    h->vp_h = height;
    return SL_OK;
}

int sl_present(sl_handle *h, void *dev_dst, int dst_pitch, int dst_width, int dst_height, void *stream) {
    if (!h || dst_width < 0 || dst_height < 0) return SL_E_ARG;
    if (!dev_dst && dst_width && dst_height) return fail(h, SL_E_ARG, "destination pointer is NULL");
    if (dst_pitch < dst_width * 4) return fail(h, SL_E_ARG, "pitch %d is below width * 4", dst_pitch);
    int rc = check_image(h);
    if (rc) return rc;
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, SL_E_HIP, "hipSetDevice failed");
    // the source rectangle (0, 0, vp_w, vp_h) of the image goes to (vp_x, vp_y), clipped
    // to both the image and the destination, one pixel for one (GL_NEAREST, same size)
    const int vw = h->vp_w ? h->vp_w : h->img_W, vh = h->vp_h ? h->vp_h : h->img_H;
    int sx0 = 0, sy0 = 0, dx0 = h->vp_x, dy0 = h->vp_y;
    int w = std::min(vw, h->img_W), hh = std::min(vh, h->img_H);
    if (dx0 < 0) { sx0 -= dx0; w += dx0; dx0 = 0; }
    if (dy0 < 0) { sy0 -= dy0; hh += dy0; dy0 = 0; }
    w = std::min(w, dst_width - dx0);
    hh = std::min(hh, dst_height - dy0);
    if (w <= 0 || hh <= 0) return SL_OK;
    const hipStream_t s = (hipStream_t)stream;
    if (h->pending && s != h->pending_stream) HIPCHK(h, hipStreamWaitEvent(s, h->done_ev, 0));
    const size_t spitch = (size_t)h->img_W * 4;
    HIPCHK(h, hipMemcpy2DAsync((uint8_t *)dev_dst + (size_t)dy0 * dst_pitch + (size_t)dx0 * 4, (size_t)dst_pitch,
                               (const uint8_t *)h->A.b_image.p + (size_t)sy0 * spitch + (size_t)sx0 * 4, spitch,
                               (size_t)w * 4, (size_t)hh, hipMemcpyDeviceToDevice, s));
    return SL_OK;
}

int sl_add_vg(sl_handle *h, const sl_scene *s) {
    if (!h) return SL_E_ARG;
    h->sized = false;
    if (!s) return fail(h, SL_E_ARG, "scene is NULL");
    const uint32_t P = s->n_paths, C = s->n_curves, V = s->n_vertices;
    if ((C && (!s->curve_vertex_pos || !s->curve_type || !s->curve_path_id)) || (V && !s->vertex) ||
        (P && (!s->fill_rule || !s->fill_color)))
        return fail(h, SL_E_ARG, "scene array pointer is NULL");
    if (P == 0 && C != 0) return fail(h, SL_E_ARG, "curves without paths");
    if ((s->gradient_ramp_size && !s->gradient_ramp) || (s->gradient_table_size && !s->gradient_table))
        return fail(h, SL_E_ARG, "gradient texture pointer is NULL");
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, SL_E_HIP, "hipSetDevice failed");
    int src = settle(h);
    if (src) return src;
    // The arrays are packed into one pinned staging block (no page faults after the first
    // scene of a size, DMA without the runtime's bounce copies), then copied asynchronously.
    auto al = [](size_t x) { return (x + 63) & ~(size_t)63; };
    const size_t o_vtx = 0, o_vpath = o_vtx + al(8 * (size_t)V), o_cvpos = o_vpath + al(4 * (size_t)V),
                 o_cpath = o_cvpos + al(4 * (size_t)C), o_cw = o_cpath + al(4 * (size_t)C),
                 o_cinfo = o_cw + al(4 * (size_t)C), o_finfo = o_cinfo + al(16 * (size_t)C),
                 o_ctype = o_finfo + al(4 * (size_t)P), o_crev = o_ctype + al(C), o_frule = o_crev + al(C),
                 total = o_frule + al(P) + 64;
    if (h->stage_cap < total) {
        if (h->stage) (void)hipHostFree(h->stage);
        h->stage = nullptr;
        h->stage_cap = 0;
        const size_t want = total + total / 4;
        if (hipHostMalloc(&h->stage, want, hipHostMallocDefault) != hipSuccess) {
            h->stage = nullptr;
            return fail(h, SL_E_HIP, "hipHostMalloc of %zu bytes failed", want);
        }
        h->stage_cap = want;
    }
    uint8_t *st = (uint8_t *)h->stage;
    float *vtx = (float *)(st + o_vtx);
    uint32_t *vpath = (uint32_t *)(st + o_vpath);
    uint8_t *crev = st + o_crev;
    float *cw = (float *)(st + o_cw);
    // Validation and packing run over chunks of curves on host threads.  Errors are the
    // ones a curve-by-curve pass would report first: the lowest failing curve, and a vertex
    // shared by two curves is reported at the later curve (vertices go to their lowest
    // curve by an atomic minimum, then every curve checks that it owns its vertices).
    uint32_t *owner = vpath;  // curve owning each vertex; turned into path ids at the end
    uint32_t *cinfo = (uint32_t *)(st + o_cinfo);
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const unsigned T = (C >= 65536u && hw > 1) ? hw : 1u;
    auto for_chunks = [&](size_t n, auto &&body) {
        // body(t, lo, hi) on T contiguous chunks of [0, n)
        if (T == 1 || n < 4096) { body(0u, (size_t)0, n); return; }
        std::thread th[16];
        bool started[16] = {};
        for (unsigned t = 0; t < T; ++t) {
            const size_t lo = n * t / T, hi = n * (t + 1) / T;
            try {
                th[t] = std::thread([&body, t, lo, hi] { body(t, lo, hi); });
                started[t] = true;
            } catch (...) {
                body(t, lo, hi);  // no thread to be had (limits): this chunk runs here
            }
        }
        for (unsigned t = 0; t < T; ++t)
            if (started[t]) th[t].join();
    };
    struct Err { uint32_t at = 0xFFFFFFFFu; int kind = 0; uint32_t a = 0, b = 0; };
    std::vector<Err> errs(T);
    auto better = [](const Err &x, const Err &y) { return x.at < y.at; };
    if (V) {
        for_chunks(2 * (size_t)V, [&](unsigned, size_t lo, size_t hi) {
            std::memcpy(vtx + lo, s->vertex + lo, (hi - lo) * 4);
        });
        for_chunks(V, [&](unsigned, size_t lo, size_t hi) { std::fill(owner + lo, owner + hi, 0xFFFFFFFFu); });
    }
    // structural checks, in the order a curve-by-curve pass applies them
    for_chunks(C, [&](unsigned t, size_t lo, size_t hi) {
        for (size_t c = lo; c < hi; ++c) {
            const uint8_t ty = s->curve_type[c];
            Err e;
            const uint32_t p = s->curve_path_id[c];
            const uint32_t vp = s->curve_vertex_pos[c];
            if (ty != SL_CURVE_LINEAR && ty != SL_CURVE_QUADRATIC && ty != SL_CURVE_CUBIC && ty != SL_CURVE_RATIONAL)
                e = {(uint32_t)c, 1, ty, 0};
            else if (p >= P)
                e = {(uint32_t)c, 2, p, 0};
            else if (c > 0 && p < s->curve_path_id[c - 1])
                e = {(uint32_t)c, 3, 0, 0};
            else if ((uint64_t)vp + (ty & 7) > V)
                e = {(uint32_t)c, 4, vp, vp + (ty & 7)};
            if (e.kind) { errs[t] = e; return; }
        }
    });
    Err first = *std::min_element(errs.begin(), errs.end(), better);
    // A decreasing path id right after an earlier failure is that failure's to report; the
    // chunk scan compares with the previous curve's id, a serial pass with the last good
    // one, and both agree up to the first failing curve.
    const uint32_t nchecked = first.kind ? first.at : C;
    // vertex ownership among the curves before the first structural failure
    for_chunks(nchecked, [&](unsigned, size_t lo, size_t hi) {
        for (size_t c = lo; c < hi; ++c) {
            const uint32_t vp = s->curve_vertex_pos[c];
            const int nv = s->curve_type[c] & 7;
            for (int k = 0; k < nv; ++k) {
                uint32_t cur = __atomic_load_n(&owner[vp + k], __ATOMIC_RELAXED);
                while ((uint32_t)c < cur &&
                       !__atomic_compare_exchange_n(&owner[vp + k], &cur, (uint32_t)c, true, __ATOMIC_RELAXED,
                                                    __ATOMIC_RELAXED)) {
                }
            }
        }
    });
    std::vector<Err> cerrs(T);
    for_chunks(nchecked, [&](unsigned t, size_t lo, size_t hi) {
        for (size_t c = lo; c < hi; ++c) {
            const uint32_t vp = s->curve_vertex_pos[c];
            const int nv = s->curve_type[c] & 7;
            for (int k = 0; k < nv; ++k)
                if (owner[vp + k] != (uint32_t)c) { cerrs[t] = {(uint32_t)c, 5, vp + (uint32_t)k, 0}; return; }
        }
    });
    const Err shared = *std::min_element(cerrs.begin(), cerrs.end(), better);
    if (shared.kind && shared.at < first.at) first = shared;
    switch (first.kind) {
    case 1: return fail(h, SL_E_ARG, "curve %u has invalid type 0x%x", first.at, first.a);
    case 2: return fail(h, SL_E_ARG, "curve %u has path id %u >= n_paths %u", first.at, first.a, P);
    case 3: return fail(h, SL_E_ARG, "curve %u: curve_path_id must be non-decreasing", first.at);
    case 4: return fail(h, SL_E_ARG, "curve %u vertices [%u, %u) exceed n_vertices %u", first.at, first.a, first.b, V);
    case 5: return fail(h, SL_E_ARG, "vertex %u belongs to two curves", first.a);
    default: break;
    }
    // orientation, weights and the packed curve words; every vertex now has one curve
    for_chunks(C, [&](unsigned, size_t lo, size_t hi) {
        for (size_t c = lo; c < hi; ++c) {
            const uint8_t ty = s->curve_type[c];
            const uint32_t vp = s->curve_vertex_pos[c];
            const int nv = ty & 7;
            cw[c] = (ty == SL_CURVE_RATIONAL) ? (s->curve_arc_w ? s->curve_arc_w[c] : 1.f) : -1.f;
            // unifyCurveVertexOrder (ras_base.cpp:523-569): order by float bit patterns
            uint32_t ixf, ixl, iyf, iyl;
            std::memcpy(&ixf, &vtx[2 * (size_t)vp], 4);
            std::memcpy(&iyf, &vtx[2 * (size_t)vp + 1], 4);
            std::memcpy(&ixl, &vtx[2 * (size_t)(vp + nv - 1)], 4);
            std::memcpy(&iyl, &vtx[2 * (size_t)(vp + nv - 1) + 1], 4);
            crev[c] = 0;
            if (!(iyf < iyl || (iyf == iyl && ixf < ixl))) {
                crev[c] = 1;
                for (int k = 0; k < nv / 2; ++k) {
                    std::swap(vtx[2 * (size_t)(vp + k)], vtx[2 * (size_t)(vp + nv - 1 - k)]);
                    std::swap(vtx[2 * (size_t)(vp + k) + 1], vtx[2 * (size_t)(vp + nv - 1 - k) + 1]);
                }
            }
            uint32_t wb;
            std::memcpy(&wb, &cw[c], 4);
            cinfo[4 * c] = vp;
            cinfo[4 * c + 1] = s->curve_path_id[c];
            cinfo[4 * c + 2] = (uint32_t)ty | ((uint32_t)crev[c] << 8);
            cinfo[4 * c + 3] = wb;
        }
    });
    // owning path of each vertex; vertices of no curve: harmless, never read
    for_chunks(V, [&](unsigned, size_t lo, size_t hi) {
        for (size_t v = lo; v < hi; ++v) owner[v] = owner[v] == 0xFFFFFFFFu ? 0u : s->curve_path_id[owner[v]];
    });
    // fill words (ras_base.cpp:144-166): a colour keeps its value unless its alpha is 0; with
    // a gradient table, a value in [1, 0x01000000] names gradient slot value - 1, which the
    // compositor evaluates per pixel (output_32.vert.glsl:55-97)
    uint32_t *finfo = (uint32_t *)(st + o_finfo);
    const uint32_t gslots = s->gradient_table_size / 12;
    bool has_grad = false;
    for (uint32_t p = 0; p < P; ++p) {
        const uint32_t c = s->fill_color[p];
        const bool grad = !(c & 0xFF000000u) && gslots && c != 0u;
        finfo[p] = ((c & 0xFF000000u) || grad) ? c : 0u;
        has_grad |= grad;
    }
    if (C) {
        std::memcpy(st + o_cvpos, s->curve_vertex_pos, 4 * (size_t)C);
        std::memcpy(st + o_cpath, s->curve_path_id, 4 * (size_t)C);
        std::memcpy(st + o_ctype, s->curve_type, C);
    }
    if (P) std::memcpy(st + o_frule, s->fill_rule, P);
    struct Part { DevBuf *b; size_t off, bytes; };
    const Part parts[] = {{&h->vtx, o_vtx, 8 * (size_t)V},       {&h->vpath, o_vpath, 4 * (size_t)V},
                          {&h->cvpos, o_cvpos, 4 * (size_t)C},   {&h->ctype, o_ctype, C},
                          {&h->cpath, o_cpath, 4 * (size_t)C},   {&h->cw, o_cw, 4 * (size_t)C},
                          {&h->crev, o_crev, C},                 {&h->cinfo, o_cinfo, 16 * (size_t)C},
                          {&h->frule, o_frule, P},               {&h->finfo, o_finfo, 4 * (size_t)P}};
    // from here on the old scene's arrays may be replaced: a failure leaves no scene
    h->has_scene = false;
    h->P = h->C = h->V = 0;
    for (const Part &q : parts) {
        int rc = ensure(h, *q.b, q.bytes);
        if (rc) return rc;
    }
    // the null stream orders these copies after earlier work as hipMemcpy does
    for (const Part &q : parts)
        if (q.bytes) HIPCHK(h, hipMemcpyAsync(q.b->p, st + q.off, q.bytes, hipMemcpyHostToDevice, nullptr));
    HIPCHK(h, hipStreamSynchronize(nullptr));
    // gradient textures (ras_base.cpp:339-371): the ramp has size / 1024 whole rows
    const uint32_t rows = s->gradient_ramp_size >> 10;
    h->gslots = has_grad ? gslots : 0;
    h->gramp_rows = has_grad ? rows : 0;
    h->has_grad = has_grad;
    if (has_grad) {
        int rc = upload(h, h->gtable, s->gradient_table, (size_t)gslots * 48);
        if (!rc && rows) rc = upload(h, h->gramp, s->gradient_ramp, (size_t)rows * 4096);
        if (rc) return rc;
    }
    h->P = P; h->C = C; h->V = (P == 0) ? 0 : V;
    h->has_scene = true;
    return ensure_tables(h);
}

int sl_clear(sl_handle *h) {
    if (!h) return SL_E_ARG;
    h->has_scene = false;
    h->P = h->C = h->V = 0;
    h->sized = false;
    return SL_OK;
}

int sl_chord_load(sl_handle *h, const sl_chord *d, int n) {
    if (!h) return SL_E_ARG;
    if (!d || n < 1 || n > sl::kMaxChordDiagrams)
        return fail(h, SL_E_ARG, "need 1 to %d chord diagrams (got %d)", sl::kMaxChordDiagrams, d ? n : 0);
    // validation first: nothing changes on failure
    uint64_t paths = 0;
    for (int i = 0; i < n; ++i) {
        const sl_chord &c = d[i];
        if (c.n_keys < 2 || c.n_nodes < 0) return fail(h, SL_E_ARG, "diagram %d: n_keys %d < 2 or n_nodes %d < 0", i, c.n_keys, c.n_nodes);
        if (!c.link_offset) return fail(h, SL_E_ARG, "diagram %d: link_offset is NULL", i);
        if (c.link_offset[0] != 0) return fail(h, SL_E_ARG, "diagram %d: link_offset[0] must be 0", i);
        uint32_t most = 0;
        for (int k = 1; k < c.n_keys; ++k) {
            if (c.link_offset[k] < c.link_offset[k - 1])
                return fail(h, SL_E_ARG, "diagram %d: link_offset decreases at keyframe %d", i, k);
            most = std::max(most, c.link_offset[k] - c.link_offset[k - 1]);
        }
        const uint32_t L = c.link_offset[c.n_keys - 1];
        if (c.n_nodes && (!c.node_t0 || !c.node_t1 || !c.node_size || !c.node_color))
            return fail(h, SL_E_ARG, "diagram %d: node array pointer is NULL", i);
        if (L && (!c.link_node || !c.link_from || !c.link_to))
            return fail(h, SL_E_ARG, "diagram %d: link array pointer is NULL", i);
        for (uint64_t l = 0; l < 2 * (uint64_t)L; ++l)
            if (c.link_node[l] >= (uint32_t)c.n_nodes)
                return fail(h, SL_E_ARG, "diagram %d: link %llu names node %u >= n_nodes %d", i,
                            (unsigned long long)(l / 2), c.link_node[l], c.n_nodes);
        paths += (uint64_t)c.n_nodes + most;
    }
    if (paths > (1u << 24)) return fail(h, SL_E_UNSUPPORTED, "chord scenes of more than 2^24 paths (%llu)", (unsigned long long)paths);
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, SL_E_HIP, "hipSetDevice failed");
    int rc = settle(h);
    if (rc) return rc;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    std::vector<sl_handle::Chord> ch(n);
    size_t total = 0;
    for (int i = 0; i < n; ++i) {
        const sl_chord &c = d[i];
        sl_handle::Chord &q = ch[i];
        const size_t KN = (size_t)c.n_keys * c.n_nodes, L = c.link_offset[c.n_keys - 1];
        q.n_keys = c.n_keys;
        q.n_nodes = c.n_nodes;
        q.link_offset.assign(c.link_offset, c.link_offset + c.n_keys);
        q.cx = c.center_x;
        q.cy = c.center_y;
        q.o_t0 = total; total += al(4 * KN);
        q.o_t1 = total; total += al(4 * KN);
        q.o_size = total; total += al(4 * KN);
        q.o_color = total; total += al(4 * (size_t)c.n_nodes);
        q.o_node = total; total += al(8 * L);
        q.o_from = total; total += al(16 * L);
        q.o_to = total; total += al(16 * L);
    }
    std::vector<uint8_t> blk;
    try {
        blk.resize(total);
    } catch (...) {  // no exception crosses the C ABI
        return fail(h, SL_E_OOM, "host staging of %zu bytes of keyframes failed", total);
    }
    for (int i = 0; i < n; ++i) {
        const sl_chord &c = d[i];
        const sl_handle::Chord &q = ch[i];
        const size_t KN = (size_t)c.n_keys * c.n_nodes, L = c.link_offset[c.n_keys - 1];
        if (KN) {
            std::memcpy(&blk[q.o_t0], c.node_t0, 4 * KN);
            std::memcpy(&blk[q.o_t1], c.node_t1, 4 * KN);
            std::memcpy(&blk[q.o_size], c.node_size, 4 * KN);
        }
        if (c.n_nodes) std::memcpy(&blk[q.o_color], c.node_color, 4 * (size_t)c.n_nodes);
        if (L) {
            std::memcpy(&blk[q.o_node], c.link_node, 8 * L);
            std::memcpy(&blk[q.o_from], c.link_from, 16 * L);
            std::memcpy(&blk[q.o_to], c.link_to, 16 * L);
        }
    }
    h->has_scene = false;
    h->P = h->C = h->V = 0;
    h->sized = false;
    h->chord.clear();
    if ((rc = upload(h, h->chord_keys, blk.data(), total))) return rc;
    h->chord = std::move(ch);
    return SL_OK;
}

int sl_chord_frame(sl_handle *h, int time_ms, void *stream) {
    if (!h) return SL_E_ARG;
    if (h->chord.empty()) return fail(h, SL_E_STATE, "no chord keyframes loaded (sl_chord_load)");
    if (time_ms < 0) return fail(h, SL_E_ARG, "time %d ms is negative", time_ms);
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, SL_E_HIP, "hipSetDevice failed");
    const int n = (int)h->chord.size();
    sl::ChordInterval iv[sl::kMaxChordDiagrams];
    uint32_t P = 0;
    const uint8_t *base = (const uint8_t *)h->chord_keys.p;
    for (int i = 0; i < n; ++i) {
        const sl_handle::Chord &q = h->chord[i];
        // keyframe pair of the frame (chord_animation, animation.cu:2457-2459); past the
        // last keyframe the diagram rests on it
        int k = time_ms / sl::kChordKeyMs + 1;
        float t = sl::chord_ease(time_ms);
        if (k > q.n_keys - 1) {
            k = q.n_keys - 1;
            t = 1.f;
        }
        const uint32_t l0 = q.link_offset[k - 1], nl = q.link_offset[k] - l0;
        const size_t a = (size_t)(k - 1) * q.n_nodes, b = (size_t)k * q.n_nodes;
        sl::ChordInterval &v = iv[i];
        v.n_nodes = q.n_nodes;
        v.n_links = (int)nl;
        v.t = t;
        v.center = make_float2(q.cx, q.cy);
        v.node_t0_a = (const float *)(base + q.o_t0) + a;
        v.node_t0_b = (const float *)(base + q.o_t0) + b;
        v.node_t1_a = (const float *)(base + q.o_t1) + a;
        v.node_t1_b = (const float *)(base + q.o_t1) + b;
        v.node_size_a = (const float *)(base + q.o_size) + a;
        v.node_size_b = (const float *)(base + q.o_size) + b;
        v.node_color = (const uint32_t *)(base + q.o_color);
        v.link_node = (const uint2 *)(base + q.o_node) + l0;
        v.link_from = (const float4 *)(base + q.o_from) + l0;
        v.link_to = (const float4 *)(base + q.o_to) + l0;
        P += (uint32_t)q.n_nodes + nl;
    }
    const uint32_t C = 6 * P, V = 24 * P;
    const hipStream_t s = (hipStream_t)stream;
    struct Part { DevBuf *b; size_t bytes; };
    const Part parts[] = {{&h->vtx, 8 * (size_t)V}, {&h->vpath, 4 * (size_t)V}, {&h->cvpos, 4 * (size_t)C},
                          {&h->ctype, C},           {&h->cpath, 4 * (size_t)C}, {&h->cw, 4 * (size_t)C},
                          {&h->crev, C},            {&h->cinfo, 16 * (size_t)C}, {&h->frule, P},
                          {&h->finfo, 4 * (size_t)P}};
    bool grow = false;
    for (const Part &q : parts) grow |= q.b->cap < std::max<size_t>(q.bytes, 16);
    if (grow) {
        // a queued frame may still read the arrays about to be replaced
        int rc = settle(h);
        if (rc) return rc;
        // the old scene's arrays may be replaced: a failure leaves no scene
        h->has_scene = false;
        h->P = h->C = h->V = 0;
        for (const Part &q : parts)
            if ((rc = ensure(h, *q.b, q.bytes))) return rc;
    }
    // the scene arrays are rewritten after every frame issued before this call
    if (h->pending && s != h->pending_stream) {
        HIPCHK(h, hipEventRecord(h->switch_ev, h->pending_stream));
        HIPCHK(h, hipStreamWaitEvent(s, h->switch_ev, 0));
    }
    sl::ChordScene out{(float2 *)h->vtx.p,   (uint32_t *)h->vpath.p, (uint32_t *)h->cvpos.p, (uint32_t *)h->cpath.p,
                       (uint32_t *)h->finfo.p, (uint8_t *)h->ctype.p, (uint8_t *)h->frule.p, (uint8_t *)h->crev.p,
                       (float *)h->cw.p,     (uint4 *)h->cinfo.p};
    sl::launch_chord_frame(iv, n, P, out, s);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipEventRecord(h->done_ev, s));
    h->pending = true;
    h->pending_stream = s;
    h->P = P;
    h->C = C;
    h->V = V;
    h->has_scene = true;
    h->sized = false;
    h->has_grad = false;
    h->gslots = 0;
    h->gramp_rows = 0;
    return ensure_tables(h);
}

int sl_rasterize(sl_handle *h, void *stream) {
    if (!h) return SL_E_ARG;
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, SL_E_HIP, "hipSetDevice failed");
    // Once a frame has sized the buffers, the same scene under the same settings is
    // re-rendered without host synchronisation: the launches are those of that frame
    // (every kernel runs again; only its counts and sizes are taken over), queued on the
    // stream, and the call returns at once.  Step timing and parity dumps need the host
    // reads of a full frame, so they keep every frame synchronous.
    // A frame still queued shares this frame's buffers.  On the same stream the order is
    // the stream's.  On another stream, this frame comes after everything queued on the
    // previous frame's stream so far (that frame, and any reads of its image queued after
    // it): on the device for a replay, on the host for a full frame (which may reallocate).
    const hipStream_t s = (hipStream_t)stream;
    const bool replay = h->sized && !h->timing && !h->debug;
    if (h->pending && s != h->pending_stream) {
        HIPCHK(h, hipEventRecord(h->switch_ev, h->pending_stream));
        if (replay) HIPCHK(h, hipStreamWaitEvent(s, h->switch_ev, 0));
        else HIPCHK(h, hipEventSynchronize(h->switch_ev));
        h->pending = false;
    } else if (h->pending && !replay) {
        const int src = settle(h);
        if (src) return src;
    }
    const int rc = render(h, s, replay, h->A);
    // Whether or not the frame succeeded, some of its kernels may be queued on s: later
    // uploads and frames must wait for them as for a completed replay.
    if (hipEventRecord(h->done_ev, s) == hipSuccess) {
        h->pending = true;
        h->pending_stream = s;
    } else {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(s);
        h->pending = false;
    }
    return rc;
}

int sl_rasterize_noalloc(sl_handle *h, void *dev_rgba, int pitch_bytes, void *stream) {
    if (!h) return SL_E_ARG;
    if (!dev_rgba) return fail(h, SL_E_ARG, "output image pointer is NULL");
    if (h->W > 0 && pitch_bytes < h->W * 4)
        return fail(h, SL_E_ARG, "pitch %d is below width * 4 = %d", pitch_bytes, h->W * 4);
    if (h->debug) return fail(h, SL_E_STATE, "debug dumps allocate: disable them for the allocation-free frame");
    if (!h->noalloc_scratch) return fail(h, SL_E_STATE, "no static scratch on this device");
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, SL_E_HIP, "hipSetDevice failed");
    const hipStream_t s = (hipStream_t)stream;
    // the handle's scene is read by a queued frame too (read only); order after it anyway,
    // so that the counts this frame reports are the last ones
    if (h->pending && s != h->pending_stream) {
        HIPCHK(h, hipEventRecord(h->switch_ev, h->pending_stream));
        HIPCHK(h, hipStreamWaitEvent(s, h->switch_ev, 0));
    }
    Arena A;
    A.carve_base = (uint8_t *)h->noalloc_scratch;
    A.carve_cap = sl::noalloc_scratch_bytes() - ((size_t)1 << 20);  // 1 MiB kept free at the end
    const int rc = render(h, s, false, A, (uint8_t *)dev_rgba, pitch_bytes);
    if (rc) {
        // kernels queued before the failure read the scratch: let them drain
        (void)hipStreamSynchronize(s);
        return rc;
    }
    // the frame's status comes back when it is settled, as for sl_rasterize
    if (hipEventRecord(h->done_ev, s) == hipSuccess) {
        h->pending = true;
        h->pending_stream = s;
    } else {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(s);
        h->pending = false;
    }
    return SL_OK;
}

int sl_output_device_ptr(sl_handle *h, void **dev_ptr, int *pitch_bytes) {
    if (!h || !dev_ptr || !pitch_bytes) return SL_E_ARG;
    if (int rc = check_image(h)) return rc;
    *dev_ptr = h->A.b_image.p;
    *pitch_bytes = h->img_W * 4;
    return SL_OK;
}

int sl_read_output(sl_handle *h, uint8_t *host_rgba, int flags) {
    if (!h || !host_rgba) return SL_E_ARG;
    int rc = check_image(h);
    if (rc) return rc;
    if ((rc = settle(h))) return rc;
    if (h->frame_rc) return fail(h, h->frame_rc, "%s", h->frame_err.c_str());
    const size_t row = (size_t)h->img_W * 4;
    HIPCHK(h, hipMemcpy(host_rgba, h->A.b_image.p, row * h->img_H, hipMemcpyDeviceToHost));
    if (flags & 1) flip_rows_alpha(host_rgba, h->img_W, h->img_H);
    return SL_OK;
}

int sl_element_number(sl_handle *h, sl_counts *out) {
    if (!h || !out) return SL_E_ARG;
    if (int rc = settle(h)) return rc;
    *out = h->counts;
    if (h->frame_rc) return fail(h, h->frame_rc, "%s", h->frame_err.c_str());
    return SL_OK;
}

int sl_enable_step_timing(sl_handle *h, int enable) {
    if (!h) return SL_E_ARG;
    h->timing = enable != 0;
    return SL_OK;
}

int sl_step_timing(sl_handle *h, double *ms, int capacity, int *n_stages) {
    if (!h || !n_stages) return SL_E_ARG;
    if (int rc = settle(h)) return rc;
    *n_stages = kStages;
    if (ms)
        for (int i = 0; i < kStages && i < capacity; ++i) ms[i] = h->stage_ms[i];
    return SL_OK;
}

const char *sl_step_name(int stage) { return (stage >= 0 && stage < kStages) ? kStageNames[stage] : ""; }

int sl_pixel_count(sl_handle *h, int64_t *out) {
    if (!h || !out) return SL_E_ARG;
    if (!h->debug) return fail(h, SL_E_STATE, "pixel count needs sl_debug_enable (it reads the draw list back)");
    *out = h->pixel_count;
    return SL_OK;
}

const char *sl_last_error(sl_handle *h) { return h ? h->err.c_str() : "null handle"; }

int sl_debug_enable(sl_handle *h, int enable) {
    if (!h) return SL_E_ARG;
    h->debug = enable != 0;
    h->sized = false;
    return SL_OK;
}

int sl_debug_requantize(sl_handle *h, int channel, const float *values, int64_t n, uint8_t *codes, float *decoded) {
    if (!h || (channel != 0 && channel != 1) || n < 0 || (n && (!values || !codes || !decoded))) return SL_E_ARG;
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, SL_E_HIP, "hipSetDevice failed");
    int rc;
    if ((rc = ensure_tables(h))) return rc;
    if (!n) return SL_OK;
    const sl::Tables tb = tables_view(h);
    void *dv = nullptr, *dc = nullptr, *dd = nullptr;
    hipError_t e = hipMalloc(&dv, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc(&dc, (size_t)n);
    if (e == hipSuccess) e = hipMalloc(&dd, (size_t)n * 4);
    if (e == hipSuccess) e = hipMemcpy(dv, values, (size_t)n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        sl::launch_encode_values(tb, channel, (const float *)dv, n, (uint8_t *)dc, (float *)dd, nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(codes, dc, (size_t)n, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(decoded, dd, (size_t)n * 4, hipMemcpyDeviceToHost);
    (void)hipFree(dv);
    (void)hipFree(dc);
    (void)hipFree(dd);
    if (e != hipSuccess) return fail(h, SL_E_HIP, "requantize: %s", hipGetErrorString(e));
    return SL_OK;
}

int sl_debug_valu_probe(sl_handle *h, int blocks, int iters, double *ms, uint64_t *wave_insts) {
    if (!h || blocks <= 0 || iters <= 0 || !ms || !wave_insts) return SL_E_ARG;
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, SL_E_HIP, "hipSetDevice failed");
    void *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&out, (size_t)blocks * 4);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        sl::launch_valu_probe(blocks, iters, (float *)out, nullptr);  // warm-up
        (void)hipEventRecord(e0, nullptr);
        sl::launch_valu_probe(blocks, iters, (float *)out, nullptr);
        (void)hipEventRecord(e1, nullptr);
        e = hipEventSynchronize(e1);
    }
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (e != hipSuccess) return fail(h, SL_E_HIP, "valu probe: %s", hipGetErrorString(e));
    *ms = t;
    *wave_insts = (uint64_t)blocks * 4ull * (uint64_t)iters * 512ull;
    return SL_OK;
}

int sl_debug_requantize_sweep(sl_handle *h, int channel, uint32_t bits_lo, uint32_t bits_hi, uint64_t *mismatches) {
    if (!h || (channel != 0 && channel != 1) || !mismatches) return SL_E_ARG;
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, SL_E_HIP, "hipSetDevice failed");
    int rc;
    if ((rc = ensure_tables(h))) return rc;
    const sl::Tables tb = tables_view(h);
    void *db = nullptr;
    unsigned long long bad = 0;
    hipError_t e = hipMalloc(&db, 8);
    if (e == hipSuccess) e = hipMemset(db, 0, 8);
    if (e == hipSuccess) {
        sl::launch_encode_sweep(tb, channel, bits_lo, bits_hi, (unsigned long long *)db, nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(&bad, db, 8, hipMemcpyDeviceToHost);
    (void)hipFree(db);
    if (e != hipSuccess) return fail(h, SL_E_HIP, "requantize sweep: %s", hipGetErrorString(e));
    *mismatches = bad;
    return SL_OK;
}

int sl_debug_dump(sl_handle *h, const char *name, void *host_dst, size_t capacity, size_t *nbytes) {
    if (!h || !name || !nbytes) return SL_E_ARG;
    int rc = settle(h);
    if (rc) return rc;
    const void *src = nullptr;
    size_t n = 0;
    std::string k(name);
    if (k == "image") {
        if ((rc = check_image(h))) return rc;
        n = (size_t)h->img_W * h->img_H * 4;
        *nbytes = n;
        if (host_dst && capacity >= n) HIPCHK(h, hipMemcpy(host_dst, h->A.b_image.p, n, hipMemcpyDeviceToHost));
        return SL_OK;
    }
    // the scene arrays as the renderer reads them (an upload's or a generated frame's)
    struct SceneDump { const char *name; const DevBuf *b; size_t bytes; };
    const SceneDump sd[] = {{"scene_vertex", &h->vtx, 8 * (size_t)h->V},  {"scene_vertex_path", &h->vpath, 4 * (size_t)h->V},
                            {"scene_curve", &h->cinfo, 16 * (size_t)h->C}, {"scene_fill_rule", &h->frule, h->P},
                            {"scene_fill", &h->finfo, 4 * (size_t)h->P}};
    for (const SceneDump &q : sd) {
        if (k != q.name) continue;
        n = h->has_scene ? q.bytes : 0;
        *nbytes = n;
        if (host_dst && capacity >= n && n) HIPCHK(h, hipMemcpy(host_dst, q.b->p, n, hipMemcpyDeviceToHost));
        return SL_OK;
    }
    if (k == "pixel_table") {
        sl::build_pixel_table(h->spp, h->d_ptable);
        src = h->d_ptable.data(); n = h->d_ptable.size() * 4;
    } else if (!h->debug) {
        return fail(h, SL_E_STATE, "debug dumps need sl_debug_enable before sl_rasterize");
    } else if (k == "ts") { src = h->d_ts.data(); n = h->d_ts.size() * 4; }
    else if (k == "frag_yx") { src = h->d_yx.data(); n = h->d_yx.size() * 4; }
    else if (k == "frag_pid") { src = h->d_pid.data(); n = h->d_pid.size() * 4; }
    else if (k == "frag_wn") { src = h->d_wn.data(); n = h->d_wn.size(); }
    else if (k == "frag_mask") { src = h->d_mask.data(); n = h->d_mask.size() * 4; }
    else if (k == "merged_mask") { src = h->d_mmask.data(); n = h->d_mmask.size() * 4; }
    else if (k == "draw_list") { src = h->d_draw.data(); n = h->d_draw.size() * 4; }
    else if (k == "bucket_info") { src = h->d_binfo.data(); n = h->d_binfo.size() * 4; }
    else if (k == "bucket_items") { src = h->d_bicnt.data(); n = h->d_bicnt.size() * 4; }
    else if (k == "bucket_ext") { src = h->d_bext.data(); n = h->d_bext.size() * 4; }
    else if (k == "bucket_off") { src = h->d_boff.data(); n = h->d_boff.size() * 4; }
    else if (k == "row_keys") { src = h->d_rowkeys.data(); n = h->d_rowkeys.size() * 4; }
    else if (k == "row_list") { src = h->d_rowvals.data(); n = h->d_rowvals.size() * 4; }
    else if (k == "row_beg") { src = h->d_rowbeg.data(); n = h->d_rowbeg.size() * 4; }
    else if (k == "row_end") { src = h->d_rowend.data(); n = h->d_rowend.size() * 4; }
    else return fail(h, SL_E_ARG, "unknown dump '%s'", name);
    *nbytes = n;
    if (host_dst && capacity >= n && n) std::memcpy(host_dst, src, n);
    return SL_OK;
}


/* ---- batched scenes ------------------------------------------------------------------ */

}  // extern "C"

struct sl_batch {
    int device = 0;
    int n = 0;
    std::vector<sl_handle *> grp;        // one internal handle per sample count present
    std::vector<int> grp_spp;
    std::vector<int> scene_grp, scene_w, scene_h;
    std::vector<uint64_t> scene_off;      // byte offset of each scene's image in its group
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int shape[4] = {0, 0, 4, 256};        // resolve grid, class grid, compositor group, fragment block
    // the last frame issued by sl_batch_render(_eager), possibly still running
    hipEvent_t done_ev = nullptr;
    hipEvent_t switch_ev = nullptr;       // recorded on last_stream when a frame moves on
    hipStream_t last_stream = nullptr;
    bool pending = false;
    std::string err;
};

namespace {

int bfail(sl_batch *b, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    return code;
}

// Every frame of a batch uses the same buffers.  A frame on another stream than the last
// one waits on the device for everything queued on that stream so far (the last frame and
// any reads of its images queued after it); tuning and recapture wait on the host.
int batch_order(sl_batch *b, hipStream_t s) {
    if (b->pending && s != b->last_stream) {
        if (hipEventRecord(b->switch_ev, b->last_stream) != hipSuccess ||
            hipStreamWaitEvent(s, b->switch_ev, 0) != hipSuccess)
            return bfail(b, SL_E_HIP, "ordering after the previous stream failed");
    }
    return SL_OK;
}

int batch_issued(sl_batch *b, hipStream_t s) {
    if (hipEventRecord(b->done_ev, s) != hipSuccess) return bfail(b, SL_E_HIP, "hipEventRecord failed");
    b->pending = true;
    b->last_stream = s;
    return SL_OK;
}

int batch_settle(sl_batch *b) {
    if (b->pending && hipEventSynchronize(b->done_ev) != hipSuccess) return bfail(b, SL_E_HIP, "render failed");
    b->pending = false;
    return SL_OK;
}

void drop_graph(sl_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void apply_shape(sl_batch *b) {
    for (sl_handle *h : b->grp) {
        h->resolve_grid = b->shape[0];
        h->class_grid = b->shape[1];
        h->comp_group = b->shape[2];
        h->frag_block = b->shape[3];
    }
}

int enqueue_all(sl_batch *b, hipStream_t s, bool replay) {
    for (sl_handle *h : b->grp) {
        const int rc = render(h, s, replay, h->A);
        if (rc) return bfail(b, rc, "%s", h->err.c_str());
    }
    return SL_OK;
}

int capture(sl_batch *b) {
    drop_graph(b);
    hipError_t e = hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return bfail(b, SL_E_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    int rc = enqueue_all(b, b->cap_stream, true);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->cap_stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return bfail(b, SL_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(b);
        return bfail(b, SL_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    return SL_OK;
}

}  // namespace

extern "C" {

int sl_batch_create(int device, const sl_scene *scenes, const int *widths, const int *heights, const int *samples,
                    int n_scenes, int srgb, sl_batch **out) {
    if (!out || n_scenes <= 0 || !scenes || !widths || !heights || !samples) return SL_E_ARG;
    *out = nullptr;
    for (int i = 0; i < n_scenes; ++i) {
        if (widths[i] <= 0 || heights[i] <= 0 || widths[i] > 16384 || heights[i] > 16384) return SL_E_ARG;
        if (samples[i] != 8 && samples[i] != 16 && samples[i] != 32) return SL_E_ARG;
    }
    sl_batch *b = new (std::nothrow) sl_batch;
    if (!b) return SL_E_OOM;
    b->device = device < 0 ? 0 : device;
    b->n = n_scenes;
    b->scene_grp.assign(n_scenes, -1);
    b->scene_w.assign(widths, widths + n_scenes);
    b->scene_h.assign(heights, heights + n_scenes);
    b->scene_off.assign(n_scenes, 0);
    int rc = SL_OK;
    const int spps[3] = {8, 16, 32};
    for (int gi = 0; gi < 3 && !rc; ++gi) {
        std::vector<int> ids;
        for (int i = 0; i < n_scenes; ++i)
            if (samples[i] == spps[gi]) ids.push_back(i);
        if (ids.empty()) continue;
        sl_handle *h = nullptr;
        rc = sl_create(b->device, &h);
        if (rc) break;
        b->grp.push_back(h);
        b->grp_spp.push_back(spps[gi]);
        const int g = (int)b->grp.size() - 1;
        h->spp = spps[gi];
        h->ms = 0;
        h->srgb = srgb ? 1 : 0;
        // concatenate the group's scenes: paths, curves and vertices renumbered
        uint64_t P = 0, C = 0, V = 0;
        for (int i : ids) { P += scenes[i].n_paths; C += scenes[i].n_curves; V += scenes[i].n_vertices; }
        if (P >= 0xFFFFFFFFull || C >= 0xFFFFFFFFull || V >= 0xFFFFFFFFull) { rc = SL_E_ARG; break; }
        std::vector<float> vtx(2 * V), cw(C, -1.f);
        std::vector<uint32_t> cvpos(C), cpath(C), color(P);
        std::vector<uint8_t> ctype(C), rule(P);
        std::vector<I4> pc(P);
        std::vector<I4> st(ids.size());
        std::vector<uint64_t> simg(ids.size());
        std::vector<uint32_t> sfirst(ids.size());
        uint64_t p0 = 0, c0 = 0, v0 = 0, row0 = 0, tile0 = 0, img0 = 0;
        int Wmax = 1, Hmax = 1;
        for (size_t k = 0; k < ids.size(); ++k) {
            const sl_scene &sc = scenes[ids[k]];
            const int W = widths[ids[k]], H = heights[ids[k]];
            if (sc.n_paths && (!sc.vertex || !sc.curve_vertex_pos || !sc.curve_type || !sc.curve_path_id ||
                               !sc.fill_rule || !sc.fill_color)) { rc = SL_E_ARG; break; }
            if (sc.gradient_table_size >= 12) {
                bool grad = false;
                for (uint32_t p = 0; p < sc.n_paths; ++p)
                    grad |= !(sc.fill_color[p] & 0xFF000000u) && sc.fill_color[p] != 0u;
                if (grad) {
                    b->err = "gradient paints are not supported in batches (scene " + std::to_string(ids[k]) + ")";
                    rc = SL_E_UNSUPPORTED;
                    break;
                }
            }
            std::memcpy(&vtx[2 * v0], sc.vertex, (size_t)sc.n_vertices * 8);
            for (uint32_t c = 0; c < sc.n_curves; ++c) {
                cvpos[c0 + c] = sc.curve_vertex_pos[c] + (uint32_t)v0;
                cpath[c0 + c] = sc.curve_path_id[c] + (uint32_t)p0;
                ctype[c0 + c] = sc.curve_type[c];
                if (sc.curve_arc_w) cw[c0 + c] = sc.curve_arc_w[c];
            }
            const int rows = (H + 1) / 2, tiles = rows * ((W + 63) / 64);
            for (uint32_t p = 0; p < sc.n_paths; ++p) {
                color[p0 + p] = sc.fill_color[p];
                rule[p0 + p] = sc.fill_rule[p];
                pc[p0 + p] = {W, H, (int)row0, (int)k};
            }
            st[k] = {W, H, (int)tile0, (int)row0};
            simg[k] = img0;
            sfirst[k] = (uint32_t)p0;
            b->scene_grp[ids[k]] = g;
            b->scene_off[ids[k]] = img0;
            Wmax = std::max(Wmax, W);
            Hmax = std::max(Hmax, H);
            p0 += sc.n_paths; c0 += sc.n_curves; v0 += sc.n_vertices;
            row0 += rows; tile0 += tiles;
            img0 += ((uint64_t)W * H * 4 + 255) & ~(uint64_t)255;
        }
        if (rc) break;
        if (row0 >= 0x7FFFFFFFull || tile0 >= 0x7FFFFFFFull) { rc = SL_E_ARG; break; }
        sl_scene all{(uint32_t)P, (uint32_t)C, (uint32_t)V, vtx.data(), cvpos.data(), ctype.data(), cpath.data(),
                     cw.data(), rule.data(), color.data()};
        h->W = Wmax;
        h->H = Hmax;
        if ((rc = sl_add_vg(h, &all))) { b->err = h->err; break; }
        h->nscenes = (int)ids.size();
        h->batch_rows = (int64_t)row0;
        h->batch_tiles = (int64_t)tile0;
        h->batch_image_bytes = (int64_t)img0;
        if ((rc = upload(h, h->pcanvas, pc.data(), pc.size() * 16)) ||
            (rc = upload(h, h->stile, st.data(), st.size() * 16)) ||
            (rc = upload(h, h->simg, simg.data(), simg.size() * 8)) ||
            (rc = upload(h, h->sfirst, sfirst.data(), sfirst.size() * 4)) ||
            (rc = ensure(h, h->scarry, ids.size() * 4))) { b->err = h->err; break; }
    }
    if (!rc && hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking) != hipSuccess) rc = SL_E_HIP;
    if (!rc && hipEventCreateWithFlags(&b->done_ev, hipEventDisableTiming) != hipSuccess) rc = SL_E_HIP;
    if (!rc && hipEventCreateWithFlags(&b->switch_ev, hipEventDisableTiming) != hipSuccess) rc = SL_E_HIP;
    if (!rc) {
        // one eager frame sizes every buffer; the graph then replays it without host syncs
        apply_shape(b);
        rc = enqueue_all(b, b->cap_stream, false);
        if (!rc && hipStreamSynchronize(b->cap_stream) != hipSuccess) rc = SL_E_HIP;
        if (!rc) rc = capture(b);
    }
    if (rc) {
        sl_batch_destroy(b);
        return rc;
    }
    *out = b;
    return SL_OK;
}

int sl_batch_destroy(sl_batch *b) {
    if (!b) return SL_OK;
    (void)hipSetDevice(b->device);
    (void)hipDeviceSynchronize();
    drop_graph(b);
    for (sl_handle *h : b->grp) sl_destroy(h);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    if (b->done_ev) (void)hipEventDestroy(b->done_ev);
    if (b->switch_ev) (void)hipEventDestroy(b->switch_ev);
    delete b;
    return SL_OK;
}

int sl_batch_render(sl_batch *b, void *stream) {
    if (!b || !b->exec) return SL_E_ARG;
    if (hipSetDevice(b->device) != hipSuccess) return bfail(b, SL_E_HIP, "hipSetDevice failed");
    int rc = batch_order(b, (hipStream_t)stream);
    if (rc) return rc;
    const hipError_t e = hipGraphLaunch(b->exec, (hipStream_t)stream);
    if (e != hipSuccess) return bfail(b, SL_E_HIP, "hipGraphLaunch: %s", hipGetErrorString(e));
    return batch_issued(b, (hipStream_t)stream);
}

int sl_batch_render_eager(sl_batch *b, void *stream) {
    if (!b) return SL_E_ARG;
    if (hipSetDevice(b->device) != hipSuccess) return bfail(b, SL_E_HIP, "hipSetDevice failed");
    int rc = batch_order(b, (hipStream_t)stream);
    if (!rc) rc = enqueue_all(b, (hipStream_t)stream, true);
    if (!rc) rc = batch_issued(b, (hipStream_t)stream);
    return rc;
}

int sl_batch_set_shape(sl_batch *b, const int shape[4]) {
    if (!b || !shape) return SL_E_ARG;
    if (shape[0] < 0 || shape[1] < 0 || (shape[2] != 4 && shape[2] != 8) ||
        (shape[3] != 64 && shape[3] != 128 && shape[3] != 256))
        return bfail(b, SL_E_ARG, "launch shape out of range");
    int rc = batch_settle(b);
    if (rc) return rc;
    for (int i = 0; i < 4; ++i) b->shape[i] = shape[i];
    apply_shape(b);
    return capture(b);
}

int sl_batch_get_shape(sl_batch *b, int shape[4]) {
    if (!b || !shape) return SL_E_ARG;
    for (int i = 0; i < 4; ++i) shape[i] = b->shape[i];
    return SL_OK;
}

int sl_batch_tune(sl_batch *b, int reps, double *best_ms) {
    if (!b) return SL_E_ARG;
    if (reps <= 0) reps = 3;
    if (hipSetDevice(b->device) != hipSuccess) return bfail(b, SL_E_HIP, "hipSetDevice failed");
    if (int rc0 = batch_settle(b)) return rc0;
    // candidate launch shapes: per-bucket resolve grid, size-class grid, compositor
    // bucket group, fragment block size
    const int rg[] = {4096, 16384, 65536};
    const int cg[] = {1024, 4096};
    const int gg[] = {4, 8};
    const int fbk[] = {128, 256};
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return bfail(b, SL_E_HIP, "hipEventCreate failed");
    double best = 1e30;
    int best_shape[4] = {b->shape[0], b->shape[1], b->shape[2], b->shape[3]};
    int rc = SL_OK;
    for (int a : rg)
        for (int c : cg)
            for (int g : gg)
                for (int f : fbk) {
                    if (rc) break;
                    b->shape[0] = a; b->shape[1] = c; b->shape[2] = g; b->shape[3] = f;
                    apply_shape(b);
                    rc = enqueue_all(b, b->cap_stream, true);  // warm-up
                    if (rc) break;
                    (void)hipEventRecord(e0, b->cap_stream);
                    for (int r = 0; r < reps && !rc; ++r) rc = enqueue_all(b, b->cap_stream, true);
                    (void)hipEventRecord(e1, b->cap_stream);
                    if (hipEventSynchronize(e1) != hipSuccess) rc = bfail(b, SL_E_HIP, "tuning run failed");
                    float ms = 0.f;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    if (!rc && ms / reps < best) {
                        best = ms / reps;
                        best_shape[0] = a; best_shape[1] = c; best_shape[2] = g; best_shape[3] = f;
                    }
                }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (int i = 0; i < 4; ++i) b->shape[i] = best_shape[i];
    apply_shape(b);
    if (rc) return rc;
    if (best_ms) *best_ms = best;
    return capture(b);
}

int sl_batch_element_number(sl_batch *b, sl_counts *out) {
    if (!b || !out) return SL_E_ARG;
    sl_counts t{};
    for (sl_handle *h : b->grp) {
        t.path += h->counts.path; t.curve += h->counts.curve; t.vertex += h->counts.vertex;
        t.curve_fragment += h->counts.curve_fragment; t.merged_fragment += h->counts.merged_fragment;
        t.span += h->counts.span;
    }
    *out = t;
    return SL_OK;
}

int sl_batch_output_device_ptr(sl_batch *b, int scene, void **dev_ptr, int *pitch_bytes) {
    if (!b || scene < 0 || scene >= b->n || !dev_ptr || !pitch_bytes) return SL_E_ARG;
    sl_handle *h = b->grp[b->scene_grp[scene]];
    *dev_ptr = (uint8_t *)h->A.b_image.p + b->scene_off[scene];
    *pitch_bytes = 4 * b->scene_w[scene];
    return SL_OK;
}

int sl_batch_read_output(sl_batch *b, int scene, uint8_t *host_rgba, int flags) {
    if (!b || scene < 0 || scene >= b->n || !host_rgba) return SL_E_ARG;
    if (hipSetDevice(b->device) != hipSuccess) return bfail(b, SL_E_HIP, "hipSetDevice failed");
    if (hipDeviceSynchronize() != hipSuccess) return bfail(b, SL_E_HIP, "render failed");
    b->pending = false;
    sl_handle *h = b->grp[b->scene_grp[scene]];
    const int W = b->scene_w[scene], H = b->scene_h[scene];
    const size_t bytes = (size_t)W * H * 4;
    const hipError_t e = hipMemcpy(host_rgba, (uint8_t *)h->A.b_image.p + b->scene_off[scene], bytes,
                                   hipMemcpyDeviceToHost);
    if (e != hipSuccess) return bfail(b, SL_E_HIP, "hipMemcpy: %s", hipGetErrorString(e));
    if (flags & 1) flip_rows_alpha(host_rgba, W, H);
    return SL_OK;
}

const char *sl_batch_last_error(sl_batch *b) { return b ? b->err.c_str() : "null batch"; }

}  // extern "C"
