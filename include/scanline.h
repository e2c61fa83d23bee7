/*
 * scanline.h -- C ABI of the MI355X (gfx950) comb-scanline path renderer.
 *
 * This is the drop-in boundary for the reference's rasterizer interface
 * RasterizerBase::VGRasterizer (gpu-scanline/src/rasterizer/shared/ras_base.h:133-378)
 * as constructed by createRasterizer(PM_Cut_Mask_Comb_Scanline)
 * (gpu-scanline/src/rasterizer/shared/ras_factory.h:18-32).  Every entry point below
 * names the reference member it replaces.  Signatures use plain C types only: no
 * exceptions cross this boundary, every call returns an SL_* status code, and the
 * text of the last failure is available from sl_last_error().
 *
 * Threading: a handle is single-threaded (like the reference's class, which keeps
 * static state in rasterizeImpl, ras_cut_mask_comb_scanline.cu:882-884); separate
 * handles are independent.  All GPU work of a handle runs on the HIP stream passed to
 * sl_rasterize (NULL = the default stream).  sl_add_vg validates and packs large scenes
 * on up to 16 host worker threads of its own, copies on the default stream and returns
 * once the copies are done.
 */
#ifndef SCANLINE_MI355X_H
#define SCANLINE_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (replace the reference's std::runtime_error throws:
 *      ras_factory.h:28, cuda_cached_allocator.h:80-81, ras_qm_mask.cu:173) ---- */
#define SL_OK             0
#define SL_E_ARG          1  /* invalid argument or malformed scene                  */
#define SL_E_OOM          2  /* device or host allocation failed                      */
#define SL_E_HIP          3  /* a HIP runtime call or kernel launch failed            */
#define SL_E_UNSUPPORTED  4  /* valid request the renderer does not implement         */
#define SL_E_STATE        5  /* call made in the wrong state (e.g. render w/o scene)  */
#define SL_E_NODEVICE     6  /* no usable GPU (the renderer has no CPU fallback)       */

/* ---- curve type codes: VGCurveType (bezier_curve_type.h:6-18); vertex count = code & 7 */
#define SL_CURVE_LINEAR    0x02
#define SL_CURVE_QUADRATIC 0x03
#define SL_CURVE_CUBIC     0x04
#define SL_CURVE_RATIONAL  0x13

/* ---- fill rules: VGFillRule (vg_container.h:43-46) ---- */
#define SL_FILL_NONZERO  0
#define SL_FILL_EVEN_ODD 1

typedef struct sl_handle sl_handle;

/*
 * Scene in the reference's GPU structure-of-arrays form (VGInputCurveDataPack /
 * VGInputPathDataPack, ras_base.h:34-112, filled by RasterizerBase::addVg,
 * ras_base.cpp:70-382).  All pointers are host pointers.  Vertices are stored per
 * curve (a curve of type T owns (T & 7) consecutive vertices starting at
 * curve_vertex_pos), exactly as VGContainer::addLinear/.../addRational lay them out
 * (vg_container.cpp:97-125).  Curves must be grouped by path with non-decreasing
 * curve_path_id; every contour must be closed (the reference's loaders close open
 * contours: rvg.cpp:383-396, svg_load.cpp:784-795).
 *
 * fill_color is the path's fill word (fill_info, ras_base.cpp:144-166): an RGBA8 colour
 * (alpha 0 draws nothing), or, when a gradient table is given, a value g in
 * [1, 0x01000000] naming the gradient in table slot g - 1 (the reference's packing gives a
 * gradient-filled path p the word p + 1).  The gradient textures are the VGContainer's
 * (vg_container.h:344-345): gradient_ramp holds RGBA8 texels in rows of 1024, and
 * gradient_table 12 words per slot {paint transform tr[6] (float), ramp position (int),
 * ramp scale (float), focal point x, y, z (float), unused}, as RVG::load builds them
 * (rvg.cpp:490-758).  Sizes count 32-bit words; both may be 0 / NULL.
 */
typedef struct sl_scene {
    uint32_t        n_paths;
    uint32_t        n_curves;
    uint32_t        n_vertices;
    const float    *vertex;            /* 2*n_vertices floats: x0,y0,x1,y1,...        */
    const uint32_t *curve_vertex_pos;  /* [n_curves] first vertex of each curve       */
    const uint8_t  *curve_type;        /* [n_curves] SL_CURVE_*                        */
    const uint32_t *curve_path_id;     /* [n_curves] owning path, non-decreasing      */
    const float    *curve_arc_w;       /* [n_curves] rational weight w1; may be NULL  */
    const uint8_t  *fill_rule;         /* [n_paths]  SL_FILL_*                         */
    const uint32_t *fill_color;        /* [n_paths]  fill word: RGBA8 r | g<<8 | b<<16 | a<<24,
                                          or gradient slot + 1                        */
    const int32_t  *gradient_ramp;     /* [gradient_ramp_size] RGBA8 texels; may be NULL */
    uint32_t        gradient_ramp_size;
    const int32_t  *gradient_table;    /* [gradient_table_size] 12 words per slot; may be NULL */
    uint32_t        gradient_table_size;
} sl_scene;

/* ElementNumber (ras_base.h:124-131), widened to 64 bits. */
typedef struct sl_counts {
    int64_t path;
    int64_t curve;
    int64_t vertex;
    int64_t curve_fragment;   /* raw fragments F (one per intersection entry)       */
    int64_t merged_fragment;  /* merged (path, 2x2 cell) fragments M                 */
    int64_t span;             /* interior spans S                                    */
} sl_counts;

/* ---- lifecycle -------------------------------------------------------------------- */

/* createRasterizer(PM_Cut_Mask_Comb_Scanline) + VGRasterizer::init()
 * (ras_factory.h:18-32; ras_cut_mask_comb_scanline.cpp:31-45).  Selects device 0
 * unless `device` >= 0.  Fails with SL_E_NODEVICE when no GPU is present. */
int sl_create(int device, sl_handle **out);

/* VGRasterizer::uninit() + delete. */
int sl_destroy(sl_handle *h);

/* ---- configuration (ras_base.h:155-205) ------------------------------------------ */

/* setSamples (ras_base.h:199). 8 and 32 follow the reference; 16, which has no
 * reference mask layout, keeps one 32-bit word per pixel column (bit y*16 + s) and a
 * bit-reversal sample set like the 8-sample one. */
int sl_set_samples(sl_handle *h, int samples);

/* setMultisampleOutput (ras_base.h:200): 0 = coverage-alpha compositing
 * (output_{8,32}.frag.glsl), 1 = per-sample compositing + resolve
 * (ms_output_{8,32}.frag.glsl, ras_base.cpp:1148-1163). */
int sl_set_multisample_output(sl_handle *h, int enable);

/* enableSRGBCorrection (ras_base.h:184): sRGB-decoded paint blended in linear space
 * into an sRGB8 framebuffer (output_32.vert.glsl:44-51, .cu:1384-1389). */
int sl_enable_srgb_correction(sl_handle *h, int enable);

/* setInputTransform (ras_base.cpp:576-583). `m` is a column-major 4x4 matrix
 * (glm layout); row i of the transform is (m[i], m[4+i], m[8+i], m[12+i]). */
int sl_set_input_transform(sl_handle *h, const float m[16]);

/* resizeOutputBuffer (ras_base.cpp:393-402). */
int sl_resize_output_buffer(sl_handle *h, int width, int height);

/* viewport / resizeViewport (ras_base.h:147-148, ras_base.cpp:385-394): the rectangle of
 * the final blit of the output image into the window (ras_base.cpp:1170-1176: source
 * (0, 0, w, h) to destination (x, y, x + w, y + h), GL_NEAREST, no scaling).  A size of
 * 0 (the default) means the image's size. */
int sl_viewport(sl_handle *h, int x, int y, int width, int height);
int sl_resize_viewport(sl_handle *h, int width, int height);

/* The window blit itself, into a caller's device RGBA8 buffer of dst_width x dst_height
 * pixels with rows dst_pitch bytes apart: the viewport rectangle of the last image copied
 * one pixel for one to (x, y), clipped to the image and to the buffer; rows in the same
 * y-up order as the image.  Queued on `stream` after the last frame. */
int sl_present(sl_handle *h, void *dev_dst, int dst_pitch, int dst_width, int dst_height, void *stream);

/* ---- scene ------------------------------------------------------------------------ */

/* addVg (ras_cut_mask_comb_scanline.cpp:224-229 -> ras_base.cpp:70-382) followed by
 * unifyCurveVertexOrder (ras_base.cpp:523-569).  The scene is validated, copied to the
 * device and oriented; the caller's arrays are not retained. */
int sl_add_vg(sl_handle *h, const sl_scene *scene);

/* clear (ras_base.cpp:572-573): drop the uploaded scene. */
int sl_clear(sl_handle *h);

/* ---- scenes generated on the device ------------------------------------------------
 * The reference's animation mode writes each frame's scene on the GPU straight into the
 * rasterizer's input packs (vg_animation -> chord_animation, animation.cu:2435-2786,
 * called from the frame loop at ras_cut_mask_comb_scanline.cu:903-921).  Here a caller
 * loads the keyframes once and asks for the scene at a time; the next sl_rasterize renders
 * it.  Keyframes of one chord diagram, as RefugeeAnimation::genFrame lays them out
 * (animation.cu:1200-1330):
 *   node_t0, node_t1, node_size  [n_keys * n_nodes]  angular extent (radians, direction
 *                                 (sin t, cos t) from the centre) and size at each keyframe
 *   node_color                   [n_nodes] RGBA8 fill word of each node (a link takes its
 *                                 larger end's colour with the alpha's top bit cleared)
 *   link_offset                  [n_keys] the links drawn while moving from keyframe k - 1
 *                                 to k are [link_offset[k - 1], link_offset[k]); entry 0 is 0
 *   link_node                    [2 * n_links] the two nodes of each link
 *   link_from, link_to           [4 * n_links] {pos0, pos1, size0, size1} of each link's
 *                                 two ends inside their nodes, at keyframe k - 1 and k
 *   center_x, center_y           the diagram's centre in scene coordinates
 * Every node becomes a ring sector between radii 320 and 350, every link a ribbon inside
 * radius 320; each path is 6 curves in 24 vertex slots, nodes before links, diagram after
 * diagram.  The country and year text of the reference's demo is not generated. */
typedef struct sl_chord {
    int32_t         n_keys;           /* >= 2 */
    int32_t         n_nodes;
    const float    *node_t0, *node_t1, *node_size;
    const uint32_t *node_color;
    const uint32_t *link_offset;
    const uint32_t *link_node;
    const float    *link_from, *link_to;
    float           center_x, center_y;
} sl_chord;

/* load_refugee_animation's result (animation.cu:1480-1706) handed to the device: copies
 * up to 8 diagrams' keyframes to device memory and drops the handle's scene.  The caller's
 * arrays are not retained. */
int sl_chord_load(sl_handle *h, const sl_chord *diagrams, int n_diagrams);

/* chord_animation at time `time_ms` >= 0 (animation.cu:2435-2760): keyframes k - 1 and
 * k = time_ms / 6000 + 1 eased by GET_ANIMATION_T (animation.cu:1710-1718); after the last
 * keyframe the diagram holds it.  The scene is written on `stream` into the handle's scene
 * arrays, already oriented as sl_add_vg leaves an upload, and replaces the scene; the next
 * sl_rasterize is a full frame.  It is queued after every frame issued before it; the call
 * waits on the host only when the scene outgrows its arrays. */
int sl_chord_frame(sl_handle *h, int time_ms, void *stream);

/* ---- rendering ---------------------------------------------------------------------- */

/* rasterize() (ras_base.cpp:1131-1195 -> rasterizeImpl<2>, .cu:879-1453): renders the
 * scene into the handle's device RGBA8 image (row 0 = y in [0,1), the reference's
 * GL framebuffer order).  `stream` is a hipStream_t or NULL.
 * The first frame after any change (scene, transform, output size, sample count, output
 * mode, colour encoding, debug dumps) reads its three sizes (fragments, chains, buckets)
 * back in one round trip early in the frame, sizes its buffers and returns with the rest
 * queued.  Later frames of the unchanged scene reuse those sizes: every kernel runs
 * again, queued on `stream` with no host synchronisation, and the call returns at once.
 * Every frame hands its element counts and error flags back without waiting; they are
 * read when the frame is settled: sl_read_output and sl_element_number wait for the frame
 * and return its error (SL_E_HIP for an internal inconsistency flagged on the device),
 * users of sl_output_device_ptr order their work after it on `stream`.
 * Parity dumps keep every frame synchronous.
 * A queued frame shares the handle's buffers with the next one: sl_add_vg, sl_debug_dump
 * and a full frame on the same stream wait for it on the host.  A frame issued on another
 * stream comes after everything queued on the previous frame's stream up to the call (that
 * frame and the caller's reads of its image queued after it): on the device for a queued
 * frame, on the host for a full one.  A frame that fails is still waited for like a queued
 * one, since some of its kernels may have been issued. */
int sl_rasterize(sl_handle *h, void *stream);

/* The frame of sl_rasterize, allocating nothing: no counterpart in the reference.  It is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The handle's scene and settings are
 * rendered into the caller's device image `dev_rgba` (height rows of `pitch_bytes` >=
 * width * 4 bytes) on `stream`; every intermediate array lives in one static device
 * scratch of 256 MiB per GPU that the library holds until the process ends (its address
 * is looked up at sl_create).  A scene needing more fails with SL_E_UNSUPPORTED and a
 * message giving the size; render it with sl_rasterize.  The mask and colour tables are
 * built by sl_add_vg, sl_set_samples and sl_enable_srgb_correction (which may allocate),
 * and debug dumps must be off (SL_E_STATE otherwise).  Like a full sl_rasterize frame it
 * reads its counts back and returns when the frame is done.  Every call on one GPU shares
 * that GPU's scratch: calls that overlap on one GPU (several host threads, several
 * handles) race and must be serialised by the caller. */
int sl_rasterize_noalloc(sl_handle *h, void *dev_rgba, int pitch_bytes, void *stream);

/* Device pointer and row pitch (bytes) of the handle's RGBA8 output image.  Fails with
 * SL_E_STATE when the output size changed since the last frame (the image then still has
 * the old size); the same holds for sl_read_output and the "image" dump. */
int sl_output_device_ptr(sl_handle *h, void **dev_ptr, int *pitch_bytes);

/* Copy the last image to host memory (width*height*4 bytes of the size it was rendered
 * at).  flags bit 0: flip rows and force alpha to 255 as saveOutputToFile does
 * (ras_base.cpp:1249-1262). */
int sl_read_output(sl_handle *h, uint8_t *host_rgba, int flags);

/* elementNumber() (ras_base.h:207). */
int sl_element_number(sl_handle *h, sl_counts *out);

/* enableStepTiming / stepTiming (ras_base.h:193-195): per-stage device times in
 * milliseconds of the last sl_rasterize, measured with HIP events. */
int sl_enable_step_timing(sl_handle *h, int enable);
int sl_step_timing(sl_handle *h, double *ms, int capacity, int *n_stages);
const char *sl_step_name(int stage);

/* countPixel / pixelCount (ras_base.h:181-182; .cu:1269-1288): sum over draw items of
 * 2 * item width. */
int sl_pixel_count(sl_handle *h, int64_t *out);

/* Text of the last error on this handle (never NULL). */
const char *sl_last_error(sl_handle *h);

/* ---- parity dumps (debug data of the reference's pipeline stages) --------------------
 * After sl_debug_enable(h, 1), sl_rasterize also materialises the intermediate arrays
 * in the reference's own layouts, and sl_debug_dump copies one of them to host memory:
 *   "ts"            int2[F]   intersections {curve id, t bits | flags}  (ras_cut.h:837-1241)
 *   "frag_yx"       int32[F]  raw fragment cell key, in ts order          (.cu:255-262)
 *   "frag_pid"      int32[F]  path id | winding change << 30              (.cu:414)
 *   "frag_wn"       int8[F]   scan winding number                         (.cu:264-278)
 *   "frag_mask"     int32[4F] coverage mask, 4 words per fragment         (.cu:330-371)
 *   "merged_mask"   int32[4M] resolved per-sample masks, draw-list order  (.cu:554-786)
 *   "draw_list"     int4[M+S] {y<<16|x, width, fill, mask index+1}        (.cu:789-868)
 *   "image"         u8[W*H*4] the output image
 * `nbytes` receives the size; a NULL or too-small destination only queries it. */
int sl_debug_enable(sl_handle *h, int enable);
int sl_debug_dump(sl_handle *h, const char *name, void *host_dst, size_t capacity,
                  size_t *nbytes);

/* Colour requantisation self-check of the 8-bit framebuffer emulation (the blend
 * re-quantisation of an RGBA8 / SRGB8_ALPHA8 colour attachment, ras_base.cpp:1044-1060).
 * channel 0 = colour (sRGB when sRGB correction is enabled, else UNORM), 1 = alpha (UNORM).
 * sl_debug_requantize runs the device encoder on n host values: codes[i] and the decoded
 * value of that code.  sl_debug_requantize_sweep compares the device encoder with the
 * threshold tables on every float bit pattern in [bits_lo, bits_hi) and returns the
 * number of disagreements.  Neither has a counterpart in the reference (its blending is
 * the GL driver's); both exist to prove the emulation exact. */
int sl_debug_requantize(sl_handle *h, int channel, const float *values, int64_t n,
                        uint8_t *codes, float *decoded);
int sl_debug_requantize_sweep(sl_handle *h, int channel, uint32_t bits_lo, uint32_t bits_hi,
                              uint64_t *mismatches);

/* VALU issue-rate probe (no reference counterpart; calibrates the instruction ceiling the
 * benchmark reports against): one timed launch of `blocks` workgroups of 256 threads, each
 * running `iters` rounds of 512 independent v_fma_f32 per wave.  Returns the launch time
 * and the wave VALU instructions it issued (blocks * 4 * iters * 512). */
int sl_debug_valu_probe(sl_handle *h, int blocks, int iters, double *ms, uint64_t *wave_insts);

/* ---- batched scenes -------------------------------------------------------------------
 * Many independent scenes of uneven size rendered by one launch sequence over all of
 * them, captured once into a HIP graph and replayed.  No reference counterpart: the
 * reference renders one scene per rasterize() call (ras_base.cpp:1131-1195); a batch
 * gives each scene exactly the image a separate handle would (coverage output).
 * sl_batch_create uploads and orients every scene (as sl_add_vg), renders one eager
 * frame to size the buffers, and captures the graph.  sl_batch_render replays it on
 * `stream`; sl_batch_render_eager issues the same kernels without the graph.
 * sl_batch_tune times candidate launch shapes {per-bucket resolve grid, size-class
 * grid, compositor bucket group, fragment block size} on this batch, keeps the fastest
 * and recaptures.  Images are read per scene.  A frame issued on another stream than the
 * previous one waits on the device for everything queued on that stream up to the call
 * (the previous frame and the caller's reads of its images); tuning, a shape change and
 * the read-back wait for it on the host. */
typedef struct sl_batch sl_batch;
int sl_batch_create(int device, const sl_scene *scenes, const int *widths, const int *heights,
                    const int *samples, int n_scenes, int srgb, sl_batch **out);
int sl_batch_destroy(sl_batch *b);
int sl_batch_render(sl_batch *b, void *stream);
int sl_batch_render_eager(sl_batch *b, void *stream);
int sl_batch_tune(sl_batch *b, int reps, double *best_ms);
int sl_batch_set_shape(sl_batch *b, const int shape[4]);
int sl_batch_get_shape(sl_batch *b, int shape[4]);
int sl_batch_element_number(sl_batch *b, sl_counts *out);
int sl_batch_output_device_ptr(sl_batch *b, int scene, void **dev_ptr, int *pitch_bytes);
int sl_batch_read_output(sl_batch *b, int scene, uint8_t *host_rgba, int flags);
const char *sl_batch_last_error(sl_batch *b);

/* ---- differentiable mode -----------------------------------------------------------------
 * No reference counterpart.  Coverage is a box filter of `sample_rows` rows per pixel,
 * each integrated exactly along x (continuous in the control points); compositing is the
 * renderer's src-over onto white in path order, unquantised, into a float RGBA image.
 * Scenes of line / quadratic / cubic segments (closed contours), vertices stored per
 * curve as in sl_scene.  `points` (float [V][2]), `colors` (float [P][4], RGBA in [0,1]),
 * `image` (float [H][W][4]) and all gradients are DEVICE pointers; work runs on `stream`.
 * sl_diff_forward keeps what the adjoint needs when `keep_for_backward` is set (the
 * points and colours must stay unchanged until the backward).  sl_diff_backward writes
 * d loss / d points and d loss / d colors given d loss / d image: the compositing chain
 * reversed per pixel, and boundary terms at every sample-row crossing with a fill change.
 * One pending backward per handle.  A call on another stream than the previous call waits
 * on the host for that call to finish (the calls share the handle's buffers). */
typedef struct sl_diff sl_diff;
int sl_diff_create(int device, int width, int height, int sample_rows, uint32_t n_paths,
                   uint32_t n_curves, uint32_t n_vertices, const uint32_t *curve_vertex_pos,
                   const uint8_t *curve_type, const uint32_t *curve_path_id,
                   const uint8_t *fill_rule, sl_diff **out);
int sl_diff_destroy(sl_diff *d);
int sl_diff_forward(sl_diff *d, const float *points, const float *colors, float *image,
                    int keep_for_backward, void *stream);
int sl_diff_backward(sl_diff *d, const float *grad_image, float *grad_points,
                     float *grad_colors, void *stream);
int sl_diff_crossings(sl_diff *d, int64_t *n);
const char *sl_diff_last_error(sl_diff *d);

/* ---- seeded synthetic scenes ------------------------------------------------------------
 * Host-only generators for the benchmark workloads (splitmix64, one substream per
 * path).  kind: 1 = closed chains of cubics with control points uniform in the canvas;
 * 2 = mixed line/quad/cubic paths with log-uniform bboxes, star polygons, alternating
 * fill rules; 3 = glyph-like contours of quadratics with reversed holes.
 * `min_segs`/`max_segs` <= 0 select the kind's defaults.  The scene stays owned by the
 * generator object until sl_scene_free. */
typedef struct sl_scene_obj sl_scene_obj;
int sl_scene_generate(int kind, int width, int height, int n_paths, uint64_t seed,
                      int min_segs, int max_segs, sl_scene_obj **out);
int sl_scene_view(const sl_scene_obj *obj, sl_scene *out);
int sl_scene_free(sl_scene_obj *obj);

/* Library version string. */
const char *sl_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SCANLINE_MI355X_H */
