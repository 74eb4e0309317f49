/*
 * pyr3hip — C ABI of the MI355X (gfx950) render path.
 *
 * Drop-in for the reference's stage interface (src/renderer.ts:163-212,
 * createRenderer / reset / iterate / present / readHistogram) and its pass
 * drivers (src/chaos.ts ChaosPass, src/density.ts DensityPass,
 * src/visualize.ts VisualizePass). The host packs the same byte blobs the
 * reference uploads to WebGPU (464-byte xform records, the u32 xform-pick
 * table, the 128-byte chaos uniform block, the 32-byte DE block, the 96-byte
 * tonemap block, the 256 x RGBA f32 palette) and hands them over by pointer.
 *
 * Device buffers are owned by the renderer handle. Every call is ordered on
 * the handle's HIP stream and returns immediately unless it reads back to
 * host memory. All functions return 0 on success or a negative P3_E* code;
 * p3_last_error() gives the message. A handle is not re-entrant: one host
 * thread at a time.
 */
#ifndef PYR3HIP_H
#define PYR3HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define P3_OK 0
#define P3_E_ARG (-1)      /* invalid argument (null, size mismatch, out of range) */
#define P3_E_HIP (-2)      /* HIP runtime failure */
#define P3_E_STATE (-3)    /* call out of order (e.g. present before reset) */

#define P3_XFORM_FLOATS 116
#define P3_GRAIN 16384
#define P3_MAX_XFORMS 128
#define P3_CHAOS_UNIFORM_WORDS 32
#define P3_DE_UNIFORM_WORDS 8
#define P3_VIZ_UNIFORM_WORDS 24
#define P3_TRACE_FLOATS 16
#define P3_NOALLOC_SCRATCH_BYTES (96u * 1024u * 1024u)  /* p3_iterate_noalloc's device array */

typedef struct p3_renderer p3_renderer;

/* createRenderer (src/renderer.ts:214, buildPipelines :332-376).
 * width/height: output pixels; oversample >= 1; taps: the spatial-filter
 * kernel (src/spatial-filter.ts:217-256), n_taps in [3, 17]. Allocates the
 * super-res histogram (W*os x H*os x 4 u32), the filtered grid (x 4 f32),
 * the RGBA8 image, and builds the DE kernel-norm LUT (src/density.ts:78-99). */
int p3_create(int width, int height, int oversample, const float* taps, int n_taps,
              p3_renderer** out);

/* Renderer.resize (src/renderer.ts:277-292): rebuild the handle at a new
 * size. Allocations are kept and grown only when the new grid needs more;
 * work already queued on the handle's stream is not waited for (new filter
 * taps wait for it before their host copy is replaced). The histogram is
 * undefined until the next p3_reset, as after p3_create. When growing a
 * buffer fails (P3_E_HIP), the handle is left without usable buffers: every
 * later call except p3_resize, p3_dims, p3_sync, p3_last_error and
 * p3_destroy returns P3_E_STATE until a p3_resize succeeds. */
int p3_resize(p3_renderer* r, int width, int height, int oversample, const float* taps,
              int n_taps);

/* Renderer.destroy. Null is accepted. */
void p3_destroy(p3_renderer* r);

/* Renderer.reset (src/renderer.ts:222-225, chaos.ts:309-320): upload the
 * 256 x RGBA f32 palette LUT and zero the histogram (device-side fill). */
int p3_reset(p3_renderer* r, const float* palette_rgba);

/* Renderer.iterate -> ChaosPass.dispatch (src/chaos.ts:333-466): one chaos
 * dispatch accumulating into the histogram.
 *   xforms:   n_slots x 116 f32 (regular xforms, then the final xform if any)
 *   distrib:  n_rows x 16384 u32; rows 0..n_rows-2 are "previous xform = i",
 *             the last row is the no-previous fallback (row 128 in the
 *             reference's 129-row buffer)
 *   uniforms: the 32-word chaos uniform block; slot 7 (seed) and slot 12
 *             (walker count) seed walker w's ISAAC stream on the device, bit
 *             for bit as packIsaacStates (src/isaac.ts:194-229) would; slots
 *             14-29 select the colour-index capture and the splat colour mode
 *             (0 palette, 1 flow, 2 trap distance, 3 phase) with its
 *             parameters, laid out as src/chaos.ts:421-450 writes them. */
int p3_iterate(p3_renderer* r, const float* xforms, int n_slots, const uint32_t* distrib,
               int n_rows, const uint32_t* uniforms);

/* Batched launch: n_sub (1..64) dispatches of one genome shape (same
 * variation slots, post-affine presence and final xform; e.g. the sub-frames
 * of a motion-blurred frame, src/animate-render.ts:145-160) in one grid. Each
 * sub-dispatch k has its own records xforms[k], pick table distribs[k] and
 * uniform block uniforms[k] (seed, walkers, iterations, camera) and numbers
 * its walkers from 0, so the histogram equals that of n_sub p3_iterate calls.
 * A sub-dispatch of a different shape is an error (P3_E_ARG). No counterpart
 * in the reference (it issues one dispatch per sub-frame). */
int p3_iterate_multi(p3_renderer* r, int n_sub, const float* const* xforms, int n_slots,
                     const uint32_t* const* distribs, int n_rows,
                     const uint32_t* const* uniforms);

/* As p3_iterate, but walkers take explicit ISAAC states (walkers x 36 u32,
 * the reference's isaac_states buffer layout) instead of device seeding. */
int p3_iterate_states(p3_renderer* r, const float* xforms, int n_slots, const uint32_t* distrib,
                      int n_rows, const uint32_t* uniforms, const uint32_t* isaac_states);

/* Renderer.present (src/renderer.ts:241-262): DE (when de_uniforms is
 * non-null, the 8-word DEUniforms block) then the tonemap (the 24-word
 * VizUniforms block) into the renderer-owned RGBA8 image. */
int p3_present(p3_renderer* r, const uint32_t* de_uniforms, const uint32_t* viz_uniforms);

/* Readbacks (block until the stream has drained). n is the element count of
 * the destination and must match the buffer exactly. */
int p3_read_histogram(p3_renderer* r, uint32_t* out, size_t n);  /* readHistogramRgba */
/* Colour-index capture (src/chaos.ts:262-291 readIndexAndCount, idx_sum half):
 * W_s x H_s u32 sums of clamp(colour, 0, 1) x opacity x 255 per bin, written
 * by dispatches whose uniform slot 14 (capture_index) is 1; cleared by
 * p3_reset once allocated. P3_E_STATE before the first capturing dispatch. */
int p3_read_index(p3_renderer* r, uint32_t* out, size_t n);
/* Channel-curves LUT for p3_present (src/visualize.ts:170-175 uploads the
 * bakeCurves output, src/channel-curves.ts:115-133): 5 x 256 floats in [0, 1],
 * channels composite, R, G, B, luma. Read only when the tonemap uniforms' word
 * 16 (curvesActive bit mask) is non-zero; p3_present fails with P3_E_STATE if
 * that mask is set and no LUT was uploaded. lut = NULL or n = 0 clears it.
 * Words 17..20 are hslActive, hue (degrees), saturation factor and lightness
 * offset (visualize.ts:186-193, applied as visualize_f32.wgsl:240-266). */
int p3_set_curves(p3_renderer* r, const float* lut, size_t n);
/* histogramToLinearRgba (src/export-linear.ts:24-47) on the device, for the
 * scene-linear EXR export (bin/pyr3-render.ts:520-537): W x H x 4 f32, each
 * output pixel the sum of its os x os histogram cells times (k1 k2) / os^2
 * (f64 arithmetic, as the reference), alpha clamped to [0, 1]. k1, k2 from
 * the calibration of the render (src/calibration.ts:37-43). Reads the
 * histogram in place: no histogram readback. */
int p3_read_linear(p3_renderer* r, double k1, double k2, float* out, size_t n);

/* Measurement only (no counterpart in the reference): the device's rate of
 * the histogram splat's atomic requests, for the chaos kernel's roofline.
 * blocks x 256 threads each add 1 to every 16-bit field of per_thread cells
 * drawn uniformly over the handle's grid, in the splat's request form: one
 * 64-bit add per 8-byte cell of the packed accumulator (merged into the
 * histogram at its next read, exact below 2^16 adds per cell). It
 * overwrites the histogram's contents: reset before the next render. Writes
 * the atomic requests per second (HIP events on the handle's stream) and
 * the requests per splat (1). */
int p3_probe_atomics(p3_renderer* r, uint32_t blocks, uint32_t per_thread,
                     double* requests_per_s, int* requests_per_splat);
int p3_read_filtered(p3_renderer* r, float* out, size_t n);      /* DE output grid */
int p3_read_image(p3_renderer* r, uint8_t* out, size_t n);       /* W x H x 4 RGBA8 */
int p3_read_image_f32(p3_renderer* r, float* out, size_t n);     /* pre-quantisation RGBA */

/* When enabled, p3_present also stores the pre-quantisation RGBA f32 image
 * (the values the reference writes to its rgba16float target for png16/EXR,
 * src/visualize.ts:105-156); off by default. */
int p3_set_image_f32(p3_renderer* r, int enabled);

/* Upload a histogram (tests: DE/tonemap on an identical histogram). */
int p3_write_histogram(p3_renderer* r, const uint32_t* data, size_t n);

/* Trace mode (src/shaders/chaos.wgsl:8412-8429 generalised): record walkers
 * [0, walkers) for loop indices [0, iters) (fuse included), 16 f32 per entry:
 * i, pick, pa.xy, pv_pre.xy, pv.xy, bad, z, draws(u32 bits), pick draw (u32
 * bits), p.xy after jitter, splat.xy. walkers = 0 turns it off. */
int p3_set_trace(p3_renderer* r, int walkers, int iters);
int p3_read_trace(p3_renderer* r, float* out, size_t n);

/* Kernel timing with HIP events on the handle's stream: when enabled, each
 * launch is bracketed and p3_read_timings writes the last durations (ms) of
 * [merge, chaos, density, tonemap, reset, denoise] (6 floats); merge is the
 * packed accumulator's merge into the histogram, run by the first reader
 * (present, a readback or the linear export) after a dispatch. */
int p3_set_timing(p3_renderer* r, int enabled);
int p3_read_timings(p3_renderer* r, float* out6);

/* Block until every queued call on the handle has completed. */
int p3_sync(p3_renderer* r);

/* Geometry of the handle: out4 = {width, height, super_w, super_h}. */
int p3_dims(p3_renderer* r, int* out4);

/* 1 if the kernels implement catalog entry `index` (src/variations.ts:24 V
 * table), else 0. p3_iterate rejects xform records that use an entry the
 * kernels do not implement instead of rendering it as (0, 0). */
int p3_variation_supported(int index);

/* Chaos-kernel specialisation (no counterpart in the reference, whose WGSL
 * kernel dispatches variations through a runtime switch): p3_iterate compiles
 * a kernel per genome shape (variation index of every slot, post-affine
 * presence, final xform) with hipRTC on first use and caches it for the
 * process. p3_chaos_prepare compiles ahead of time without a device (for
 * warm-up outside a timed region); p3_chaos_source returns the generated
 * source (len = full length; out, when non-null, gets up to cap - 1 bytes). */
int p3_chaos_prepare(const float* xforms, int n_slots, int n_rows, int final_idx);
/* 1 when the specialised kernel of this shape is already compiled (in this
 * process or in the code-object cache), 0 when not; never compiles. No
 * counterpart in the reference (its kernel is generic). */
int p3_chaos_ready(const float* xforms, int n_slots, int n_rows, int final_idx);
int p3_chaos_source(const float* xforms, int n_slots, int n_rows, int final_idx, char* out,
                    size_t cap, size_t* len);

/* Learned denoiser (no counterpart in the reference): a 4-level U-Net
 * (pyr3_amd/denoise/net.py, ~0.49 M parameters) run in BF16 with MFMA
 * kernels on the handle's stream. Parameters are float32, per 3x3 layer
 * (e0 e1 d1 e2 d2 e3 d3 b0 u3 c3 u2 c2 u1 c1) its weight [c_out][c_in][3][3]
 * then its bias, then the 1x1 head's weight [3][16] and bias;
 * p3_denoiser_param_count() gives their number. p3_denoise reads the
 * renderer's histogram (oversample 1) and its last present, which must have
 * run without DE and with the f32 image enabled (the base image the network
 * corrects; the present's k2 scales the density input), and overwrites the
 * RGBA8 image with the denoised one. */
typedef struct p3_denoiser p3_denoiser;
size_t p3_denoiser_param_count(void);
int p3_denoiser_create(const float* params, size_t n, p3_denoiser** out);
int p3_denoiser_destroy(p3_denoiser* d);
int p3_denoise(p3_renderer* r, p3_denoiser* d);

/* p3_iterate without any device allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The
 * inputs are those of p3_iterate plus the 256 x RGBA f32 palette (host
 * memory); the output is hist, a caller-owned device buffer of
 * n_hist = u[0] x u[1] x 4 u32 (the super-res grid) to which the dispatch's
 * splats are added, saturating, on `stream` (a hipStream_t; null = the null
 * stream) before the call returns. The packed accumulator, its carry counts
 * and the uploaded blobs live in a static device array of
 * P3_NOALLOC_SCRATCH_BYTES (96 MiB) that every GPU the library runs on holds
 * until the process ends: 40 bytes per bin plus the blobs, so grids up to
 * about 2.3 M bins (1920 x 1080 at oversample 1); a larger grid is P3_E_ARG.
 * The kernel of the genome's shape must already be loaded on the device
 * (one p3_iterate of that shape on it; loading code allocates), else
 * P3_E_STATE. The colour-index capture (u[14] = 1) is P3_E_ARG. Calls share
 * the device's array, which keeps the last call's contents (each call
 * clears what it needs): calls that overlap on one GPU race. Errors are
 * read with p3_last_error(NULL). */
int p3_iterate_noalloc(const float* xforms, int n_slots, const uint32_t* distrib, int n_rows,
                       const uint32_t* uniforms, const float* palette_rgba, uint32_t* hist,
                       size_t n_hist, void* stream);

/* Batched rendering (no counterpart in the reference, whose batch loop
 * renders one genome or sub-frame per dispatch, src/animate-render.ts:144-160):
 * many output frames of any genome shapes and sizes on one captured HIP
 * graph with a few persistent chaos launches.
 *
 * p3_batch_add: one output frame, with the arguments of p3_create (size,
 * oversample, taps), p3_reset (palette), p3_iterate_multi (n_sub
 * sub-dispatches of one genome shape, e.g. the sub-frames of a motion-blurred
 * frame, each with its own records, pick table and uniform block: seed,
 * walkers, iterations, camera) and p3_present (DE uniforms or null, tonemap
 * uniforms). Returns the frame's index (>= 0) or a P3_E* code. Colour-index
 * capture and channel curves are refused (P3_E_ARG).
 *
 * p3_batch_commit(b, max_shapes, lanes): allocates every frame's buffers,
 * uploads the blobs, compiles one union chaos kernel per launch group (the
 * frames in order, at most max_shapes distinct genome shapes per group;
 * frames with more than 16 xforms group apart) and captures the batch in a
 * HIP graph: per group a histogram fill, one chaos launch over all of its
 * sub-dispatches, the packed merge, then DE and tonemap per frame; groups
 * alternate over `lanes` (1..4) streams. Frames cannot be added afterwards.
 *
 * p3_batch_launch replays the graph on the batch's stream (returns at once);
 * every replay renders the same frames, bit-identically, and each frame's
 * histogram equals that of p3_reset + p3_iterate_multi of its sub-dispatches
 * on a renderer. p3_batch_elapsed gives the last launch's device time (ms);
 * p3_batch_stats {groups, shapes, sub-dispatches, chaos blocks}; the reads
 * wait for the launch and copy one frame's RGBA8 image (w x h x 4 bytes) or
 * histogram (W_s x H_s x 4 u32). Errors: p3_batch_last_error(b). */
typedef struct p3_batch p3_batch;
int p3_batch_create(p3_batch** out);
void p3_batch_destroy(p3_batch* b);
int p3_batch_add(p3_batch* b, int width, int height, int oversample, const float* taps,
                 int n_taps, const float* palette_rgba, int n_sub, const float* const* xforms,
                 int n_slots, const uint32_t* const* distribs, int n_rows,
                 const uint32_t* const* uniforms, const uint32_t* de_uniforms,
                 const uint32_t* viz_uniforms);
int p3_batch_commit(p3_batch* b, int max_shapes, int lanes);
int p3_batch_launch(p3_batch* b);
int p3_batch_sync(p3_batch* b);
int p3_batch_elapsed(p3_batch* b, float* ms);
int p3_batch_frames(p3_batch* b, int* out);
int p3_batch_stats(p3_batch* b, uint32_t* out4);
int p3_batch_read_image(p3_batch* b, int frame, uint8_t* out, size_t n);
int p3_batch_read_histogram(p3_batch* b, int frame, uint32_t* out, size_t n);
const char* p3_batch_last_error(p3_batch* b);
/* Compiles the union chaos kernel of n shapes ahead of time (no device
 * needed), as p3_batch_commit would for a group of those shapes in that
 * order; each shape is given as for p3_chaos_prepare. */
int p3_union_prepare(int n, const float* const* xforms, const int* n_slots, const int* n_rows,
                     const int* final_idx);

/* Message for the last failing call on this handle (or of p3_create, and of
 * p3_iterate_noalloc, when r is null). Never null. */
const char* p3_last_error(p3_renderer* r);

#ifdef __cplusplus
}
#endif

#endif /* PYR3HIP_H */
