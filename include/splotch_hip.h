/*
 * splotch_hip.h — C-ABI of the MI355X (gfx950) Splotch particle-splat renderer.
 *
 * Drop-in boundary for the render dispatch of Splotch (splotch/splotch.cc:248-264,
 * server/server.cc:158-171): it replaces the CUDA backend entry points
 *   cuda_rendercontext_init   cuda/cuda_splotch.h:33   (cuda/cuda_splotch.cc:26-64)
 *   cuda_rendering            cuda/cuda_splotch.h:35   (cuda/cuda_splotch.cc:121-192)
 * and computes what the host path host_rendering() computes
 *   splotch/splotch_host.h:43 (splotch/splotch_host.cc:711-834).
 *
 * Conventions
 *  - Plain pointers and sizes only; every function returns 0 on success and a
 *    negative SPLHIP_E* code on error (message via splhip_last_error).
 *  - Device pointers unless stated otherwise ("host" in the name / comment).
 *  - Work is queued on the given hipStream_t (passed as void*; NULL = the
 *    null stream), in call order.  A render call returns once its work is
 *    queued; to size its launches it waits on the host for the stream up to
 *    three times (pass sizes, bin entries, work items), except under stream
 *    capture of a reserved batch (splhip_capture_reserve), where it queues
 *    everything without waiting.  A context is used by one host thread at a
 *    time.
 *  - Images are x-major RGB float32: pic[(x*yres + y)*3 + c], the layout of
 *    arr2<COLOUR> pic (cxxsupport/arr.h:474-488) that Splotch composites and
 *    writes.  The raw image is what render_new leaves in pic
 *    (splotch_host.cc:675-691); the tone map of splotch.cc:275-284 is applied
 *    by the caller, as in the reference.
 */
#ifndef SPLOTCH_HIP_H
#define SPLOTCH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPLHIP_MAX_TYPES 8

#define SPLHIP_OK 0
#define SPLHIP_E_ARG (-1)       /* invalid argument */
#define SPLHIP_E_HIP (-2)       /* HIP runtime error */
#define SPLHIP_E_DEGENERATE (-3) /* degenerate camera matrix (kernel/transform.cxx:84) */
#define SPLHIP_E_CAPACITY (-4)  /* a size exceeds what the library supports */
#define SPLHIP_E_STATE (-5)     /* call out of order (e.g. backward before forward) */

typedef struct splhip_ctx splhip_ctx;

/* Per-frame camera: what splotch_host.cc:62-109 derives from the .par keys and
   the scene's campos/lookat/sky/centerpos.  Filled by splhip_camera_setup. */
typedef struct {
  float mat[12];        /* TRANSFORM::Matrix().p (kernel/transform.h:36-40) */
  double xshift, yshift;
  float res2, ycorr, xfac, fovfct, dist;
  float zminval, zmaxval, minrad_pix;
  int32_t xres, yres, projection;
} splhip_camera;

/* Render parameters (the .par keys of splotch_host.cc:76-81,256-265,806-825). */
typedef struct {
  int32_t ptypes;                          /* ptypes (1) */
  float smooth_fac[SPLHIP_MAX_TYPES];      /* smooth_factor{t} (1) */
  float brightness[SPLHIP_MAX_TYPES];      /* brightness{t} * b_brightness (1) */
  int32_t col_vector[SPLHIP_MAX_TYPES];    /* color_is_vector{t} (0) */
  int32_t a_eq_e[SPLHIP_MAX_TYPES];        /* a_eq_e (1); per type (see splhip_render):
                                              1 additive, 0 ordered, 2 both */
  int32_t sort_type;                       /* sort_type (1) */
  float gray_absorption;                   /* gray_absorption (0.2) */
  int32_t tile_size;                       /* tile_size_cpu (50): reference tile */
  int32_t hip_region;                      /* hip_region: GPU work region, px: 8, 16, 32, 64 (0 = auto) */
  int32_t hip_split;                       /* hip_split: list length from which an 8x8 block is
                                              splatted as one work item per channel
                                              (0 = auto: 32768) */
  int32_t hip_variant;                     /* hip_variant: 8x8 splat inner loop, 0 = auto (4),
                                              1 = per particle, 2 = 4-particle groups,
                                              3 = software pipelined, 4 = chunked branch-free */
  int32_t hip_heavy;                       /* hip_heavy: list length from which an 8x8 block is
                                              splatted by a producer/consumer workgroup
                                              (0 = auto: 32768, < 0 = never) */
  int32_t hip_heavy_max;                   /* hip_heavy_max: at most this many such workgroups
                                              per pass, the heaviest (0 = auto: 192) */
  int32_t hip_chunk_entries;               /* hip_chunk_entries: a single-view pass with more
                                              (block, particle) entries is rendered in chunks
                                              of consecutive render order, each continuing
                                              the pixels' chains (0 = auto: 2^31 - 2^20) */
  int32_t hip_autotune;                    /* hip_autotune: 1 = on the first render of a
                                              batch bucket (particle count, pixel count and
                                              pass mix, each by factors of sqrt(2)), time the
                                              work region hip_region (8, 16, 32, 64 px) and,
                                              for 8 px, the 8x8 splat's decomposition
                                              (hip_heavy_max, hip_split, hip_variant and
                                              hip_ppw), and reuse the fastest for the bucket
                                              (0 = off) */
  int32_t hip_bin_group;                   /* hip_bin_group: 8 px units are binned through
                                              regions of g x g units, then split: 4, 2, or
                                              1 = sort the unit entries directly (0 = auto: the
                                              largest of 4, 2 whose region ids and unit masks
                                              fit 32-bit keys, when it merges entries) */
  int32_t hip_ppw;                         /* hip_ppw: particles per chunk of the 8x8 splat's
                                              walk (hip_variant 4): 1, 2, 4 or 8 (0 = auto:
                                              8 for one channel, 8 additive / 4 ordered for
                                              three) */
  int32_t hip_stage;                       /* hip_stage: list length from which an 8x8 block
                                              of an additive pass is staged: its per-pixel
                                              terms are computed across the GPU first, then
                                              summed in list order by one wave per channel
                                              (0 = auto: 8192, but none for a single view
                                              with both passes in one chunk, whose additive
                                              splat runs beside the ordered pass; < 0 =
                                              never) */
  int32_t hip_stage_ordered;               /* hip_stage_ordered: the same for an ordered pass
                                              (per-pixel operands of the compositing update
                                              computed first; not when a particle's q_c is
                                              non-finite) (0 = auto: when the longest list
                                              x 216 reaches the pass's bin entries, the
                                              lists from about the longest / 16 on; else,
                                              when it x 432 does, the lists from entries /
                                              432 on; else none; thresholds are levels
                                              8192 x 2^(j/2); < 0 = never) */
} splhip_params;

/* SoA particle fields (device pointers), post particle_normalize:
   x,y,z world position; hsml smoothing length (particle_sim::r); inten (I);
   col = colour scalar (e.r) or, for color_is_vector types, red with c1/c2 the
   green/blue fields (may be NULL when no type is vector-coloured);
   type: particle type per particle (NULL = all type 0). */
typedef struct {
  const float *x, *y, *z, *hsml, *inten, *col, *c1, *c2;
  const uint8_t *type;
  int64_t n;
} splhip_particles;

/* Reference particle_sim record (splotch/splotchutils.h:52-72), 36 bytes. */
typedef struct {
  float er, eg, eb, x, y, z, r, I;
  uint16_t type;
  uint8_t active;
  uint8_t pad_;
} splhip_particle_sim;

/* --- context ------------------------------------------------------------- */

/* Replaces cuda_rendercontext_init + cuda_renderer_init
   (cuda/cuda_splotch.cc:26-118): binds a device and builds the exp tables
   (splotch/splotchutils.h:174-215) on the host. */
int splhip_create(int device, splhip_ctx **out);
void splhip_destroy(splhip_ctx *ctx);
const char *splhip_last_error(const splhip_ctx *ctx);

/* --- host-side setup ------------------------------------------------------ */

/* Camera setup of splotch_host.cc:62-109 (kernel/transform.cxx:28-219);
   replaces cu_get_trans_params (cuda/cuda_utils.cu:47-118).  Host only. */
int splhip_camera_setup(const double campos[3], const double lookat[3], const double sky[3],
                        const double centerpos[3], int32_t xres, int32_t yres, float fov,
                        float zmin, float zmax, int32_t projection, float minrad_pix,
                        splhip_camera *out);

/* Defaults of the .par keys (the find<T>(key, default) values). Host only. */
int splhip_params_default(splhip_params *p);

/* Reads the hot-path .par keys from `par_text` ("key=value" lines, '#'
   comments; the paramfile format, cxxsupport/paramfile.cc) into p, cam_*
   (camera keys) and b_brightness-free brightness.  Keys and defaults as read
   by splotch_host.cc:62-109,256-265,806-825.  Host only.
   scalars_out = {xres, yres, fov, zmin, zmax, projection, minrad_pix} (float). */
int splhip_params_from_par(const char *par_text, float b_brightness, splhip_params *p,
                           float scalars_out[7]);

/* Palette for one type: n nodes x[i] (host double, ascending) and RGB (host
   float, 3n).  Replaces setup_colormap / cu_init_colormap
   (cuda/cuda_splotch.cc:195-234, cuda/cuda_utils.cu:259-271). */
int splhip_set_palette(splhip_ctx *ctx, int32_t type, int32_t n, const double *x,
                       const float *rgb);

/* --- plain render --------------------------------------------------------- */

/* Full hot path on device-resident SoA particles: project + colorize
   (splotch_host.cc:58-287), depth sort for ordered types (:289-321),
   binning and splat (render_new, :330-697).  Writes the raw image `pic`
   (device, xres*yres*3 floats).  Particles are not modified.
   Types with a_eq_e[t]=1 are splatted additively; types with a_eq_e[t]=0 are
   composited in sort order; types with a_eq_e[t]=2 take part in both passes.
   When both kinds are present the image is pic = P_ordered + P_additive, each
   rendered from zero as render_new does for its mode (the reference's a_eq_e
   is one global flag). */
int splhip_render(splhip_ctx *ctx, const splhip_particles *parts, const splhip_camera *cam,
                  const splhip_params *prm, float *pic, void *stream);

/* cuda_rendering semantics (cuda/cuda_splotch.cc:121-192): host AoS particle
   records in, host image out; the host records are not modified. */
int splhip_render_host(splhip_ctx *ctx, const splhip_particle_sim *host_particles, int64_t n,
                       const splhip_camera *cam, const splhip_params *prm, float *host_pic,
                       void *stream);

/* --- batched render ----------------------------------------------------------- */

/* One view of a batch: its particles (device SoA), camera, render keys and
   output image (device, xres*yres*3 floats). */
typedef struct {
  splhip_particles parts;
  splhip_camera cam;
  splhip_params prm;
  float *pic;
} splhip_job;

/* Many views in one pipeline (no reference counterpart as one call: it is the
   reference's per-frame host_rendering / cuda_rendering (splotch.cc:248-264,
   scenemaker.cc:599-673 frame loop) applied to njobs frames).  Every image is
   bit-identical to splhip_render of that job alone.  All jobs share the
   context's palettes; the hip_* tuning keys are taken from jobs[0].prm, and
   batches use 8 px work regions.  Stages run once over all jobs (one launch
   per stage, heaviest work items of all views first); the host waits three
   times per pipeline for sizes.  Consecutive views go through one pipeline
   up to 2^28 particles (a larger view alone); a pipeline with a pass past
   hip_chunk_entries bin entries is split into smaller groups of views, down
   to single views rendered in chunks.
   Under stream capture (hipStreamBeginCapture on `stream`) the call queues
   the reserved batch (splhip_capture_reserve) with device-side counts: no
   host wait, allocation, host copy or memset, so the captured graph can be
   replayed; each replay renders the reserved views from the current values
   of their particle arrays. */
int splhip_render_batched(splhip_ctx *ctx, int32_t njobs, const splhip_job *jobs, void *stream);

/* --- capture (no reference counterpart) ---------------------------------------- */

/* Reserve a batch for rendering under stream capture: renders it once as
   splhip_render_batched does (images written), sizes the per-pass bin-entry
   capacities from its counts (x headroom >= 1, + 4096), fixes its binning
   form and tuning, then renders it once more with device-side counts into a
   workspace of its own (images written again), which is then frozen.  A
   later splhip_render_batched (or splhip_render, njobs = 1) of the same views
   (same particle arrays, counts, cameras, params and images) on a capturing
   stream queues that device-count pipeline.  One reservation per context: a
   new one invalidates graphs captured under the previous one, as does
   destroying the context or giving a palette more entries.  Needs 8 px work
   regions and passes within hip_chunk_entries. */
int splhip_capture_reserve(splhip_ctx *ctx, int32_t njobs, const splhip_job *jobs,
                           float headroom, void *stream);

/* After a replay (waits for the device): *overflow bit p (0 ordered, 1
   additive) is set when a pass had more bin entries than reserved; that pass
   was left out of the images (nothing is written past a buffer).  Reserve
   again with more headroom. */
int splhip_capture_status(splhip_ctx *ctx, int32_t *overflow);

/* --- differentiable form ------------------------------------------------------ */

/* Gradient outputs (device, n floats each, original particle order).  c1, c2:
   green / blue fields of color_is_vector types (may be NULL when no type is
   vector-coloured).  screen (optional, n*6 floats): the gradient with respect
   to the screen-space state {x, y, r, e.r, e.g, e.b} of each projected
   particle. */
typedef struct {
  float *x, *y, *z, *hsml, *inten, *col, *c1, *c2;
  float *screen;
} splhip_grads;

/* Backward of the last splhip_render on this context (no reference
   counterpart: Splotch has no gradient path).  dL_dpic: device, xres*yres*3
   floats, the gradient of a scalar loss with respect to that render's raw
   image.  The particles, camera and params of that render must be unchanged
   (the forward's bins and projected particles are reused; they stay valid
   until the next call that renders or bins on this context).
   The derivative is that of the smooth model on the forward's frozen
   footprint: each (pixel, particle) pair that render visits contributes with
   the true exponential, att = exp(stp*(y-sy)^2)*exp(stp*(x-sx)^2),
   additive P += -e*att, ordered P = q + (P-q)*exp(-e*att) (the table-free
   limit of splotch_host.cc:615,648-658); outputs d/d(x, y, z, hsml, I,
   colour) through particle_project/particle_colorize (splotch_host.cc:111-287).
   Requires 8 px work regions (hip_region 0 or 8) in that render.
   Returns SPLHIP_E_STATE when there is no such forward. */
int splhip_render_backward(splhip_ctx *ctx, const float *dL_dpic, const splhip_grads *out,
                           void *stream);

/* --- parity hooks ----------------------------------------------------------- */

/* particle_project + particle_colorize only.  rec_out: device, n*8 floats per
   particle {x, y, r, e.r, e.g, e.b, z, I} (state after the reference loops);
   active_out: device n bytes, 0 inactive, else the passes of the particle
   (bits: 1 ordered, 2 additive). */
int splhip_project(splhip_ctx *ctx, const splhip_particles *parts, const splhip_camera *cam,
                   const splhip_params *prm, float *rec_out, uint8_t *active_out,
                   void *stream);

/* Tile index lists of render_new's binning (splotch_host.cc:337-406) for the
   particles of one mode (ordered=1: a_eq_e=0 types, in sort order; 0: a_eq_e=1
   types, ascending index), at tile size prm->tile_size.  Tile id = iy*ncx+ix.
   Host outputs: tile_off[ntiles+1] (int64) and idx[] = original particle
   indices; *n_entries returns the total (call with idx=NULL to size).
   order_out (host, optional): the render order (original indices). */
int splhip_bins(splhip_ctx *ctx, const splhip_particles *parts, const splhip_camera *cam,
                const splhip_params *prm, int32_t ordered, int64_t *tile_off,
                uint32_t *idx, int64_t idx_cap, int64_t *n_entries, uint32_t *order_out,
                int64_t *n_order, void *stream);

/* --- stage timing ------------------------------------------------------------ */

/* The reference times its stages with the tstack timers "3D transform",
   "Particle coloring", "Particle sorting", "Host Chunk preparation" and
   "Host Rendering proper" (splotch_host.cc:737-833, cuda/cuda_render.cu:251-285).
   With profiling on, each render records HIP events at the stage boundaries
   (a stage may span several intervals of the stream; their times are summed);
   splhip_stage_times returns the last frame's GPU time per stage in ms:
     [0] project + colorize
     [1] ordered pass: compaction + depth sort     [4] additive pass: compaction
     [2] ordered pass: binning + work items        [5] additive pass: binning + work items
     [3] ordered pass: splat (all its kernels)     [6] additive pass: splat (all its kernels)
   and single kernels, each bracketed by events on the stream it runs on:
     [7] ordered k_splat8 launch (normal items)    [9] additive k_splat8 launch
     [8] ordered k_splat_heavy launch (auxiliary   [10] additive k_splat_heavy launch
         stream, concurrent with [7])
     [11] ordered staged long lists: operands and  [12] additive staged long lists
          chains (third stream, concurrent with [7])
   counts (optional, 4 values): active particles and bin entries of the
   ordered pass, then of the additive pass. */
#define SPLHIP_NSTAGES 13
int splhip_set_profiling(splhip_ctx *ctx, int32_t enable);
int splhip_stage_times(splhip_ctx *ctx, float *ms, int64_t *counts);

/* Work-item trace of the 8x8-block splat (no reference counterpart; the
   reference's tstack timers stop at stage granularity).  With tracing on, each
   render records per work item {start, end} (s_memrealtime, 100 MHz), the list
   length | channel << 32 (channel 3 = all) and the hardware CU id;
   splhip_trace_items copies the last frame's items of one pass (4 uint64 each;
   call with out=NULL to get the count in *n). */
int splhip_set_trace(splhip_ctx *ctx, int32_t enable);
int splhip_trace_items(splhip_ctx *ctx, int32_t ordered, uint64_t *out, int64_t cap, int64_t *n);

/* The splat tuning of the last render: {hip_variant, hip_split, heavy
   workgroup cap (0 = none), 1 if the autotuner chose it, work region in px
   (hip_region), hip_ppw (0 = auto)}.  No reference counterpart. */
int splhip_last_tuning(splhip_ctx *ctx, int64_t out[6]);

/* How the last render binned its passes (no reference counterpart): out[p] =
   the (16x16 region, particle) entries pass p (0 ordered, 1 additive) was
   binned through, or 0 where its (8x8 block, particle) entries were sorted
   directly.  Both give the same block lists; regions are used where they
   merge at least a third of the entries. */
int splhip_last_binning(splhip_ctx *ctx, int64_t out[2]);

/* --- post-processing and 8-bit output --------------------------------------- */

/* The master's epilogue after the render (splotch.cc:272-310), which the
   reference runs on the host over the float image:
     a_eq_e      c = -xexp.expm1(c) per channel            (splotch.cc:275-284)
     gamma, brightness, contrast (pic_gamma 1, pic_brightness 0, pic_contrast 1):
                 c = float(contrast * pow(double(c), gamma) + brightness), only
                 when one differs from its default   (colour_adjust, splotchutils.cc:230-250)
     colorbar    one 10-row bar per type t < ptypes with col_vector[t] == 0, from
                 the image's last row up, with the palette colour at
                 ix / double(xres)                   (add_colorbar, splotchutils.cc:35-56;
                 rows that would fall below row 0 are skipped)
   then the 8-bit conversion of LS_Image::put_pixel, byte = max(0, min(255,
   int(c * 256))) with x86 truncation (ls_image.h:65-75). */
typedef struct {
  int32_t a_eq_e;
  int32_t colorbar;
  int32_t ptypes;
  int32_t col_vector[SPLHIP_MAX_TYPES];
  double gamma, brightness, contrast;
} splhip_post;

/* The .par keys a_eq_e, pic_gamma, pic_brightness, pic_contrast, colorbar,
   ptypes and color_is_vector{t}, with the reference's defaults. */
int splhip_post_from_par(const char *par_text, splhip_post *out);

/* One pass over the device image pic (x-major RGB float, xres * yres * 3;
   3 xres yres < 2^31).  pic_out (device, may equal pic, may be NULL): the
   post-processed float image.  rgb8 (device, may be NULL): 3 xres yres bytes,
   layout 0 = x-major RGB (LS_Image's pixel[ix][iy]), layout 1 = the body of
   the file LS_Image::write_TGA writes (ls_image.cc:178-202: rows iy
   ascending, pixels ix ascending, B G R), to follow its 18-byte header
   {0,0,2, 0,0,0,0,0, 0,0,0,0, xres%256, xres/256, yres%256, yres/256, 24, 32}.
   Colorbars use the context's palettes (splhip_set_palette).  Stream-ordered. */
int splhip_postprocess(splhip_ctx *ctx, const float *pic, int32_t xres, int32_t yres,
                       const splhip_post *post, float *pic_out, uint8_t *rgb8, int32_t layout,
                       void *stream);

/* --- particle normalization --------------------------------------------------- */

/* sceneMaker::particle_normalize (splotch/scenemaker.cc:40-296, host branch),
   which the scene maker runs on the particles before each render: per type t,
   intensity_log{t} replaces I by log10(I) (I <= 0 becomes -38 and stays out of
   the range), color_log{t} does the same to the colour (the second and third
   components of a vector colour are logged unchecked), color_asinh{t} applies
   asinh otherwise; then the type's data ranges (min / max over the values,
   NaN never enters) are replaced by intensity_min{t} ... color_max{t} where
   given, and Normalizer::normAndClamp maps each value v to
   (max(min, min(max, v)) - min) / (max - min), or 1 when max == min
   (splotch/splotchutils.h:138-171).  The intensity is clamped too unless the
   reference is built with NO_I_NORM (norm_intensity = 0). */
typedef struct {
  int32_t ptypes;
  int32_t intensity_log[SPLHIP_MAX_TYPES], color_log[SPLHIP_MAX_TYPES];
  int32_t color_asinh[SPLHIP_MAX_TYPES], col_vector[SPLHIP_MAX_TYPES];
  int32_t has_intensity_min[SPLHIP_MAX_TYPES], has_intensity_max[SPLHIP_MAX_TYPES];
  int32_t has_color_min[SPLHIP_MAX_TYPES], has_color_max[SPLHIP_MAX_TYPES];
  float intensity_min[SPLHIP_MAX_TYPES], intensity_max[SPLHIP_MAX_TYPES];
  float color_min[SPLHIP_MAX_TYPES], color_max[SPLHIP_MAX_TYPES];
  int32_t norm_intensity;
} splhip_norm;

/* Per type, the ranges particle_normalize finds and applies (it records the
   colour ones as the parameters range_min{t}, range_max{t}, clamp_min{t} and
   clamp_max{t}).  A type without particles keeps 1e37 / -1e37. */
typedef struct {
  float range_min, range_max;          /* colour values of the data */
  float clamp_min, clamp_max;          /* colour range applied */
  float intensity_min, intensity_max;  /* intensity range applied */
  float size_min, size_max;            /* smoothing lengths of the data */
} splhip_norm_ranges;

/* The keys above from .par text (ptypes, intensity_log{t}, color_log{t},
   color_asinh{t}, color_is_vector{t}, intensity_min{t}, intensity_max{t},
   color_min{t}, color_max{t}); norm_intensity = 1. */
int splhip_norm_from_par(const char *par_text, splhip_norm *out);

/* Normalizes n particles' device fields in place: inten, col (the scalar
   colour or first component), c1 and c2 (second and third components; needed
   when a type has a vector colour, else may be NULL); hsml and type (NULL:
   all type 0) are read.  Particles of a type >= ptypes are left as they are.
   ranges (host, ptypes entries, may be NULL): when given, the call waits for
   the stream.  log10 and asinh are evaluated in double and rounded to float:
   they may differ from glibc's log10f / the reference's double log by one
   float ulp (log10f is not correctly rounded; about 1.4% of arguments). */
int splhip_normalize(splhip_ctx *ctx, int64_t n, float *inten, float *col, float *c1, float *c2,
                     const float *hsml, const uint8_t *type, const splhip_norm *norm,
                     splhip_norm_ranges *ranges, void *stream);

/* splhip_normalize without allocation (no reference counterpart; for callers
   that cannot allocate device memory: code under HIP graph capture, tight
   serving loops, harnesses that forbid hipMalloc).  Same arguments, results
   and errors as splhip_normalize; the per-type extremes live in a file-scope
   device array of 64 uint32 (256 bytes) per GPU instead of the context's
   workspace, cleared by the call itself.  Limit: ptypes <= 8 (6 words per
   type), else SPLHIP_E_CAPACITY.  Every call on one GPU shares that array:
   calls that overlap on one GPU (other streams or contexts) race. */
int splhip_normalize_noalloc(splhip_ctx *ctx, int64_t n, float *inten, float *col, float *c1,
                             float *c2, const float *hsml, const uint8_t *type,
                             const splhip_norm *norm, splhip_norm_ranges *ranges, void *stream);

/* --- SoA ingest of block binary files ------------------------------------------ */

/* bin_reader_block (reader/bin_reader.cc:34-88,147-190): the file `infile`
   holds num_blocks float columns of npart values each (npart = file size /
   (4 num_blocks)); the .par keys x, y, z, r, I, C1, C2, C3 give 1-based
   column numbers (r, I, C2, C3 optional), swap_endian (default T) swaps each
   float's bytes, smooth_factor (default 1) scales r.  Missing columns take the
   reader's defaults: r = smooth_factor, I = 0.5, C1 = 1, C2 = C3 = 0; type 0.
   splhip_bin_block_count returns npart; splhip_read_bin_block reads the
   columns straight into the caller's device arrays (any may be NULL to skip
   it; n must equal npart), through pinned staging, and fixes byte order and
   scale on the GPU.  The whole file is one rank's share (the reference's MPI
   split is calcShare over ranks).  Stream-ordered; the call returns when the
   file has been read and the copies queued. */
int splhip_bin_block_count(const char *par_text, int64_t *npart);
int splhip_read_bin_block(splhip_ctx *ctx, const char *par_text, int64_t n, float *x, float *y,
                          float *z, float *hsml, float *inten, float *col, float *c1, float *c2,
                          uint8_t *type, void *stream);

/* --- scene files: the frames of an animation ----------------------------------- */

/* One scene of a scene file: the camera getNextScene takes from the
   parameters after applying the scene's keys (scenemaker.cc:1140-1166:
   camera_*, lookat_*, sky_* (default 0, 0, 1), center_* (default the camera
   position)), its fidx (default the .par fidx, else 0), keep_particles (the
   next scene has the same fidx within 1e-5 relative, so its particles are
   reused: scenemaker.cc:656-663) and line, the 0-based line of scene_text the
   scene came from (for callers that apply its other keys). */
typedef struct {
  double campos[3], lookat[3], sky[3], center[3];
  double fidx;
  int32_t keep_particles;
  int32_t line;
} splhip_scene;

/* The scene list sceneMaker builds from scene_text (scenemaker.cc:599-673):
   a header line of parameter names (or, when the first line holds ten
   numbers, the old format: camera_x camera_y camera_z lookat_x lookat_y
   lookat_z sky_x sky_y sky_z fidx, with that line as data), then one line of
   values per scene, from line scene_start (par key, default 0) in steps of
   scene_incr (1) while the scene number is below last_scene (1000000).
   Scenes after the capacity are counted in *n but not written.  Host only. */
int splhip_scene_file_parse(const char *par_text, const char *scene_text, splhip_scene *out,
                            int32_t cap, int32_t *n);

/* Host-built tables (device copies are used by the kernels): tab[0..1023] =
   tab1, tab[1024..2047] = tab2; consts = {expfac, taylorlimit, h2sigma, bfak,
   rfac}.  Host only. */
int splhip_exptable(float tab[2048], float consts[5]);

#ifdef __cplusplus
}
#endif

#endif /* SPLOTCH_HIP_H */
