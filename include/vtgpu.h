/* vtgpu.h: the C-ABI of the MI355X vision path: Gaussian pyramid, sparse pyramidal Lucas-Kanade, NCC template
 * search. Plain pointers and sizes; no HIP or torch types cross it. Every entry point returns 0 or a negative
 * errno-style code (VTG_E*) and leaves a reason in vtg_ctx_error(). One host thread per context; contexts are
 * independent. All work of a context is ordered on its one HIP stream.
 *
 * Two surfaces:
 *   (1) reference-shaped calls over host planes (struct vt_plane, fp32 Q6 with replicated padding), the drop-in for
 *       the reference's kernel-level functions: each copies in, runs the GPU kernels, copies out.
 *   (2) device-resident calls: pyramids live in HBM (vtg_pyr), points and results are device buffers. This is the
 *       hot path a host program batches through.
 * The reference interface each entry point replaces is cited beside it (paths relative to the reference tree). */
#ifndef VTGPU_H
#define VTGPU_H

#include <stddef.h>
#include <stdint.h>

#include "vt_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VTG_MAX_LEVELS 6 /* VT_MAX_LEVELS, src/common/vt_proto.h:59 */

enum {
    VTG_OK = 0,
    VTG_EINVAL = -22, /* bad argument: sizes, levels, window, pointers */
    VTG_ENOMEM = -12, /* device or host allocation failed */
    VTG_EIO = -5,     /* a HIP call failed (launch, copy, device lost) */
    VTG_ERANGE = -34, /* a size beyond what the kernels support */
    VTG_EAGAIN = -11, /* a served request still in flight */
};

/* ------------------------------------------------------------------ context and memory */

typedef struct vtg_ctx vtg_ctx;

/* A context on HIP device `device` with its own stream (default priority). */
int vtg_ctx_create(int device, vtg_ctx **out);
int vtg_ctx_destroy(vtg_ctx *ctx);
/* Wait for all work of the context. */
int vtg_ctx_sync(vtg_ctx *ctx);
/* The last failure's reason, "" if none. */
const char *vtg_ctx_error(const vtg_ctx *ctx);
/* Work submitted to ctx from now on starts only after all work submitted to `other` so far (both on one device):
 * two contexts are two streams, and this is how a caller overlaps independent stages, e.g. the NCC search on one
 * context beside LK on another once both pyramids exist. No host wait. */
int vtg_ctx_wait(vtg_ctx *ctx, vtg_ctx *other);
/* A point in the context's stream, for a host that waits for some of its work while later work stays queued (no
 * counterpart in the reference, whose work is synchronous). vtg_ctx_mark records one after the work submitted so far
 * and returns its number (1, 2, ... in order); vtg_ctx_mark_wait returns once the work before that mark has completed.
 * The last 64 marks are exact; waiting on an older one may wait for a later point. Not under graph capture. */
int vtg_ctx_mark(vtg_ctx *ctx, uint64_t *mark);
int vtg_ctx_mark_wait(vtg_ctx *ctx, uint64_t mark);
/* The context's hipStream_t, as an opaque pointer (for callers that order their own work after it). */
void *vtg_ctx_stream(vtg_ctx *ctx);

int vtg_alloc(vtg_ctx *ctx, size_t bytes, void **dptr);
int vtg_free(vtg_ctx *ctx, void *dptr);
/* Stream-ordered copies; these two wait for completion (the host buffer may be pageable). */
int vtg_upload(vtg_ctx *ctx, void *dst_dev, const void *src_host, size_t bytes);
int vtg_download(vtg_ctx *ctx, void *dst_host, const void *src_dev, size_t bytes);
/* Stream-ordered download with no wait: dst_host must be pinned (vtg_host_alloc) and stay untouched until the
 * context's work up to this call is done (vtg_ctx_sync, or a later context that waited for this one). */
int vtg_download_async(vtg_ctx *ctx, void *dst_host, const void *src_dev, size_t bytes);
/* Stream-ordered upload with no wait: src_host must be pinned (vtg_host_alloc) and stay unchanged until the
 * context's work up to this call is done. */
int vtg_upload_async(vtg_ctx *ctx, void *dst_dev, const void *src_host, size_t bytes);
/* Pinned (page-locked) host memory: the target of vtg_download_async and a faster source/target of the waiting
 * copies. */
int vtg_host_alloc(vtg_ctx *ctx, size_t bytes, void **host);
int vtg_host_free(vtg_ctx *ctx, void *host);
/* Stream-ordered device-to-device copy and fill (no wait). */
int vtg_copy_d2d(vtg_ctx *ctx, void *dst_dev, const void *src_dev, size_t bytes);

/* Per-kernel timing with HIP events on the context's stream. While on, every launch is bracketed by two events;
 * vtg_timing_read waits for them and returns the summed milliseconds and launch count of kernels named `kernel`. */
int vtg_timing_enable(vtg_ctx *ctx, int on);
int vtg_timing_read(vtg_ctx *ctx, const char *kernel, double *total_ms, long *launches);
int vtg_timing_reset(vtg_ctx *ctx);

/* Capture the context's stream-ordered work into a replayable HIP graph: begin, make the calls (stream-ordered ones
 * only, with buffers that outlive the graph; contexts joined by vtg_ctx_wait join the capture), end. Replaying
 * repeats the same launches on the same buffers. Timing must be off while capturing. */
int vtg_graph_begin(vtg_ctx *ctx);
int vtg_graph_end(vtg_ctx *ctx, void **graph);
int vtg_graph_launch(vtg_ctx *ctx, void *graph);
int vtg_graph_destroy(vtg_ctx *ctx, void *graph);

/* ------------------------------------------------------------------ pyramids (tk_build, src/kern/track.c:584-629) */

/* Storage of the levels in HBM. Q6 is exact: the reference's integer-valued fp32 planes hold at most 16,320, so a
 * u8 level 0 (the frame itself, Q6 = 64 x byte) and u16 upper levels hold the same values in 1 and 2 bytes. F32 is
 * the reference's own fp32 Q6 (4 B per pixel, the plain port). F16 and BF16 hold grey = Q6 / 64 above level 0 and
 * are not exact. */
enum { VTG_STORE_Q6 = 0, VTG_STORE_F32 = 1, VTG_STORE_F16 = 2, VTG_STORE_BF16 = 3, VTG_STORE_Q6W = 4, VTG_STORE_Q6F = 5,
       VTG_STORE_F32_U8 = 6 };
/* VTG_STORE_F32_U8: every level fp32 Q6 (as VTG_STORE_F32) plus the u8 frame kept beside level 0, which the NCC
 * search's matrix-core path reads: LK reads fp32 at every level. */
/* VTG_STORE_Q6F: level 0 the u8 frame (as VTG_STORE_Q6, which the NCC search's matrix-core path reads), every level
 * above fp32 Q6 (exact integers, the reference's own plane format): LK on those levels needs no conversions. */
/* VTG_STORE_Q6W: every level u16 Q6, exact; the storage for frames whose level 0 is filtered (below). */

/* Level-0 filters of tk_build (src/kern/track.c:600-623; the tracker's `contrast` and `pre_gauss` options,
 * src/common/vt_proto.h:89-90): VTG_BUILD_CONTRAST stretches the frame's Q6 values to 0 .. 16320 by its own min and
 * max (img_minmax + img_norm, gain 16320 / span and offset -min * gain in long double, as tk_build computes them);
 * VTG_BUILD_PRE_GAUSS then blurs level 0 with the 5 x 5 binomial (img_gauss5). Both in the reference's arithmetic,
 * bit for bit. Level 0 is then no longer 64 x a byte: the pyramid must be VTG_STORE_Q6W or VTG_STORE_F32. */
enum { VTG_BUILD_CONTRAST = 1, VTG_BUILD_PRE_GAUSS = 2 };

typedef struct vtg_pyr vtg_pyr;

/* A pyramid of `levels` levels of a w x h frame, level l of size ((w + 2^l - 1) >> l) x ((h + 2^l - 1) >> l) as
 * (w + 1) / 2 per step (src/kern/img.h:30-31). */
int vtg_pyr_create(vtg_ctx *ctx, int w, int h, int levels, int storage, vtg_pyr **out);
int vtg_pyr_destroy(vtg_ctx *ctx, vtg_pyr *p);
/* No counterpart in the reference (whose level 0 is always tk_build's own plane): a one-level VTG_STORE_Q6 pyramid
 * whose level 0 is the caller's u8 device frame itself, read in place and never copied, so a search of a frame
 * (template cut, vtg_ncc_search*) can run before or beside the build of its pyramid. frame: device memory, 4-byte
 * aligned, pitch a multiple of 4 bytes; it must outlive the pyramid and stay unchanged while work reads it. Building
 * a wrapped pyramid is refused (VTG_EINVAL); vtg_pyr_destroy frees only its descriptor. */
int vtg_pyr_wrap(vtg_ctx *ctx, int w, int h, const uint8_t *frame, int pitch, vtg_pyr **out);
/* Build every level from a u8 frame (w bytes per row, `pitch` apart): level 0 = 64 x byte
 * (img_u8_to_q6_*, src/kern/img.h:16), level l = reduce(level l-1) (img_reduce_*, src/kern/img.h:30-33):
 * (sum of (1 4 6 4 1)^2 x src over the 5 x 5 at (2x, 2y), replicated border, + 128) >> 8.
 * on_device: frame is a device pointer (else host memory, copied up first). */
int vtg_pyr_build(vtg_ctx *ctx, vtg_pyr *p, const uint8_t *frame, int pitch, int on_device);
/* n pyramids of one storage at once (the batched form of tk_build for frames of any sizes): one launch per level
 * covers every pyramid that has it. frames[i], pitches[i] as for vtg_pyr_build. */
int vtg_pyr_build_many(vtg_ctx *ctx, int n, vtg_pyr *const *pyrs, const uint8_t *const *frames, const int *pitches,
                       int on_device);
/* vtg_pyr_build / vtg_pyr_build_many with the level-0 filters `flags` (VTG_BUILD_*). */
int vtg_pyr_build_ex(vtg_ctx *ctx, vtg_pyr *p, const uint8_t *frame, int pitch, int on_device, int flags);
int vtg_pyr_build_many_ex(vtg_ctx *ctx, int n, vtg_pyr *const *pyrs, const uint8_t *const *frames,
                          const int *pitches, int on_device, int flags);
/* Level l's device pointer, size, pitch in bytes and element size. */
int vtg_pyr_level(const vtg_pyr *p, int level, void **dptr, int *w, int *h, size_t *pitch_bytes, int *elem_bytes);
/* Level l into a host fp32 Q6 plane, with the reference's replicated padding filled (img_pad_rows, img_pad_tb,
 * src/kern/img.h:10-12): the form the reference's kernels read. */
int vtg_pyr_read_level(vtg_ctx *ctx, const vtg_pyr *p, int level, const struct vt_plane *dst);

/* ------------------------------------------------------------------ LK (lk_track_*, lk_flowback_*, src/kern/lk.h:45-53) */

/* The 2x2 solve: VTG_SOLVE_F64 runs the whole per-point driver in fp64 (the reference's is x87 long double);
 * VTG_SOLVE_MIXED runs it in fp32 and falls back to fp64 for a level whose G has condition number above 1e4. */
enum { VTG_SOLVE_F64 = 0, VTG_SOLVE_MIXED = 1 };

/* n points tracked from pyramid I to pyramid J from level `top` down (the OP_LK batch loop, src/kern/track.c:275-292):
 * xy[2i], xy[2i+1] the point in level-0 pixels, guess the same shape or NULL for zero. fb_max > 0 adds the
 * forward-backward check (lk_flowback_*). out: n records of struct lk_out (32 B). All device pointers. Window
 * 5..48 x 5..31; top < levels of both pyramids. The launch orders its points in the context's LK scratch (see
 * vtg_lk_set_order), grown on a call that needs more: a launch to be captured in a graph is made once eagerly
 * first, at its largest size, and a captured launch keeps the scratch it was captured with. */
int vtg_lk_track(vtg_ctx *ctx, const vtg_pyr *I, const vtg_pyr *J, int top, const struct lk_cfg *cfg, int n,
                 const float *xy, const float *guess, float fb_max, struct lk_out *out);
/* The same with the solve mode chosen. */
int vtg_lk_track_ex(vtg_ctx *ctx, const vtg_pyr *I, const vtg_pyr *J, int top, const struct lk_cfg *cfg, int n,
                    const float *xy, const float *guess, float fb_max, int solve_mode, struct lk_out *out);

/* Points of many frame pairs of any sizes in one launch per 64 pairs (the batched OP_LK): pair k tracks points
 * offsets[k] .. offsets[k+1] - 1 of xy (and guess, out) from I[k] to J[k] starting at level top[k]. One window
 * configuration and one storage for the whole batch. offsets: host array of npairs + 1 entries. */
int vtg_lk_track_many(vtg_ctx *ctx, int npairs, const vtg_pyr *const *I, const vtg_pyr *const *J, const int *top,
                      const int *offsets, const struct lk_cfg *cfg, const float *xy, const float *guess, float fb_max,
                      int solve_mode, struct lk_out *out);
/* Both pyramids of npairs frame pairs built AND every pair's points tracked in ONE persistent launch (the batched
 * tk_build + OP_LK, track.c:584-629 and 275-292): pair k builds I[k] from device frame0[k] and J[k] from frame1[k]
 * (u8, pitch0[k] / pitch1[k] bytes per row), then tracks points offsets[k] .. offsets[k+1] - 1 of xy from I[k] to J[k]
 * from level top[k]. Workgroups take work items in dependency order (a pair's reduce tiles level by level, then its
 * LK blocks of 4 points) and an item waits only for the items it reads, so one pair's LK overlaps later pairs'
 * pyramids. Results equal vtg_pyr_build_many + vtg_lk_track_many bit for bit. Q6 pyramids with at least 2 levels,
 * up to 64 pairs. Stream-ordered; graph-capturable (no allocation after the first call). */
int vtg_pyr_lk_fused(vtg_ctx *ctx, int npairs, vtg_pyr *const *I, vtg_pyr *const *J, const uint8_t *const *frame0,
                     const uint8_t *const *frame1, const int *pitch0, const int *pitch1, const int *top,
                     const int *offsets, const struct lk_cfg *cfg, const float *xy, const float *guess, float fb_max,
                     int solve_mode, struct lk_out *out);
/* The context's device error word (waits for the context): nonzero when a fused launch's bounded wait gave up, whose
 * results are then not valid; reading clears it. */
int vtg_ctx_device_error(vtg_ctx *ctx, int *code);
/* The LK schedule where no autotuner pick applies: wavefronts per workgroup, 1 .. 8 (default 1). A wavefront tracks 4
 * points, one per row of 16 lanes. Results are the same bit for bit under every schedule; only time differs. */
int vtg_lk_set_waves(vtg_ctx *ctx, int waves);
/* The LK points per wavefront where no autotuner pick applies: 4 (one point per row of 16 lanes), 3 (21 x 21
 * windows: three points' vector lanes in three rows, the fourth row computing their second block's 5 columns), or 0
 * (default: 3 where the window allows it, else 4). Results are the same bit for bit; only time differs. */
int vtg_lk_set_ppw(vtg_ctx *ctx, int ppw);
/* Whether an LK launch tracks its points in spatial order (1, the default: each pair's points by 32 x 32-pixel cell of
 * their position, so a wavefront's points are neighbours) or in the order given (0). Records land at each point's own
 * index either way and are the same bit for bit; only time differs. */
int vtg_lk_set_order(vtg_ctx *ctx, int on);
/* The template's gradient operator for the context's later LK launches. VTG_GRAD_SOBEL (the default) is the
 * reference's: [1 2 1] smoothing across, [-1 0 1] along, / 8 per grey level (lk_ref.c:30-31, lk_vx.c:60-75), and only
 * it reproduces the reference's records. VTG_GRAD_SCHARR is a labelled variant with no counterpart in the reference:
 * [3 10 3] across, [-1 0 1] along, / 32, computed as fmaf(centre difference, 10, 3 (sum of the outer differences))
 * / 2048 in Q6 units. Everything else (window, sums, driver, statuses) is unchanged. */
enum { VTG_GRAD_SOBEL = 0, VTG_GRAD_SCHARR = 1 };
int vtg_lk_set_gradient(vtg_ctx *ctx, int grad);

/* ------------------------------------------------------------------ autotuner (no counterpart in the reference, whose
 * schedule is fixed at build time: one thread per row range or point, track.c:210-292) */

/* What the autotuner measured and keeps for a size class (level-0 size, levels, storage; for LK also top level and
 * window). */
struct vtg_tune {
    int pyr_rows;                         /* pyramid reduce tile height: 8, 16 or 32 output rows */
    int lk_waves;                         /* LK schedule: wavefronts per workgroup */
    float pyr_us, lk_us;                  /* median times of the picks: both pyramids; the given points */
    float pyr_us_default, lk_us_default;  /* the same under the defaults (16 rows; the context's LK waves) */
    int candidates;                       /* schedules timed */
    int lk_ppw;                           /* LK schedule: points per wavefront (3 or 4) */
};

/* Time every candidate schedule of the pyramid reduce and of the LK kernel on this pair (device frames f0, f1 of
 * `pitch` bytes per row, built into I and J) and these n device points, and keep the fastest for the pair's size
 * class in the context: later builds and LK launches of pyramids of that class use them (the context's waves,
 * vtg_lk_set_waves, apply to classes not tuned; a batch of several classes takes the pick of its pair with the most
 * points, so it stays one launch). Waits; allocates a temporary output. */
int vtg_tune_pair(vtg_ctx *ctx, vtg_pyr *I, vtg_pyr *J, const uint8_t *f0, const uint8_t *f1, int pitch, int top,
                  const struct lk_cfg *cfg, int n, const float *xy, int solve_mode, struct vtg_tune *res);
/* The picks in force for a size class (the defaults when it was not tuned; times left 0). */
int vtg_tune_get(vtg_ctx *ctx, const vtg_pyr *I, int top, const struct lk_cfg *cfg, struct vtg_tune *res);
/* Forget every pick. */
int vtg_tune_reset(vtg_ctx *ctx);

/* ------------------------------------------------------------------ NCC search (src/kern/track.c:184-208, 301-329, 935-974, 1615-1734) */

typedef struct vtg_tmpl vtg_tmpl;

/* One cell of a search's map: the best score of its 16 x 16 template corners and where (OP_SEARCH, track.c:301-329). */
struct vtg_cell { float v; int x, y; };

/* What a search found (search_slot, track.c:1661-1726): the best corner in half pixels of the level after the top-8
 * half-pixel refinement, its score, the best score clear of it by half a template, the parabola offsets of the peak
 * (peak_offset, track.c:1650-1657) and the whole-pixel argmax of the map. */
struct vtg_found {
    int hx, hy;          /* half pixels; -1 score when nothing was searched */
    double ncc, second;
    double ox, oy;       /* sub-half-pixel offsets of the peak, each in [-0.5, 0.5] */
    int cells;           /* cells in the map */
    int best_x, best_y;  /* whole-pixel argmax over the map, first in cell order */
    float best_v;
};

/* A zero-mean template of the tw x th window at corner (x0, y0) of pyramid level `level` (store_appearance,
 * track.c:935-974; the reference's is 32 x 21, here tw a multiple of 16 up to 64, th up to 64). */
int vtg_ncc_template(vtg_ctx *ctx, const vtg_pyr *p, int level, int x0, int y0, int tw, int th, vtg_tmpl **out);
int vtg_ncc_template_destroy(vtg_ctx *ctx, vtg_tmpl *t);
/* Re-cut an existing template of the same size from another window or pyramid (store_appearance into a slot that
 * already exists, track.c:935-974), stream-ordered, no allocation. */
int vtg_ncc_template_update(vtg_ctx *ctx, vtg_tmpl *t, const vtg_pyr *p, int level, int x0, int y0);
/* n re-cuts at once (store_appearance for several slots or streams in one round, track.c:935-974): template t[k] from
 * level[k] of p[k] at corner (x0[k], y0[k]), host arrays; one launch per run of up to 8 that share template size and
 * level element type. Stream-ordered, no allocation. */
int vtg_ncc_template_update_many(vtg_ctx *ctx, int n, vtg_tmpl *const *t, const vtg_pyr *const *p, const int *level,
                                 const int *x0, const int *y0);
/* The template's norm (sum of squares of its zero-mean values) and fp32 pixels (zero-mean, row-major) to host. */
int vtg_ncc_template_read(vtg_ctx *ctx, const vtg_tmpl *t, float *app_host, float *norm);
/* Search corners [x0, x1] x [y0, y1] at `step` 1 or 2 (the tracker's steps, track.c:1796, 1813, 1836; others are
 * VTG_EINVAL), clipped to where the template fits as search_slot clips, on the
 * template's level of pyramid p. cells: device buffer of at least ((x1-x0)/16+1) * ((y1-y0)/16+1) records, or NULL.
 * found: host record, written on return (this call waits for the search). */
int vtg_ncc_search(vtg_ctx *ctx, const vtg_pyr *p, const vtg_tmpl *t, int x0, int x1, int y0, int y1, int step,
                   struct vtg_cell *cells, struct vtg_found *found);
/* The same, stream-ordered: the pick goes to the device record found_dev and the call returns at once (no host
 * wait, no allocation once the context's scratch has grown to the map's size, so it can be captured in a graph). */
int vtg_ncc_search_async(vtg_ctx *ctx, const vtg_pyr *p, const vtg_tmpl *t, int x0, int x1, int y0, int y1, int step,
                         struct vtg_cell *cells, struct vtg_found *found_dev);
/* No counterpart in the reference: for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid it). vtg_ncc_search_async with every byte of working memory (the per-search
 * record, the cell map when cells is NULL) in a fixed device array of 256 KiB, one per device, looked up when a
 * context is created and held until the process ends: enough for one exhaustive 64 x 64 search of a 1920 x 1080
 * level (94 KiB), not for one of a 3840 x 2160 level. Allocates nothing, ever; a search that needs more is refused with VTG_ERANGE. The
 * caller passes the pyramid, the template and the outputs. Every call on a device shares that device's array, which
 * keeps the last call's contents: calls that overlap on one device (another context's stream, another host thread)
 * race, so order them. Results equal vtg_ncc_search_async's bit for bit. */
int vtg_ncc_search_fixed(vtg_ctx *ctx, const vtg_pyr *p, const vtg_tmpl *t, int x0, int x1, int y0, int y1, int step,
                         struct vtg_cell *cells, struct vtg_found *found_dev);
/* n searches at once, one launch per stage for each group of 8 (OP_SEARCH over several slots or streams in one
 * round, track.c:301-329 and search_slot, track.c:1661-1726): search k covers corners region[4k .. 4k+3] =
 * x0, x1, y0, y1 of pyramid p[k] with template t[k]. All templates have one size, all searched levels one element
 * type. cells: NULL, or n device maps (an entry may be NULL: the context's scratch). found_dev: n device records.
 * Stream-ordered, no host wait; a search whose window clips to nothing gets score -2 and cells 0. */
int vtg_ncc_search_many(vtg_ctx *ctx, int n, const vtg_pyr *const *p, const vtg_tmpl *const *t, const int *region,
                        int step, struct vtg_cell *const *cells, struct vtg_found *found_dev);
/* vtg_ncc_search_many run on finish_ctx, ordered after the work submitted to map_ctx so far (both contexts on one
 * device): the cell maps, top 8, refinement and pick all run on finish_ctx, beside whatever map_ctx does next (e.g.
 * LK of the same pyramids). A caller joins finish_ctx (vtg_ctx_wait) before reading found_dev or rebuilding the
 * searched pyramids on map_ctx. The working memory is finish_ctx's scratch, reused in finish_ctx's stream order.
 * finish_ctx must not be destroyed while the search is in flight. */
int vtg_ncc_search_many_split(vtg_ctx *map_ctx, vtg_ctx *finish_ctx, int n, const vtg_pyr *const *p,
                              const vtg_tmpl *const *t, const int *region, int step, struct vtg_cell *const *cells,
                              struct vtg_found *found_dev);
/* Scores at n corners (tk_correlate, track.h:57-59; half != 0: half-pixel corners, search_half): xy device int
 * pairs, out device doubles: -1 outside, 0 for a window under 1/16 of the template's energy. */
int vtg_ncc_score(vtg_ctx *ctx, const vtg_pyr *p, const vtg_tmpl *t, int n, const int *xy, int half, double *out);

/* ------------------------------------------------------------------ serving (the OP_LK / OP_SEARCH job ops of
 * tk_build / tk_track, track.c:210-429, 584-629, driven as a request loop over several streams instead of one
 * tracker's frame sequence) */

typedef struct vtg_srv vtg_srv;

struct vtg_srv_cfg {
    int device;
    size_t hbm_budget;     /* device bytes the server may hold: every stream's resident pyramids and template, every
                            * in-flight batch's frames, points and results */
    int lanes;             /* HIP streams batches are spread over, 1 .. 8 */
    int max_batch;         /* frames per batch, 1 .. 64 */
    struct lk_cfg lk;      /* the window and stopping rule of every request */
    int solve_mode;        /* VTG_SOLVE_F64 or VTG_SOLVE_MIXED */
    float fb_max;          /* > 0: flow-back check (lk_flowback) */
    int tmpl_w, tmpl_h;    /* NCC template size (level 0) */
};

struct vtg_srv_stats {
    long submitted, completed, batches, lk_points;
    int max_batch, max_inflight;  /* largest batch; most batches on the device at once */
    size_t hbm_budget, hbm_resident, hbm_in_use, hbm_peak;
};

int vtg_srv_create(const struct vtg_srv_cfg *cfg, vtg_srv **out);
int vtg_srv_destroy(vtg_srv *srv);
const char *vtg_srv_error(vtg_srv *srv);
/* A stream of w x h frames with `levels`-level pyramids and up to max_points points per frame. Its two pyramids and
 * template stay resident; VTG_ENOMEM when they do not fit the budget left. */
int vtg_srv_open(vtg_srv *srv, int w, int h, int levels, int max_points, int *stream);
/* Waits for the stream's requests, then releases its memory. */
int vtg_srv_close(vtg_srv *srv, int stream);
/* A frame of a stream (host bytes, `pitch` per row) with n points (host x, y pairs; optional guesses) tracked from the
 * stream's previous frame to this one (none for its first frame); ncc_x0 >= 0: also cut the template at (ncc_x0,
 * ncc_y0) of the previous frame and search this whole frame for it. Copies its inputs and returns a ticket at once. */
int vtg_srv_submit(vtg_srv *srv, int stream, const uint8_t *frame, int pitch, int n, const float *xy,
                   const float *guess, int ncc_x0, int ncc_y0, uint64_t *ticket);
/* 0 with the results when the ticket is done (out: room for the request's n records; *nrec: how many were written, n
 * when the frame had a previous one, else 0; found: the pick, or score -2 when no search ran), VTG_EAGAIN while it is
 * in flight. latency_ms: submit to results on the host. A ticket's results are delivered once. */
int vtg_srv_poll(vtg_srv *srv, uint64_t ticket, struct lk_out *out, int *nrec, struct vtg_found *found,
                 double *latency_ms);
int vtg_srv_wait(vtg_srv *srv, uint64_t ticket, struct lk_out *out, int *nrec, struct vtg_found *found,
                 double *latency_ms);
int vtg_srv_get_stats(vtg_srv *srv, struct vtg_srv_stats *stats);

/* --- the front door (src/common/vt_proto.h:143-197; src/card/session_track.c:76-139, 188-222): frames arrive as the
 * reference's wire format describes them, a ring of VTG_DESC_RING descriptors (struct vtg_msg = struct vt_msg)
 * published by a count, each naming a host slot holding its frame. vtg_door_ingest takes the descriptors published
 * so far and checks each before acting on it, in the reference's order (magic and check word, type / destination /
 * camera, slot, length, the camera's instance order): a bad one counts as an exception (the last kept) and its slot is
 * free at once. At most dev_slots frames are taken and unanswered (backpressure: the rest wait for the next call);
 * when the newest frame a call takes asks for VTG_PRIO_LIVE, the frames taken before it in that call are skipped.
 * Good frames go to the serving loop with their camera's points; vtg_door_result answers them in order. */
uint64_t vtg_msg_check(const struct vtg_msg *m);                       /* vt_msg_check, vt_proto.h:181-191 */
struct vtg_door_cfg {
    int cameras;              /* 1 .. 3 (vt_cameras) */
    int w[3], h[3], levels[3];/* each camera's frames (u8, pitch = w) and pyramid levels */
    int npoints[3];           /* each camera's points (xy given at creation) */
    int max_points;           /* the server streams' capacity */
    uint32_t slots;           /* the host's frame slots (vt_config.slots) */
    uint64_t slot_bytes;      /* bytes from one slot to the next */
    int dev_slots;            /* frames taken and unanswered, at most */
    int stereo;               /* camera 0's frames are stereo pairs (vt_config.stereo): two images per slot, the left
                                 one tracked */
    int max_inflight;         /* frames with the server at once, at most (0: 1); the other taken frames wait landed,
                                 where a LIVE frame can skip them */
};
struct vtg_door_state {       /* what the reference reports back in its control area (vt_ctl) */
    uint64_t frames_taken;    /* descriptors consumed: their slots are free */
    uint64_t exceptions;      /* descriptors dropped as malformed */
    uint32_t last_exception;  /* VTG_EXC_ */
    uint32_t pad;
    uint64_t skipped;         /* frames passed over for a LIVE one */
    uint64_t submitted;       /* frames handed to the serving loop */
    uint64_t in_flight;       /* of those, not yet answered */
};
struct vtg_door_result {
    uint64_t instance, acq_ns;
    uint32_t camera, slot;
    int nrec;                 /* LK records written (0 for a camera's first frame) */
    int pad;
    double latency_ms;        /* submit to results on the host */
};
typedef struct vtg_door vtg_door;
/* Opens one server stream per camera; xy[k]: camera k's npoints[k] points (x, y pairs, host). */
int vtg_door_create(vtg_srv *srv, const struct vtg_door_cfg *cfg, const float *const *xy, vtg_door **out);
int vtg_door_destroy(vtg_door *d);
int vtg_door_ingest(vtg_door *d, const struct vtg_msg *desc, uint64_t frames_pub, const uint8_t *slots,
                    struct vtg_door_state *state);
/* The oldest unanswered frame's answer (wait = 0: VTG_EAGAIN while it runs or when nothing is in flight); out: room for
 * its camera's npoints records (max_out). */
int vtg_door_result(vtg_door *d, int wait, struct vtg_door_result *r, struct lk_out *out, int max_out);

/* ------------------------------------------------------------------ reference-shaped calls over host planes */

/* img_u8_to_q6_* (src/kern/img.h:16-17): dst rows y0..y1 = 64 x src, padding of those rows filled. */
int vtg_img_u8_to_q6(vtg_ctx *ctx, const struct vt_plane *dst, const uint8_t *src, int src_stride, int y0, int y1);
/* img_reduce_* (src/kern/img.h:30-33): dst rows y0..y1 from src (whose padding need not be current: the border is
 * replicated by the kernel), padding of those rows filled. */
int vtg_img_reduce(vtg_ctx *ctx, const struct vt_plane *dst, const struct vt_plane *src, int y0, int y1);
/* img_gauss5_* (src/kern/img.h, img_ref.c:83-96): dst rows y0..y1 = the 5 x 5 binomial of src, (h + 128) >> 8 (the
 * border replicated), padding of those rows filled. */
int vtg_img_gauss5(vtg_ctx *ctx, const struct vt_plane *dst, const struct vt_plane *src, int y0, int y1);
/* img_norm_* (img_ref.c:42-55): dst rows y0..y1 = clamp(rint(src * gain_offset[0] + gain_offset[1]), 0, 16320) with the
 * reference's long double sum reproduced exactly, padding of those rows filled. */
int vtg_img_norm(vtg_ctx *ctx, const struct vt_plane *dst, const struct vt_plane *src, const float *gain_offset,
                 int y0, int y1);
/* img_minmax_* (img_ref.c:57-71): out[0], out[1] = min and max of rows y0..y1 of src (host floats). */
int vtg_img_minmax(vtg_ctx *ctx, const struct vt_plane *src, float *out, int y0, int y1);
/* lk_track_* / lk_flowback_* (src/kern/lk.h:45-53) for n points over host pyramids I[0..top], J[0..top]:
 * gx, gy may be NULL (zero guess); fb_max > 0 runs the flowback check. */
int vtg_lk_track_planes(vtg_ctx *ctx, const struct vt_plane *I, const struct vt_plane *J, int top,
                        const struct lk_cfg *cfg, int n, const float *x, const float *y, const float *gx,
                        const float *gy, float fb_max, struct lk_out *out);
/* tk_correlate (src/kern/track.h:57-59) at n corners of a host plane, template tw x th (tw a multiple of 16):
 * out[i] as the reference returns it. */
int vtg_correlate_planes(vtg_ctx *ctx, const struct vt_plane *pl, const float *app, int tw, int th, double norm,
                         int n, const int *x, const int *y, double *out);

/* --- still-cell vectors (src/kern/still.h:18-53, still.c): one camera's running-mean picture of the level it reads,
 * and per 16 x 8 cell, LK (15 x 15, one level, flowback 0.25) from that mean to each new frame. A frame is the
 * reference's four phases in order; `vtg_still_frame` runs them in one call. The frame is a level of a device pyramid
 * (any storage, read as Q6 fp32) or, in the _plane forms, a padded host plane. Cells, motion and hash match the
 * reference's: u, v within LK's tolerance of the reference's LK; the differences, decisions, counts, weights and the
 * mean bit for bit given the same vectors. */
typedef struct vtg_still vtg_still;
int vtg_still_create(vtg_ctx *ctx, const struct vtg_still_cfg *cfg, vtg_still **out); /* still_create, still.c:60-96 */
int vtg_still_destroy(vtg_ctx *ctx, vtg_still *s);
int vtg_still_first(const vtg_still *s);                                            /* still_first, still.c:97 */
int vtg_still_size(const vtg_still *s, int *cols, int *rows);
/* still_vectors (still.c:113-138): the cells' vectors, differences and sigmas against `level` of `pyr` */
int vtg_still_vectors(vtg_ctx *ctx, vtg_still *s, const vtg_pyr *pyr, int level);
int vtg_still_vectors_plane(vtg_ctx *ctx, vtg_still *s, const struct vt_plane *cur);
/* still_decide (still.c:140-237): noise median, evidence, 3 x 3 vote, counts, weights, drift; waits for the vectors */
int vtg_still_decide(vtg_ctx *ctx, vtg_still *s);
/* still_update (still.c:239-262): the weighted mean with this frame */
int vtg_still_update(vtg_ctx *ctx, vtg_still *s, const vtg_pyr *pyr, int level);
int vtg_still_update_plane(vtg_ctx *ctx, vtg_still *s, const struct vt_plane *cur);
int vtg_still_finish(vtg_ctx *ctx, vtg_still *s);                                   /* still_finish, still.c:264-270 */
/* the four phases in order (vectors skipped on the first frame, as the reference's caller does: m3_ref.c:273-277) */
int vtg_still_frame(vtg_ctx *ctx, vtg_still *s, const vtg_pyr *pyr, int level);
/* n cameras' frames at once (a camera at most once): every camera's cells in one LK launch, one wait for all cells;
 * each camera ends as after vtg_still_frame with its own level `level[i]` of pyr[i]. */
int vtg_still_frame_many(vtg_ctx *ctx, int n, vtg_still *const *s, const vtg_pyr *const *pyr, const int *level);
int vtg_still_cells(vtg_ctx *ctx, vtg_still *s, struct vtg_still_cell *out);             /* still_cells: cols x rows */
int vtg_still_mean(vtg_ctx *ctx, const vtg_still *s, const struct vt_plane *out);  /* still_mean, into a host plane */
int vtg_still_motion(const vtg_still *s, struct vtg_motion *m);                    /* still_motion, still.c:280-293 */
int vtg_still_hash(vtg_ctx *ctx, vtg_still *s, uint64_t h, uint64_t *out);         /* still_hash, still.c:272-278 */

/* --- stereo row scan (src/kern/stereo.c:36-80 stereo_scan_vx, called per feature by the tracker's stereo step,
 * track.c:1866-1895): a 16 x 11 zero-mean template of `level` of the left pyramid at each feature against integer
 * disparities dmin .. dmax (at most 512) of the right pyramid's row. n features in one launch; xy (2 floats a feature)
 * and out are device buffers. The _planes form takes the reference's host planes and host arrays. */
int vtg_stereo_scan_many(vtg_ctx *ctx, const vtg_pyr *left, const vtg_pyr *right, int level, int n, const float *xy,
                         int dmin, int dmax, struct vtg_stereo_scan *out);
int vtg_stereo_scan_planes(vtg_ctx *ctx, const struct vt_plane *left, const struct vt_plane *right, int n,
                           const float *x, const float *y, int dmin, int dmax, struct vtg_stereo_scan *out);
/* The tracker's stereo step (stereo_one, track.c:1859-1914) for n features (a host array, read and written): the scan
 * at cfg->stereo_level for features without a prior disparity, LK from the left to the right pyramid with the disparity
 * as x guess (flowback from level 0), the match, and the 3D point where it triangulates and reprojects within
 * max_reproj. Waits for its results. */
int vtg_stereo_match(vtg_ctx *ctx, const vtg_pyr *left, const vtg_pyr *right, const struct vtg_stereo_cfg *cfg, int n,
                     struct vtg_stereo_feature *feat);

#ifdef __cplusplus
}
#endif

#endif
