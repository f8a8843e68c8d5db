/* vtg_track.h: the tracker's job-op layer on the GPU, behind the reference's tracker-level surface (src/kern/track.h).
 *
 * What a frame's tracking does to the path (track.c:584-629 tk_build; track.c:2072-2118 tk_track; the job ops
 * OP_U8 / OP_REDUCE / OP_LK / OP_SEARCH, track.c:210-429): a frame's pyramid goes into pyramid set k mod the set count;
 * the next frame's tracking runs LK of the feature list from the previous set to the current one, every feature
 * starting from one guess, the head's rate times the time since the last frame (track.c:2113-2114), with the
 * forward-backward check when fb_max > 0, drops the features LK lost (track.c:1305), and searches a stored
 * appearance when one is armed (search_slot, track.c:1661-1726).
 *
 * Kept as the reference has it: the injected executor and allocator (track.h:12-16: tk_fn, tk_run_fn, tk_alloc_fn,
 * the same types), the two halves (build, then track; frame k reads sets k and k - 1), the set count, tk_level, and
 * the per-feature output record (struct vtg_feat_out has struct vt_feat_out's layout, vt_proto.h:332-340). The region
 * state machine, corner seeding, stereo, the Kalman filter and frame generation are not here: the caller owns the
 * feature list (vtg_tk_set_features, vtg_tk_add_features) and decides what to track and when to search.
 *
 * Errors: create returns NULL with the reason on stderr (as tk_create, track.c:452-454); the other calls return 0 or
 * a negative VTG_E* code (vtgpu.h), the reason in vtg_tk_error. One host thread per tracker. */
#ifndef VTG_TRACK_H
#define VTG_TRACK_H

#include <stddef.h>
#include <stdint.h>

#include "vtgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The reference's executor and allocator types (track.h:12-16): run fn on every thread of a team, or once with
 * (arg, 0, 1); memory 64 B aligned, never freed by the tracker. */
typedef void (*tk_fn)(void *arg, int tid, int n);
typedef void (*tk_run_fn)(void *ctx, tk_fn fn, void *arg);
typedef void *(*tk_alloc_fn)(void *ctx, size_t bytes);

#define VTG_TK_MAX_FEATURES 512 /* VT_MAX_FEATURES, vt_proto.h:60 */

/* The fields of struct vt_config (vt_proto.h:67-99) this layer acts on, same names and meanings, and where the GPU
 * needs more: the pyramid sets, the level storage and the device. */
struct vtg_tk_config {
    uint32_t width, height;      /* u8 frames, one camera */
    uint32_t levels;             /* pyramid levels, 1 = the image alone */
    uint32_t patch_w, patch_h;   /* the LK window */
    uint32_t max_iter;
    float eps;                   /* LK stop, pixels */
    float min_eig;               /* smaller eigenvalue per patch pixel */
    float max_residual;          /* mean |T - J| in 8-bit grey above which a feature is dropped */
    float fb_max;                /* forward-backward miss in pixels above which a feature is dropped; 0 = no check */
    uint32_t max_features;       /* at most VTG_TK_MAX_FEATURES */
    uint32_t sets;               /* pyramid sets in the ring: 2 .. 4 (the reference allocates 2, 3 or 4) */
    int storage;                 /* VTG_STORE_* of the levels */
    int device;
};

/* = struct vt_feat_out (vt_proto.h:332-340). This layer fills x, y, u, v, sigma and flags; the stereo fields are 0. */
struct vtg_feat_out {
    float x, y;       /* left image */
    float u, v;       /* displacement this frame */
    float xr, yr;     /* right image */
    float disparity;
    float sigma;      /* sqrt of the larger diagonal entry of LK's covariance at level 0, pixels */
    uint32_t flags;   /* VTG_FF_NEW: not tracked this frame (no previous set, or no flow asked): no displacement */
    float z;
};

/* An appearance search's answer (search_slot, track.c:1661-1726): the best corner in half pixels of the appearance's
 * level, its correlation and the best at least half a template away. */
struct vtg_tk_found {
    int armed;               /* a search ran this frame */
    int hx, hy;
    double ncc, second;
    float x, y;              /* the match's centre in level-0 pixels */
};

/* One frame's result. */
struct vtg_tk_result {
    uint64_t instance;       /* the frame's message instance */
    uint32_t nfeat;          /* features in feat[] */
    uint32_t tracked;        /* of them, tracked from the previous frame (the rest are new) */
    uint32_t dropped;        /* features LK lost this frame */
    float guess_u, guess_v;  /* the start every feature had: the head's rate times the time since the last frame */
    double rate_u, rate_v;   /* the head's rate after this frame, pixels per second: the median displacement / dt */
    struct vtg_tk_found search;
    struct vtg_feat_out feat[VTG_TK_MAX_FEATURES];
};

struct vtg_tracker;

/* tk_create (track.h:22): everything the config asks for, allocated now (device memory on the config's device, host
 * bookkeeping through alloc). threads_pyr / threads_track size the executor teams; run_pyr runs the build's host
 * side, run_track the per-feature host side of tracking. NULL when it cannot be done, the reason on stderr. */
struct vtg_tracker *vtg_tk_create(const struct vtg_tk_config *cfg, int threads_pyr, int threads_track,
                                  tk_alloc_fn alloc, void *alloc_ctx, tk_run_fn run_pyr, void *pyr_ctx,
                                  tk_run_fn run_track, void *track_ctx);
void vtg_tk_destroy(struct vtg_tracker *t);
/* The reason for the last failed call. */
const char *vtg_tk_error(const struct vtg_tracker *t);
/* tk_sets (track.h:25): how many pyramid sets exist; frame k lives in set k mod that count. */
int vtg_tk_sets(const struct vtg_tracker *t);
/* tk_build (track.h:27): the frame's pyramid into set `set` (OP_U8 + OP_REDUCE). left: width-stride bytes per row
 * (host memory). Reads only the message's build keys (none acted on here). Returns when the frame has been copied;
 * the build itself is stream-ordered before the next tracking. */
int vtg_tk_build(struct vtg_tracker *t, int set, const uint8_t *left, const struct vtg_msg *m);
/* tk_track (track.h:29): the rest of a frame whose pyramid is in `set`: OP_LK of the feature list from the previous
 * built set (when the message's dest holds VTG_PIPE_FLOW and there is one) with the head-rate guess, then the armed
 * search. out is complete on return. */
int vtg_tk_track(struct vtg_tracker *t, int set, const struct vtg_msg *m, struct vtg_tk_result *out);
/* The feature list: replace it, or add to it (positions in level-0 pixels of the last tracked frame); the next
 * vtg_tk_track tracks them with the rest. At most max_features in all (VTG_ERANGE beyond). */
int vtg_tk_set_features(struct vtg_tracker *t, int n, const float *x, const float *y);
int vtg_tk_add_features(struct vtg_tracker *t, int n, const float *x, const float *y);
/* The head's rate, pixels per second (the reference keeps it from the region's flow; a caller may set it). */
int vtg_tk_set_rate(struct vtg_tracker *t, double rate_u, double rate_v);
/* store_appearance (track.c:935-974): the tw x th template of set `set`'s level `level` with corner (x0, y0) of that
 * level becomes the appearance; every later vtg_tk_track searches it in its set over corners [cx - r, cx + r] x
 * [cy - r, cy + r] of that level (step 1) until disarmed (r < 0). */
int vtg_tk_arm_search(struct vtg_tracker *t, int set, int level, int x0, int y0, int tw, int th, int cx, int cy,
                      int r);
/* tk_level (track.h:66-67): one level of a set's pyramid as the reference's host plane (fp32 Q6, padded), valid
 * until vtg_tk_build fills that set again or the next vtg_tk_level call. NULL on error. */
const struct vt_plane *vtg_tk_level(struct vtg_tracker *t, int set, int level);
/* The tracker's device context (for callers that order their own work after it). */
vtg_ctx *vtg_tk_ctx(struct vtg_tracker *t);

#ifdef __cplusplus
}
#endif

#endif
