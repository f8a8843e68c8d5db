/* vt_oracle.h: TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A CPU restatement of the reference's pyramid, sparse pyramidal LK and NCC search, written from the reference's
 * semantics. Only the tests, the smoke check and bench.py's cpu_baseline leg load it; the GPU library never does.
 *
 * Tiers:
 *   lane  (reference-faithful)  fp32 patch arithmetic in the reference's 16-lane order (fmaf per lane, the
 *                               i^1,i^2,i^4,i^8 horizontal sum), long double driver: reproduces the reference's
 *                               committed hash lines bit for bit (src/kern/lk_vx.c, src/common/vx.h:158-252).
 *   lane64                      the lane tier's fp32 patch arithmetic with the driver in double instead of x87
 *                               long double: the arithmetic of the GPU's LK. Equal to the lane tier except on the
 *                               rare records where a 53-bit driver step rounds a position differently.
 *   f64   (FP64 oracle)         every patch and driver operation in double: the yardstick each flow result's
 *                               error is reported against. The reference has no such tier.
 *   mt    (threaded baseline)   pthreads over rows (pyramid) and over points / cells (LK, NCC) for timing.
 *
 * Pinning of each function is listed beside it in vt_oracle.c. */
#ifndef VT_ORACLE_H
#define VT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/vt_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* --- helpers (src/common/vt_util.h:29-50) --- */
uint64_t vo_fnv(uint64_t h, const void *p, size_t n);
uint64_t vo_rand(uint64_t *s);
uint64_t vo_hash_plane(const struct vt_plane *pl);
uint64_t vo_hash_lk(const struct lk_out *o, int n);

/* --- deterministic gate inputs (src/kern/m1_input.h:10-21, src/kern/m2_input.h:13-58) --- */
void vo_m1_image(uint8_t *img, int w, int h, int stride);
void vo_m2_render(uint8_t *img, int w, int h, int stride, int ox, int oy);

/* --- planes: memory is plane_bytes(w, h), any alignment of 64 B --- */
struct vt_plane vo_plane_new(int w, int h); /* zeroed, free with vo_plane_free */
void vo_plane_free(struct vt_plane *pl);

/* --- image primitives (src/kern/img_ref.c) --- */
void vo_pad_rows(const struct vt_plane *pl, int y0, int y1);
void vo_pad_tb(const struct vt_plane *pl);
void vo_u8_to_q6(const struct vt_plane *dst, const uint8_t *src, int src_stride, int y0, int y1);
void vo_reduce(const struct vt_plane *dst, const struct vt_plane *src, int y0, int y1);
void vo_gauss5(const struct vt_plane *dst, const struct vt_plane *src, int y0, int y1);
void vo_sobel(const struct vt_plane *ix, const struct vt_plane *iy, const struct vt_plane *src, int y0, int y1);
void vo_absdiff(const struct vt_plane *dst, const struct vt_plane *a, const struct vt_plane *b, int y0, int y1);
void vo_norm(const struct vt_plane *dst, const struct vt_plane *src, const float *gain_offset, int y0, int y1);
void vo_minmax(const struct vt_plane *src, float *out, int y0, int y1);

/* A whole pyramid: u8 frame -> levels[0..levels-1], each padded. */
void vo_pyr_build(struct vt_plane *lv, const uint8_t *img, int stride, int levels);
void vo_pyr_build_mt(struct vt_plane *lv, const uint8_t *img, int stride, int levels, int threads);

/* --- LK (src/kern/lk_driver.inc, lk_vx.c, lk_ref.c) --- */
/* VO_LANE64_SCHARR and VO_F64_SCHARR: the tiers VO_LANE64 and VO_F64 with the Scharr gradient variant in place of
 * the reference's Sobel (no counterpart in the reference; the GPU's VTG_GRAD_SCHARR): [3 10 3] across, [-1 0 1] along,
 * / 32 per grey level, fp32 as fmaf(centre difference, 10, 3 (outer differences' sum)) / 2048 in Q6 units. */
enum { VO_LANE = 0, VO_F64 = 1, VO_LANE64 = 2, VO_LANE64_SCHARR = 3, VO_F64_SCHARR = 4 };
void vo_lk_track(int tier, const struct vt_plane *I, const struct vt_plane *J, int top, const struct lk_cfg *c,
                 float x, float y, float gx, float gy, struct lk_out *o);
void vo_lk_flowback(int tier, const struct vt_plane *I, const struct vt_plane *J, int top, const struct lk_cfg *c,
                    float x, float y, float max_fb, struct lk_out *o);
/* n points, guess may be NULL (zero); fb_max > 0 runs the flowback check; threads >= 1. */
void vo_lk_batch(int tier, const struct vt_plane *I, const struct vt_plane *J, int top, const struct lk_cfg *c, int n,
                 const float *x, const float *y, const float *gx, const float *gy, float fb_max, struct lk_out *o,
                 int threads);

/* --- NCC search (src/kern/track.c:184-208, 301-329, 935-974, 1615-1734), template tw x th, tw a multiple of 16 --- */
/* Zero-mean template from the tw x th window at corner (x0, y0): 0 on success, -1 if the window does not fit. */
int vo_ncc_template(const struct vt_plane *pl, int x0, int y0, int tw, int th, float *app, float *norm);
long double vo_ncc_at(const struct vt_plane *pl, const float *app, int tw, int th, long double norm, int x, int y);
long double vo_ncc_half(const struct vt_plane *pl, const float *app, int tw, int th, long double norm, int hx, int hy);
long double vo_peak_offset(long double l, long double c, long double r);

/* The per-cell map of one search: cells of 16 x 16 corners over [x0, x1] x [y0, y1] (already clipped), step. */
struct vo_cell { float v; int x, y; };
void vo_ncc_cells(const struct vt_plane *pl, const float *app, int tw, int th, float norm, int x0, int x1, int y0,
                  int y1, int step, struct vo_cell *cells, int threads);

/* search_slot: clip, cell map, top 8, half-pixel refinement, second peak. */
struct vo_found { int hx, hy; long double ncc, second; int cells; };
void vo_ncc_search(const struct vt_plane *pl, const float *app, int tw, int th, float norm, int x0, int x1, int y0,
                   int y1, int step, struct vo_found *f, struct vo_cell *cells_out, int threads);

/* --- still-cell vectors (src/kern/still.c): the state held explicitly so a test can set it from the GPU's --- */
struct vo_still {
    int w, h, cols, rows, ncells, avg;
    float motion_px, motion_grey;
    struct vt_plane mean;    /* padded, zeroed at creation */
    struct vtg_still_cell *cell;   /* ncells */
    uint8_t *raw;            /* (cols + 2) x (rows + 2): this frame's own evidence, a zero border */
    uint16_t *count;         /* frames in a cell's mean */
    float *weight;           /* this frame's weight of a cell */
    uint64_t frames;
    uint32_t tracked, moving;
    float count_mean, drift_u, drift_v, drift_rms, diff_median, diff_limit;
    long double noise_usual;
    int lk_tier;             /* the LK tier of still_vectors: VO_LANE (the reference's x87 driver) at creation */
};
struct vo_still *vo_still_new(const struct vtg_still_cfg *c);
void vo_still_free(struct vo_still *s);
struct lk_cfg vo_still_lk_cfg(void);
void vo_still_centre(const struct vo_still *s, int i, float *x, float *y);
float vo_still_cell_diff(const struct vo_still *s, const struct vt_plane *cur, int col, int row);
void vo_still_vectors(struct vo_still *s, const struct vt_plane *cur, int threads);
void vo_still_decide(struct vo_still *s);
void vo_still_update(struct vo_still *s, const struct vt_plane *cur);
void vo_still_finish(struct vo_still *s);
uint64_t vo_still_hash(const struct vo_still *s, uint64_t h);

/* --- stereo row scan (src/kern/stereo.h:9-20): a 16 x 11 zero-mean template of the left level against integer
 * disparities dmin .. dmax of the right row --- */
struct vo_stereo_scan { float disparity, ncc, second; }; /* = struct stereo_scan */
void vo_stereo_scan(const struct vt_plane *left, const struct vt_plane *right, float x, float y, int dmin, int dmax,
                    struct vo_stereo_scan *out);

int vo_stereo_triangulate(const struct vtg_stereo_geom *g, long double xl, long double yl, long double xr,
                          long double yr, long double *p);
long double vo_stereo_reprojection(const struct vtg_stereo_geom *g, const long double *p, long double xl,
                                   long double yl, long double xr, long double yr);
/* the tracker's stereo step over n features; L, R: the two pyramids' levels 0 .. max(stereo_level, 0) */
void vo_stereo_match(const struct vt_plane *L, const struct vt_plane *R, const struct vtg_stereo_cfg *c, int n,
                     struct vtg_stereo_feature *f);
/* the same with the LK of the given tier (VO_LANE in vo_stereo_match) */
void vo_stereo_match_tier(int tier, const struct vt_plane *L, const struct vt_plane *R, const struct vtg_stereo_cfg *c,
                          int n, struct vtg_stereo_feature *f);

#ifdef __cplusplus
}
#endif

#endif
