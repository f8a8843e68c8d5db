/* vt_oracle.c: TEST INFRASTRUCTURE, NOT PRODUCT CODE. See vt_oracle.h for the tiers.
 *
 * Only tests/, the smoke check and bench.py's cpu_baseline leg load this file's library. It restates the reference's
 * semantics in plain C; it never runs the reference. What pins each part (checks in tests/test_oracle.py):
 *   pyramid / image primitives   the reference-held hash lines logs/m1/host/m1-host-{640x480,1280x720,1920x1080}.txt
 *                                (u8_to_q6, norm, minmax, gauss5, reduce, sobel_x/y, absdiff) and the closed form
 *                                (a constant plane stays constant; an integer brute-force 5x5 on tiny planes).
 *   LK lane tier                 logs/m2/host/640x480.txt: ten displacements x (point 0's raw bits + FNV of all 330
 *                                lk_out), multi-level pyramids included; plus oracle/_ref on the seeded blob-scene
 *                                frames where that build exists (tests/golden/lk_ref.json).
 *   LK f64 tier                  agreement with the lane tier within 1e-3 px on converged points, and exact
 *                                translation truth on the m2 inputs (|d| <= 14.4 px: EPE within the reference's
 *                                measured 0.0209 px p95).
 *   LK Scharr tiers              the closed form of G's smaller eigenvalue on the plane x y^2 + x^2 y (whose exact
 *                                gradients differ from both operators' by their smoothing's second moment: 1/2 for
 *                                Sobel, 3/8 for Scharr), the lane64 and f64 Scharr tiers within 1e-3 px of each
 *                                other, and exact translation truth on the m2 inputs.
 *   NCC                          golden vectors from oracle/_ref's own tk_correlate (32 x 21, tests/golden/ncc_ref.json),
 *                                the invariants of src/host/vt_ncc_check.c, closed forms (affine copies score +-1),
 *                                and a textbook double ZNCC by brute force on small windows.
 *
 * Compiled with -ffp-contract=off: every fused multiply-add below is an explicit fmaf, as in the reference's lane model
 * (src/common/vx.h:180); long double is x87 80-bit on x86-64, as in the reference's driver (src/kern/lk.h:4-6). */
#include "vt_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define NL 16 /* lanes of the reference's vector model */

/* ------------------------------------------------------------------ helpers */

/* FNV-1a 64 over bytes: src/common/vt_util.h:29-39. */
uint64_t vo_fnv(uint64_t h, const void *p, size_t n)
{
    const uint8_t *b = (const uint8_t *)p;
    for (size_t i = 0; i < n; i++) {
        h ^= b[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

/* xorshift64*: src/common/vt_util.h:42-50. */
uint64_t vo_rand(uint64_t *s)
{
    uint64_t x = *s;
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    *s = x;
    return x * 0x2545F4914F6CDD1Dull;
}

#define FNV0 0xcbf29ce484222325ull

/* Hash of the w visible floats of every row: src/host/m1_ref.c:21-27. */
uint64_t vo_hash_plane(const struct vt_plane *pl)
{
    uint64_t h = FNV0;
    for (int y = 0; y < pl->h; y++)
        h = vo_fnv(h, plane_row(pl, y), (size_t)pl->w * sizeof(float));
    return h;
}

/* FNV over the raw 32-byte records: src/host/m2_ref.c:69-73. */
uint64_t vo_hash_lk(const struct lk_out *o, int n)
{
    uint64_t h = FNV0;
    for (int i = 0; i < n; i++)
        h = vo_fnv(h, &o[i], sizeof o[i]);
    return h;
}

/* The horizontal sum of the lane model: pairs i^1, i^2, i^4, i^8, every lane updated each step (vx.h:213-221). */
static float lane_hsum(const float *v)
{
    float a[NL], b[NL];
    memcpy(a, v, sizeof a);
    for (int step = 1; step < NL; step <<= 1) {
        for (int i = 0; i < NL; i++)
            b[i] = a[i] + a[i ^ step];
        memcpy(a, b, sizeof a);
    }
    return a[0];
}

/* ------------------------------------------------------------------ gate inputs */

/* m1 image: ramp + blocks + 5 bits of xorshift noise (src/kern/m1_input.h:10-21). */
void vo_m1_image(uint8_t *img, int w, int h, int stride)
{
    uint64_t s = 0x9e3779b97f4a7c15ull;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const int ramp = (x * 96) / w + (y * 64) / h;
            const int block = (((x >> 5) ^ (y >> 5)) & 7) * 8;
            const int noise = (int)(vo_rand(&s) & 31);
            img[(size_t)y * (size_t)stride + (size_t)x] = (uint8_t)(ramp + block + noise);
        }
}

/* m2 image: integer value noise on a 4x fine grid, box-averaged (src/kern/m2_input.h:13-50). */
static uint32_t m2_mix(int x, int y)
{
    uint32_t h = (uint32_t)x * 0x9E3779B1u ^ (uint32_t)y * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0xC2B2AE3Du;
    return h ^ (h >> 13);
}

static uint32_t m2_value(int x, int y, int bits)
{
    const int xi = x >> bits, yi = y >> bits, m = (1 << bits) - 1;
    const uint32_t fx = (uint32_t)(x & m), fy = (uint32_t)(y & m), one = 1u << bits;
    const uint32_t a = m2_mix(xi, yi) & 255, b = m2_mix(xi + 1, yi) & 255;
    const uint32_t c = m2_mix(xi, yi + 1) & 255, d = m2_mix(xi + 1, yi + 1) & 255;
    return (a * (one - fx) + b * fx) * (one - fy) + (c * (one - fx) + d * fx) * fy;
}

void vo_m2_render(uint8_t *img, int w, int h, int stride, int ox, int oy)
{
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            uint32_t sum = 0;
            for (int j = 0; j < 4; j++)
                for (int i = 0; i < 4; i++) {
                    const int fx = 4 * x + i + ox, fy = 4 * y + j + oy;
                    const uint32_t coarse = m2_value(fx, fy, 5) >> 6, fine = m2_value(fx, fy, 4) >> 4;
                    sum += (coarse * 3 + fine) >> 2;
                }
            img[(long)y * stride + x] = (uint8_t)((sum + 128) >> 8);
        }
}

/* ------------------------------------------------------------------ planes */

struct vt_plane vo_plane_new(int w, int h)
{
    const size_t bytes = (plane_bytes(w, h) + 63) & ~(size_t)63;
    void *m = aligned_alloc(64, bytes);
    if (!m)
        abort();
    memset(m, 0, bytes);
    return plane_at(m, w, h);
}

void vo_plane_free(struct vt_plane *pl)
{
    if (pl->p)
        free(pl->p - (size_t)pl->stride * VT_PAD_Y - VT_PAD_X);
    pl->p = NULL;
}

/* ------------------------------------------------------------------ image primitives */

/* Replicated left/right border of rows y0..y1: 32 lanes left, out to round16(w) + 32 right (img_ref.c:9-20). */
void vo_pad_rows(const struct vt_plane *pl, int y0, int y1)
{
    const int right_end = ((pl->w + 15) & ~15) + VT_PAD_X;
    for (int y = y0; y < y1; y++) {
        float *r = plane_row(pl, y);
        const float first = r[0], last = r[pl->w - 1];
        for (int x = -VT_PAD_X; x < 0; x++)
            r[x] = first;
        for (int x = pl->w; x < right_end; x++)
            r[x] = last;
    }
}

/* The 4 rows above and below, whole stride, copied from the edge rows (img_ref.c:22-30). */
void vo_pad_tb(const struct vt_plane *pl)
{
    const size_t row_bytes = (size_t)pl->stride * sizeof(float);
    for (int k = 1; k <= VT_PAD_Y; k++) {
        memcpy(plane_row(pl, -k) - VT_PAD_X, plane_row(pl, 0) - VT_PAD_X, row_bytes);
        memcpy(plane_row(pl, pl->h - 1 + k) - VT_PAD_X, plane_row(pl, pl->h - 1) - VT_PAD_X, row_bytes);
    }
}

/* L0 = 64 x byte (img_ref.c:32-40). */
void vo_u8_to_q6(const struct vt_plane *dst, const uint8_t *src, int src_stride, int y0, int y1)
{
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < dst->w; x++)
            plane_row(dst, y)[x] = (float)(src[(ptrdiff_t)y * src_stride + x] * VT_Q6);
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

/* Integer value of a Q6 pixel with the replicated border (the reference reads its padding; with current padding
 * that is the clamp). */
static inline int32_t q6_at(const struct vt_plane *pl, int x, int y)
{
    return (int32_t)plane_row(pl, clampi(y, 0, pl->h - 1))[clampi(x, 0, pl->w - 1)];
}

static const int32_t binom5[5] = {1, 4, 6, 4, 1};

/* 5x5 binomial sum at (cx, cy): exact integer, at most 4,177,920. */
static int32_t binom_sum(const struct vt_plane *src, int cx, int cy)
{
    int32_t acc = 0;
    for (int j = 0; j < 5; j++) {
        int32_t row = 0;
        for (int i = 0; i < 5; i++)
            row += binom5[i] * q6_at(src, cx + i - 2, cy + j - 2);
        acc += binom5[j] * row;
    }
    return acc;
}

/* dst[y][x] = (sum_{j,i} b_j b_i src[2y+j-2][2x+i-2] + 128) >> 8, one rounding (img_ref.c:98-114). */
void vo_reduce(const struct vt_plane *dst, const struct vt_plane *src, int y0, int y1)
{
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < dst->w; x++)
            plane_row(dst, y)[x] = (float)((binom_sum(src, 2 * x, 2 * y) + 128) >> 8);
}

/* The same filter at every pixel (img_ref.c:83-96). */
void vo_gauss5(const struct vt_plane *dst, const struct vt_plane *src, int y0, int y1)
{
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < dst->w; x++)
            plane_row(dst, y)[x] = (float)((binom_sum(src, x, y) + 128) >> 8);
}

/* [-1 0 1] x [1 2 1] and its transpose, exact integers (img_ref.c:116-127). */
void vo_sobel(const struct vt_plane *ix, const struct vt_plane *iy, const struct vt_plane *src, int y0, int y1)
{
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < src->w; x++) {
            const int32_t a = q6_at(src, x - 1, y - 1), b = q6_at(src, x, y - 1), c = q6_at(src, x + 1, y - 1);
            const int32_t e = q6_at(src, x - 1, y), g = q6_at(src, x + 1, y);
            const int32_t m = q6_at(src, x - 1, y + 1), n = q6_at(src, x, y + 1), o = q6_at(src, x + 1, y + 1);
            plane_row(ix, y)[x] = (float)((c + 2 * g + o) - (a + 2 * e + m));
            plane_row(iy, y)[x] = (float)((m + 2 * n + o) - (a + 2 * b + c));
        }
}

void vo_absdiff(const struct vt_plane *dst, const struct vt_plane *a, const struct vt_plane *b, int y0, int y1)
{
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < dst->w; x++)
            plane_row(dst, y)[x] = fabsf(plane_row(a, y)[x] - plane_row(b, y)[x]);
}

/* clamp(rint(s * g + o), 0, 16320): exact product and one 64-bit sum in long double, then float (img_ref.c:42-55). */
void vo_norm(const struct vt_plane *dst, const struct vt_plane *src, const float *gain_offset, int y0, int y1)
{
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < dst->w; x++) {
            float v = rintf((float)((long double)plane_row(src, y)[x] * gain_offset[0] + gain_offset[1]));
            plane_row(dst, y)[x] = v < 0.0f ? 0.0f : v > 16320.0f ? 16320.0f : v;
        }
}

void vo_minmax(const struct vt_plane *src, float *out, int y0, int y1)
{
    float lo = plane_row(src, y0)[0], hi = lo;
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < src->w; x++) {
            const float v = plane_row(src, y)[x];
            lo = v < lo ? v : lo;
            hi = v > hi ? v : hi;
        }
    out[0] = lo;
    out[1] = hi;
}

/* ------------------------------------------------------------------ threads */

typedef void (*vo_job_fn)(void *arg, int tid, int n);
struct vo_job_arg { vo_job_fn fn; void *arg; int tid, n; };

static void *vo_thread_main(void *p)
{
    struct vo_job_arg *a = (struct vo_job_arg *)p;
    a->fn(a->arg, a->tid, a->n);
    return NULL;
}

static void vo_run(vo_job_fn fn, void *arg, int threads)
{
    if (threads <= 1) {
        fn(arg, 0, 1);
        return;
    }
    if (threads > 512)
        threads = 512;
    pthread_t th[512];
    struct vo_job_arg a[512];
    for (int i = 0; i < threads; i++) {
        a[i] = (struct vo_job_arg){fn, arg, i, threads};
        if (i > 0)
            pthread_create(&th[i], NULL, vo_thread_main, &a[i]);
    }
    fn(arg, 0, threads);
    for (int i = 1; i < threads; i++)
        pthread_join(th[i], NULL);
}

/* ------------------------------------------------------------------ pyramid */

struct pyr_job { struct vt_plane *lv; const uint8_t *img; int stride, level; };

static void pyr_rows(void *p, int tid, int n)
{
    struct pyr_job *j = (struct pyr_job *)p;
    const struct vt_plane *d = &j->lv[j->level];
    const int y0 = (int)((long)d->h * tid / n), y1 = (int)((long)d->h * (tid + 1) / n);
    if (j->level == 0)
        vo_u8_to_q6(d, j->img, j->stride, y0, y1);
    else
        vo_reduce(d, &j->lv[j->level - 1], y0, y1);
    vo_pad_rows(d, y0, y1);
}

void vo_pyr_build_mt(struct vt_plane *lv, const uint8_t *img, int stride, int levels, int threads)
{
    for (int L = 0; L < levels; L++) {
        struct pyr_job j = {lv, img, stride, L};
        vo_run(pyr_rows, &j, threads);
        vo_pad_tb(&lv[L]);
    }
}

void vo_pyr_build(struct vt_plane *lv, const uint8_t *img, int stride, int levels)
{
    vo_pyr_build_mt(lv, img, stride, levels, 1);
}

/* ------------------------------------------------------------------ LK, lane tier */

#define TILE_COLS 48          /* LK_MAX_W rounded to whole 16-lane vectors */
#define S_COLS (TILE_COLS + 32)

struct lane_tile {
    float t[LK_MAX_H][TILE_COLS], tx[LK_MAX_H][TILE_COLS], ty[LK_MAX_H][TILE_COLS];
};

static const long double pow2m_ld[8] = {1.0L, 0.5L, 0.25L, 0.125L, 0.0625L, 0.03125L, 0.015625L, 0.0078125L};

/* floor by round-to-nearest and a fix-up (lk_common.h:24-29). */
static long floor_ld(long double v)
{
    const long i = lrintl(v);
    return i - ((long double)i > v);
}

/* The template window at (px, py): bilinear S over rows -1..H and 16-lane blocks from one block left of the window
 * to one right of it; Sobel / 512 on S (or the Scharr variant, / 2048); G summed per lane in row-then-block order
 * (lk_vx.c:36-97). Sobel's outer weight is 1, so its product is exact and the arithmetic is the reference's. */
static void lane_build(const struct vt_plane *pl, long double px, long double py, const struct lk_cfg *c, int scharr,
                       struct lane_tile *t, float *g)
{
    const float we = scharr ? 3.0f : 1.0f, wc = scharr ? 10.0f : 2.0f;
    const float gscale = scharr ? 1.0f / 2048.0f : 1.0f / 512.0f;
    const long ix = floor_ld(px), iy = floor_ld(py);
    const float fx = (float)(px - (long double)ix), fy = (float)(py - (long double)iy);
    const int W = c->w, H = c->h, nv = (W + 15) / 16, ncol = 16 * (nv + 2);
    const long xs = ix - W / 2 - 16, ytop = iy - H / 2;
    float s[LK_MAX_H + 2][S_COLS], above[S_COLS];
    {
        const float *a = plane_row(pl, (int)(ytop - 1)) + xs;
        for (int q = 0; q < ncol; q++)
            above[q] = fmaf(a[q + 1] - a[q], fx, a[q]);
    }
    for (int r = 0; r < H + 2; r++) {
        const float *a = plane_row(pl, (int)(ytop + r)) + xs;
        for (int q = 0; q < ncol; q++) {
            const float here = fmaf(a[q + 1] - a[q], fx, a[q]);
            s[r][q] = fmaf(here - above[q], fy, above[q]);
            above[q] = here;
        }
    }
    const int tail = W & 15;
    float gxx[NL] = {0}, gxy[NL] = {0}, gyy[NL] = {0};
    for (int r = 0; r < H; r++) {
        const float *u = s[r], *m = s[r + 1], *d = s[r + 2];
        for (int k = 0; k < nv; k++)
            for (int j = 0; j < NL; j++) {
                const int q = 16 * (k + 1) + j;
                const float du = u[q + 1] - u[q - 1], dm = m[q + 1] - m[q - 1], dd = d[q + 1] - d[q - 1];
                float tx = fmaf(dm, wc, (du + dd) * we) * gscale;
                const float vl = d[q - 1] - u[q - 1], vc = d[q] - u[q], vr = d[q + 1] - u[q + 1];
                float ty = fmaf(vc, wc, (vl + vr) * we) * gscale;
                if (k == nv - 1 && tail && j >= tail)
                    tx = ty = 0.0f;
                t->t[r][16 * k + j] = m[q];
                t->tx[r][16 * k + j] = tx;
                t->ty[r][16 * k + j] = ty;
                gxx[j] = fmaf(tx, tx, gxx[j]);
                gxy[j] = fmaf(tx, ty, gxy[j]);
                gyy[j] = fmaf(ty, ty, gyy[j]);
            }
    }
    g[0] = lane_hsum(gxx);
    g[1] = lane_hsum(gxy);
    g[2] = lane_hsum(gyy);
}

/* J bilinear at (qx, qy) with the x-interpolated row carried down; b = (sum It Tx, sum It Ty) (lk_vx.c:100-140). */
static void lane_iterate(const struct vt_plane *pl, long double qx, long double qy, const struct lk_cfg *c,
                         const struct lane_tile *t, float *b)
{
    const long jx = floor_ld(qx), jy = floor_ld(qy);
    const float fx = (float)(qx - (long double)jx), fy = (float)(qy - (long double)jy);
    const int nv = (c->w + 15) / 16, ncol = 16 * nv;
    const long x0 = jx - c->w / 2, ytop = jy - c->h / 2;
    float above[TILE_COLS], bx[NL] = {0}, by[NL] = {0};
    const float *a = plane_row(pl, (int)ytop) + x0;
    for (int q = 0; q < ncol; q++)
        above[q] = fmaf(a[q + 1] - a[q], fx, a[q]);
    for (int r = 0; r < c->h; r++) {
        a += pl->stride;
        for (int k = 0; k < nv; k++)
            for (int j = 0; j < NL; j++) {
                const int q = 16 * k + j;
                const float here = fmaf(a[q + 1] - a[q], fx, a[q]);
                const float it = t->t[r][q] - fmaf(here - above[q], fy, above[q]);
                bx[j] = fmaf(it, t->tx[r][q], bx[j]);
                by[j] = fmaf(it, t->ty[r][q], by[j]);
                above[q] = here;
            }
    }
    b[0] = lane_hsum(bx);
    b[1] = lane_hsum(by);
}

/* mean |It| and mean It^2 in 8-bit grey, lanes past the window masked, long double divide (lk_vx.c:142-175). */
static void lane_residual(const struct vt_plane *pl, long double qx, long double qy, const struct lk_cfg *c,
                          const struct lane_tile *t, float *out)
{
    static const float q6inv = 1.0f / 64.0f;
    const long jx = floor_ld(qx), jy = floor_ld(qy);
    const float fx = (float)(qx - (long double)jx), fy = (float)(qy - (long double)jy);
    const int nv = (c->w + 15) / 16, ncol = 16 * nv, tail = c->w & 15;
    const long x0 = jx - c->w / 2, ytop = jy - c->h / 2;
    float above[TILE_COLS], sa[NL] = {0}, ss[NL] = {0};
    const float *a = plane_row(pl, (int)ytop) + x0;
    for (int q = 0; q < ncol; q++)
        above[q] = fmaf(a[q + 1] - a[q], fx, a[q]);
    for (int r = 0; r < c->h; r++) {
        a += pl->stride;
        for (int k = 0; k < nv; k++)
            for (int j = 0; j < NL; j++) {
                const int q = 16 * k + j;
                const float here = fmaf(a[q + 1] - a[q], fx, a[q]);
                float it = (t->t[r][q] - fmaf(here - above[q], fy, above[q])) * q6inv;
                if (k == nv - 1 && tail && j >= tail)
                    it = 0.0f;
                sa[j] = sa[j] + fabsf(it);
                ss[j] = fmaf(it, it, ss[j]);
                above[q] = here;
            }
    }
    const long double n = (long double)(c->w * c->h);
    out[0] = (float)((long double)lane_hsum(sa) / n);
    out[1] = (float)((long double)lane_hsum(ss) / n);
}

/* The pyramid walk in long double (lk_driver.inc:7-109): per level, skip if the window does not fit or G is too
 * weak (doubling the guess), else iterate with the level's inverse of G until the step is below eps. */
#define REAL long double
#define SQRT sqrtl
#define LANE_NAME lane_track
#define LANE_SCHARR 0
#include "lane_track.inc"
#undef REAL
#undef SQRT
#undef LANE_NAME
#undef LANE_SCHARR

/* The same walk with an fp64 driver: the patch arithmetic (and so every fp32 sum) is the lane tier's; only the
 * driver's steps round to 53 instead of 64 bits. */
#define REAL double
#define SQRT sqrt
#define LANE_NAME lane64_track
#define LANE_SCHARR 0
#include "lane_track.inc"
#undef LANE_NAME
#undef LANE_SCHARR

/* The fp64-driver walk with the Scharr gradient variant (no counterpart in the reference). */
#define LANE_NAME lane64_scharr_track
#define LANE_SCHARR 1
#include "lane_track.inc"
#undef REAL
#undef SQRT
#undef LANE_NAME
#undef LANE_SCHARR

/* ------------------------------------------------------------------ LK, f64 tier */

struct f64_tile { double t[LK_MAX_H][LK_MAX_W], tx[LK_MAX_H][LK_MAX_W], ty[LK_MAX_H][LK_MAX_W]; };

static inline double px_at(const struct vt_plane *pl, long x, long y) { return (double)plane_row(pl, (int)y)[x]; }

/* Bilinear sample in double at integer corner (x, y) and fractions (fx, fy). */
static inline double bilin(const struct vt_plane *pl, long x, long y, double fx, double fy)
{
    const double top = px_at(pl, x, y) + fx * (px_at(pl, x + 1, y) - px_at(pl, x, y));
    const double bot = px_at(pl, x, y + 1) + fx * (px_at(pl, x + 1, y + 1) - px_at(pl, x, y + 1));
    return top + fy * (bot - top);
}

static void f64_build(const struct vt_plane *pl, double px, double py, const struct lk_cfg *c, int scharr,
                      struct f64_tile *t, double *g)
{
    /* Sobel [1 2 1] / 8 or Scharr [3 10 3] / 32 across, [-1 0 1] along, per Q6 unit (/ 64) */
    const double we = scharr ? 3.0 : 1.0, wc = scharr ? 10.0 : 2.0, sc = scharr ? 2048.0 : 512.0;
    const double fl_x = floor(px), fl_y = floor(py);
    const long ix = (long)fl_x, iy = (long)fl_y;
    const double fx = px - fl_x, fy = py - fl_y;
    const int W = c->w, H = c->h;
    double s[LK_MAX_H + 2][LK_MAX_W + 2];
    for (int r = 0; r < H + 2; r++)
        for (int k = 0; k < W + 2; k++)
            s[r][k] = bilin(pl, ix - W / 2 - 1 + k, iy - H / 2 - 1 + r, fx, fy);
    double gxx = 0, gxy = 0, gyy = 0;
    for (int r = 0; r < H; r++)
        for (int k = 0; k < W; k++) {
            const double *u = s[r], *m = s[r + 1], *d = s[r + 2];
            const int i = k + 1;
            const double dx = (we * (u[i + 1] - u[i - 1]) + wc * (m[i + 1] - m[i - 1]) + we * (d[i + 1] - d[i - 1])) / sc;
            const double dy = (we * (d[i - 1] - u[i - 1]) + wc * (d[i] - u[i]) + we * (d[i + 1] - u[i + 1])) / sc;
            t->t[r][k] = m[i];
            t->tx[r][k] = dx;
            t->ty[r][k] = dy;
            gxx += dx * dx;
            gxy += dx * dy;
            gyy += dy * dy;
        }
    g[0] = gxx, g[1] = gxy, g[2] = gyy;
}

static void f64_iterate(const struct vt_plane *pl, double qx, double qy, const struct lk_cfg *c,
                        const struct f64_tile *t, double *b, double *res)
{
    const double fl_x = floor(qx), fl_y = floor(qy);
    const double fx = qx - fl_x, fy = qy - fl_y;
    const long x0 = (long)fl_x - c->w / 2, ytop = (long)fl_y - c->h / 2;
    double bx = 0, by = 0, sa = 0, ss = 0;
    for (int r = 0; r < c->h; r++)
        for (int k = 0; k < c->w; k++) {
            const double it = t->t[r][k] - bilin(pl, x0 + k, ytop + r, fx, fy);
            bx += it * t->tx[r][k];
            by += it * t->ty[r][k];
            sa += fabs(it / 64.0);
            ss += (it / 64.0) * (it / 64.0);
        }
    b[0] = bx, b[1] = by;
    if (res) {
        res[0] = sa / (double)(c->w * c->h);
        res[1] = ss / (double)(c->w * c->h);
    }
}

static void f64_track(const struct vt_plane *I, const struct vt_plane *J, int top, const struct lk_cfg *c, int scharr,
                      float x, float y, float gx0, float gy0, struct lk_out *o)
{
    struct f64_tile tile;
    double gx = (double)gx0 * ldexp(1.0, -top), gy = (double)gy0 * ldexp(1.0, -top);
    const double area = (double)(c->w * c->h), eps2 = (double)c->eps * (double)c->eps;
    const int mx = c->w / 2 + 1, my = c->h / 2 + 1;
    double det0 = 0, gxx0 = 0, gxy0 = 0, gyy0 = 0;
    int status = 0, iters = 0, valid0 = 0;
    double eig0 = 0;
    for (int L = top; L >= 0; L--) {
        const double px = ldexp((double)x, -L), py = ldexp((double)y, -L);
        if (px < mx || py < my || px >= (double)(I[L].w - mx - 1) || py >= (double)(I[L].h - my - 1)) {
            if (L == 0)
                status |= LK_OUT;
            else
                gx *= 2, gy *= 2;
            continue;
        }
        double g[3];
        f64_build(&I[L], px, py, c, scharr, &tile, g);
        const double det = g[0] * g[2] - g[1] * g[1], tr = g[0] + g[2], df = g[0] - g[2];
        const double eig = (tr - sqrt(df * df + 4 * g[1] * g[1])) * 0.5 / area;
        if (L == 0)
            eig0 = eig;
        if (eig < (double)c->min_eig || det <= 0) {
            if (L == 0)
                status |= LK_LOWEIG;
            else
                gx *= 2, gy *= 2;
            continue;
        }
        if (L == 0) {
            det0 = det, gxx0 = g[0], gxy0 = g[1], gyy0 = g[2];
            valid0 = 1;
        }
        const double ia = g[2] / det, ib = -g[1] / det, ic = g[0] / det;
        const double lx = (double)(J[L].w - mx - 1), ly = (double)(J[L].h - my - 1);
        double vx = 0, vy = 0;
        for (int it = 0; it < c->max_iter; it++) {
            const double qx = px + gx + vx, qy = py + gy + vy;
            if (qx < mx || qy < my || qx >= lx || qy >= ly) {
                if (L == 0)
                    status |= LK_OUT, valid0 = 0;
                break;
            }
            double b[2];
            f64_iterate(&J[L], qx, qy, c, &tile, b, NULL);
            const double bx = b[0] / 64.0, by = b[1] / 64.0;
            const double ex = ia * bx + ib * by, ey = ib * bx + ic * by;
            vx += ex;
            vy += ey;
            iters++;
            if (ex * ex + ey * ey < eps2)
                break;
        }
        gx += vx;
        gy += vy;
        if (L > 0)
            gx *= 2, gy *= 2;
    }
    const double qx = (double)x + gx, qy = (double)y + gy;
    o->x = (float)qx;
    o->y = (float)qy;
    o->min_eig = (float)eig0;
    o->iters = (uint16_t)iters;
    o->residual = 0.0f;
    o->cov_xx = o->cov_xy = o->cov_yy = 0.0f;
    if (valid0 && !status) {
        if (qx < mx || qy < my || qx >= (double)(J[0].w - mx - 1) || qy >= (double)(J[0].h - my - 1)) {
            status |= LK_OUT;
        } else {
            double b[2], r[2];
            f64_iterate(&J[0], qx, qy, c, &tile, b, r);
            o->residual = (float)r[0];
            if ((float)r[0] > c->max_residual)
                status |= LK_RESIDUAL;
            const double s2 = r[1] * area / (area - 2);
            o->cov_xx = (float)(s2 * gyy0 / det0);
            o->cov_xy = (float)(-s2 * gxy0 / det0);
            o->cov_yy = (float)(s2 * gxx0 / det0);
        }
    }
    o->status = (uint16_t)status;
}

/* ------------------------------------------------------------------ LK entry points */

void vo_lk_track(int tier, const struct vt_plane *I, const struct vt_plane *J, int top, const struct lk_cfg *c,
                 float x, float y, float gx, float gy, struct lk_out *o)
{
    if (tier == VO_F64 || tier == VO_F64_SCHARR)
        f64_track(I, J, top, c, tier == VO_F64_SCHARR, x, y, gx, gy, o);
    else if (tier == VO_LANE64)
        lane64_track(I, J, top, c, x, y, gx, gy, o);
    else if (tier == VO_LANE64_SCHARR)
        lane64_scharr_track(I, J, top, c, x, y, gx, gy, o);
    else
        lane_track(I, J, top, c, x, y, gx, gy, o);
}

/* Track the estimate back from J to I with guess (start - estimate); mark LK_FLOWBACK when that fails or misses the
 * start by more than max_fb (lk_driver.inc:111-121). */
void vo_lk_flowback(int tier, const struct vt_plane *I, const struct vt_plane *J, int top, const struct lk_cfg *c,
                    float x, float y, float max_fb, struct lk_out *o)
{
    if (o->status)
        return;
    struct lk_out back;
    vo_lk_track(tier, J, I, top, c, o->x, o->y, (float)((long double)x - (long double)o->x),
                (float)((long double)y - (long double)o->y), &back);
    const long double ex = (long double)back.x - (long double)x, ey = (long double)back.y - (long double)y;
    if (back.status || ex * ex + ey * ey > (long double)max_fb * (long double)max_fb)
        o->status |= LK_FLOWBACK;
}

struct lk_job {
    int tier, top, n;
    const struct vt_plane *I, *J;
    const struct lk_cfg *c;
    const float *x, *y, *gx, *gy;
    float fb_max;
    struct lk_out *o;
    int next;
};

static void lk_points(void *p, int tid, int nth)
{
    (void)tid, (void)nth;
    struct lk_job *j = (struct lk_job *)p;
    for (;;) {
        const int i = __atomic_fetch_add(&j->next, 1, __ATOMIC_RELAXED);
        if (i >= j->n)
            break;
        vo_lk_track(j->tier, j->I, j->J, j->top, j->c, j->x[i], j->y[i], j->gx ? j->gx[i] : 0.0f,
                    j->gy ? j->gy[i] : 0.0f, &j->o[i]);
        if (j->fb_max > 0.0f)
            vo_lk_flowback(j->tier, j->I, j->J, j->top, j->c, j->x[i], j->y[i], j->fb_max, &j->o[i]);
    }
}

void vo_lk_batch(int tier, const struct vt_plane *I, const struct vt_plane *J, int top, const struct lk_cfg *c, int n,
                 const float *x, const float *y, const float *gx, const float *gy, float fb_max, struct lk_out *o,
                 int threads)
{
    struct lk_job j = {tier, top, n, I, J, c, x, y, gx, gy, fb_max, o, 0};
    vo_run(lk_points, &j, threads);
}

/* ------------------------------------------------------------------ NCC */

/* Zero-mean template of the tw x th window at (x0, y0): lane sums of the pixels, mean = float(sum / n) in long
 * double, z = pixel - mean, norm = lane sum of z^2 (store_appearance, track.c:935-974). */
int vo_ncc_template(const struct vt_plane *pl, int x0, int y0, int tw, int th, float *app, float *norm)
{
    if (x0 < 1 || y0 < 1 || x0 + tw >= pl->w - 1 || y0 + th >= pl->h - 1 || (tw & 15))
        return -1;
    float sum[NL] = {0}, ss[NL] = {0};
    for (int r = 0; r < th; r++)
        for (int k = 0; k < tw; k += NL)
            for (int j = 0; j < NL; j++) {
                const float v = plane_row(pl, y0 + r)[x0 + k + j];
                app[r * tw + k + j] = v;
                sum[j] = sum[j] + v;
            }
    const float mean = (float)((long double)lane_hsum(sum) / (long double)(tw * th));
    for (int r = 0; r < th; r++)
        for (int k = 0; k < tw; k += NL)
            for (int j = 0; j < NL; j++) {
                const float z = app[r * tw + k + j] - mean;
                app[r * tw + k + j] = z;
                ss[j] = fmaf(z, z, ss[j]);
            }
    *norm = lane_hsum(ss);
    return 0;
}

/* Zero-mean NCC of the template at corner (x, y): window values relative to its first pixel, lane sums
 * st = sum v t, sj = sum v, sjj = sum v^2, finished in long double; -1 outside, 0 for a window with under 1/16 of the
 * template's energy (search_at, track.c:184-206, generalised from 32 x 21 to tw x th). */
long double vo_ncc_at(const struct vt_plane *pl, const float *app, int tw, int th, long double norm, int x, int y)
{
    if (x < 1 || y < 1 || x + tw >= pl->w - 1 || y + th >= pl->h - 1)
        return -1.0L;
    const float first = plane_row(pl, y)[x];
    float st[NL] = {0}, sj[NL] = {0}, sjj[NL] = {0};
    for (int r = 0; r < th; r++) {
        const float *row = plane_row(pl, y + r) + x;
        for (int k = 0; k < tw; k += NL)
            for (int j = 0; j < NL; j++) {
                const float v = row[k + j] - first;
                st[j] = fmaf(v, app[r * tw + k + j], st[j]);
                sj[j] = sj[j] + v;
                sjj[j] = fmaf(v, v, sjj[j]);
            }
    }
    const long double n = (long double)(tw * th);
    const long double a = lane_hsum(st), b = lane_hsum(sj), c = lane_hsum(sjj);
    const long double var = c - b * b / n;
    return (var <= norm * 0.0625L || norm <= 0.0L) ? 0.0L : a / sqrtl(norm * var);
}

/* The same at a corner in half pixels: a half step averages two neighbours, both halves four (search_half,
 * track.c:1615-1647). */
long double vo_ncc_half(const struct vt_plane *pl, const float *app, int tw, int th, long double norm, int hx, int hy)
{
    const int x = hx >> 1, y = hy >> 1, ox = hx & 1, oy = hy & 1;
    if (x < 1 || y < 1 || x + tw + 1 >= pl->w - 1 || y + th + 1 >= pl->h - 1)
        return -1.0L;
    const float first = plane_row(pl, y)[x];
    float st[NL] = {0}, sj[NL] = {0}, sjj[NL] = {0};
    for (int r = 0; r < th; r++) {
        const float *row = plane_row(pl, y + r) + x, *below = plane_row(pl, y + r + 1) + x;
        for (int k = 0; k < tw; k += NL)
            for (int j = 0; j < NL; j++) {
                float v = row[k + j];
                if (ox)
                    v = (v + row[k + j + 1]) * 0.5f;
                if (oy) {
                    float w = below[k + j];
                    if (ox)
                        w = (w + below[k + j + 1]) * 0.5f;
                    v = (v + w) * 0.5f;
                }
                v = v - first;
                st[j] = fmaf(v, app[r * tw + k + j], st[j]);
                sj[j] = sj[j] + v;
                sjj[j] = fmaf(v, v, sjj[j]);
            }
    }
    const long double n = (long double)(tw * th);
    const long double a = lane_hsum(st), b = lane_hsum(sj), c = lane_hsum(sjj);
    const long double var = c - b * b / n;
    return (var <= norm * 0.0625L || norm <= 0.0L) ? 0.0L : a / sqrtl(norm * var);
}

/* Parabola through three samples: the peak's offset, clamped to half a step (peak_offset, track.c:1650-1657). */
long double vo_peak_offset(long double l, long double c, long double r)
{
    const long double den = l - 2.0L * c + r;
    if (den >= 0.0L)
        return 0.0L;
    const long double o = 0.5L * (l - r) / den;
    return o < -0.5L ? -0.5L : o > 0.5L ? 0.5L : o;
}

struct cell_job {
    const struct vt_plane *pl;
    const float *app;
    int tw, th, x0, x1, y0, y1, step, cols, cells;
    float norm;
    struct vo_cell *out;
    int next;
};

/* Best corner of each 16 x 16 cell by strict > in row-major order (OP_SEARCH, track.c:301-329). */
static void cell_worker(void *p, int tid, int nth)
{
    (void)tid, (void)nth;
    struct cell_job *j = (struct cell_job *)p;
    for (;;) {
        const int cell = __atomic_fetch_add(&j->next, 1, __ATOMIC_RELAXED);
        if (cell >= j->cells)
            break;
        const int cy0 = j->y0 + (cell / j->cols) * 16, cx0 = j->x0 + (cell % j->cols) * 16;
        const int cy1 = cy0 + 16 <= j->y1 + 1 ? cy0 + 16 : j->y1 + 1;
        const int cx1 = cx0 + 16 <= j->x1 + 1 ? cx0 + 16 : j->x1 + 1;
        float top = -2.0f;
        int bx = cx0, by = cy0;
        for (int y = cy0; y < cy1; y += j->step)
            for (int x = cx0; x < cx1; x += j->step) {
                const float v = (float)vo_ncc_at(j->pl, j->app, j->tw, j->th, (long double)j->norm, x, y);
                if (v > top)
                    top = v, bx = x, by = y;
            }
        j->out[cell] = (struct vo_cell){top, bx, by};
    }
}

void vo_ncc_cells(const struct vt_plane *pl, const float *app, int tw, int th, float norm, int x0, int x1, int y0,
                  int y1, int step, struct vo_cell *cells, int threads)
{
    const int cols = (x1 - x0) / 16 + 1, rows = (y1 - y0) / 16 + 1;
    struct cell_job j = {pl, app, tw, th, x0, x1, y0, y1, step, cols, cols * rows, norm, cells, 0};
    vo_run(cell_worker, &j, threads);
}

/* search_slot (track.c:1661-1726): clip the corner window, map the cells, keep the 8 best cells (ties keep the earlier
 * cell), refine each in half pixels within 2 step - 1 (1 at step 1), take the best, and the best clear of it by half a
 * template as the second peak. */
void vo_ncc_search(const struct vt_plane *pl, const float *app, int tw, int th, float norm, int x0, int x1, int y0,
                   int y1, int step, struct vo_found *f, struct vo_cell *cells_out, int threads)
{
    const int hi_x = pl->w - tw - 2, hi_y = pl->h - th - 2;
    x0 = x0 < 1 ? 1 : x0;
    x1 = x1 > hi_x ? hi_x : x1;
    y0 = y0 < 1 ? 1 : y0;
    y1 = y1 > hi_y ? hi_y : y1;
    f->ncc = -2.0L;
    f->second = -2.0L;
    f->hx = f->hy = 0;
    f->cells = 0;
    if (x1 < x0 || y1 < y0)
        return;
    const int cells = ((x1 - x0) / 16 + 1) * ((y1 - y0) / 16 + 1);
    struct vo_cell *sc = cells_out ? cells_out : (struct vo_cell *)malloc(sizeof(struct vo_cell) * (size_t)cells);
    vo_ncc_cells(pl, app, tw, th, norm, x0, x1, y0, y1, step, sc, threads);
    f->cells = cells;
    int top[8], ntop = 0;
    for (int i = 0; i < cells; i++) {
        if (sc[i].v <= -2.0f || (ntop == 8 && sc[i].v <= sc[top[ntop - 1]].v))
            continue;
        int at = ntop < 8 ? ntop++ : 7;
        for (; at > 0 && sc[top[at - 1]].v < sc[i].v; at--)
            top[at] = top[at - 1];
        top[at] = i;
    }
    const int around = step > 1 ? 2 * step - 1 : 1;
    long double score[8];
    int shx[8], shy[8];
    for (int c = 0; c < ntop; c++) {
        const int cx = 2 * sc[top[c]].x, cy = 2 * sc[top[c]].y;
        score[c] = -2.0L;
        shx[c] = cx, shy[c] = cy;
        for (int hy = cy - around; hy <= cy + around; hy++)
            for (int hx = cx - around; hx <= cx + around; hx++) {
                const long double v = vo_ncc_half(pl, app, tw, th, (long double)norm, hx, hy);
                if (v > score[c])
                    score[c] = v, shx[c] = hx, shy[c] = hy;
            }
        if (score[c] > f->ncc)
            f->ncc = score[c], f->hx = shx[c], f->hy = shy[c];
    }
    for (int c = 0; c < ntop; c++) {
        const int dx = shx[c] - f->hx, dy = shy[c] - f->hy;
        if ((dx * dx > tw * tw || dy * dy > th * th) && score[c] > f->second)
            f->second = score[c];
    }
    if (!cells_out)
        free(sc);
}

/* ------------------------------------------------------------------ still-cell vectors (src/kern/still.c)
 * Pinned against the reference's own still.c built from source in oracle/_ref (tests/test_oracle.py: the hash of
 * every cell and the mean after each of a run of frames, and the motion summary, equal bit for bit), and by closed
 * forms (identical frames: zero vectors and differences; a constant offset: the exact difference). */
#define ST_PATCH 15          /* still.c:12-18: the LK patch and its limits */
#define ST_ITER 6
#define ST_EPS 0.01f
#define ST_MIN_EIG 0.1f
#define ST_MAX_RES 16.0f
#define ST_FB 0.25f
#define ST_VEC_SIGMA 12.0L   /* still.c:19-22: how far past LK's own sigma a vector must reach, the noise rules */
#define ST_DIFF_NOISE 3.0L
#define ST_NOISE_FRAMES 32.0L
#define ST_EVIDENCE 2
#define ST_MARGIN 9.0f       /* still.c:23 */
#define ST_BINS 256          /* still.c:29-30 */
#define ST_PER_GREY 16

struct vo_still *vo_still_new(const struct vtg_still_cfg *c)
{
    struct vo_still *s = calloc(1, sizeof *s);
    s->w = c->w;
    s->h = c->h;
    s->avg = c->avg;
    s->motion_px = c->motion_px;
    s->motion_grey = c->motion_grey;
    s->lk_tier = VO_LANE;
    s->cols = (c->w + VTG_CELL_W - 1) / VTG_CELL_W;
    s->rows = (c->h + VTG_CELL_H - 1) / VTG_CELL_H;
    s->ncells = s->cols * s->rows;
    s->mean = vo_plane_new(c->w, c->h);
    s->cell = calloc((size_t)s->ncells, sizeof *s->cell);
    s->raw = calloc((size_t)(s->cols + 2) * (size_t)(s->rows + 2), 1);
    s->count = calloc((size_t)s->ncells, sizeof *s->count);
    s->weight = calloc((size_t)s->ncells, sizeof *s->weight);
    return s;
}

void vo_still_free(struct vo_still *s)
{
    if (!s)
        return;
    vo_plane_free(&s->mean);
    free(s->cell);
    free(s->raw);
    free(s->count);
    free(s->weight);
    free(s);
}

struct lk_cfg vo_still_lk_cfg(void)
{
    struct lk_cfg c = {ST_PATCH, ST_PATCH, ST_ITER, ST_EPS, ST_MIN_EIG, ST_MAX_RES, 0};
    return c;
}

void vo_still_centre(const struct vo_still *s, int i, float *x, float *y)
{
    const float xm = (float)s->w - 1.0f - ST_MARGIN, ym = (float)s->h - 1.0f - ST_MARGIN;
    float cx = (float)((i % s->cols) * VTG_CELL_W + VTG_CELL_W / 2), cy = (float)((i / s->cols) * VTG_CELL_H + VTG_CELL_H / 2);
    *x = cx < ST_MARGIN ? ST_MARGIN : cx > xm ? xm : cx;
    *y = cy < ST_MARGIN ? ST_MARGIN : cy > ym ? ym : cy;
}

/* cell_diff (still.c:99-111): per row of the cell, the 16 lanes' |frame - mean| summed in the lane order; the rows
 * summed in long double; the mean over rows x 16 (the lanes past the plane's edge read its padding). */
float vo_still_cell_diff(const struct vo_still *s, const struct vt_plane *cur, int col, int row)
{
    const int x = col * VTG_CELL_W, y0 = row * VTG_CELL_H, y1 = y0 + VTG_CELL_H < s->h ? y0 + VTG_CELL_H : s->h;
    long double sum = 0.0L;
    for (int y = y0; y < y1; y++) {
        float d[NL];
        for (int i = 0; i < NL; i++)
            d[i] = fabsf(plane_row(cur, y)[x + i] - plane_row(&s->mean, y)[x + i]);
        sum += (long double)lane_hsum(d);
    }
    return (float)(sum / (long double)((y1 - y0) * VTG_CELL_W));
}

/* still_vectors (still.c:113-138): LK (of the tier s->lk_tier names) from the mean to the frame at each cell's centre, with the flowback
 * check; the vector, the difference in 8-bit grey, LK's status and sigma. */
void vo_still_vectors(struct vo_still *s, const struct vt_plane *cur, int threads)
{
    const int n = s->ncells;
    float *xs = malloc(sizeof(float) * (size_t)n), *ys = malloc(sizeof(float) * (size_t)n);
    struct lk_out *o = malloc(sizeof *o * (size_t)n);
    for (int i = 0; i < n; i++)
        vo_still_centre(s, i, &xs[i], &ys[i]);
    const struct lk_cfg cfg = vo_still_lk_cfg();
    vo_lk_batch(s->lk_tier, &s->mean, cur, 0, &cfg, n, xs, ys, NULL, NULL, ST_FB, o, threads);
    for (int i = 0; i < n; i++) {
        struct vtg_still_cell *c = &s->cell[i];
        c->u = (float)((long double)o[i].x - (long double)xs[i]);
        c->v = (float)((long double)o[i].y - (long double)ys[i]);
        c->diff = (float)((long double)vo_still_cell_diff(s, cur, i % s->cols, i / s->cols) / (long double)VT_Q6);
        c->state = o[i].status;
        c->sigma = o[i].status ? 0.0f : (float)sqrtl((long double)o[i].cov_xx + (long double)o[i].cov_yy);
    }
    free(xs);
    free(ys);
    free(o);
}

/* still_decide (still.c:140-237): the camera's noise as the median cell difference (a 1/16-grey histogram), the
 * usual noise followed slowly unless everything changed, each cell's own evidence (moved past max(motion_px,
 * 12 sigma), LK lost its patch, or the difference over the limit), a 3 x 3 vote of at least two, the cell's frame
 * count and weight, its state, and the drift of the cells that held still. */
void vo_still_decide(struct vo_still *s)
{
    const int cols = s->cols, rows = s->rows, wide = cols + 2, n = s->ncells;
    uint32_t hist[ST_BINS] = {0};
    for (int i = 0; i < n; i++) {
        const long b = (long)((long double)s->cell[i].diff * (long double)ST_PER_GREY);
        hist[b < ST_BINS - 1 ? b : ST_BINS - 1]++;
    }
    int med = 0;
    uint32_t below = 0;
    while (med < ST_BINS - 1 && below + hist[med] < (uint32_t)n / 2u)
        below += hist[med++];
    const long double noise = (long double)med / (long double)ST_PER_GREY, three = noise * ST_DIFF_NOISE;
    const long double limit = three > (long double)s->motion_grey ? three : (long double)s->motion_grey;
    s->diff_median = (float)noise;
    s->diff_limit = (float)limit;
    const long double usual3 = s->noise_usual * ST_DIFF_NOISE;
    const long double usual = usual3 > (long double)s->motion_grey ? usual3 : (long double)s->motion_grey;
    const int everything = s->frames >= 2 && noise > usual;
    if (!everything && s->frames >= 1)
        s->noise_usual = s->frames == 1 ? noise : s->noise_usual + (noise - s->noise_usual) / ST_NOISE_FRAMES;
    const long double px2 = (long double)s->motion_px * (long double)s->motion_px;
    for (int i = 0; i < n; i++) {
        const struct vtg_still_cell *c = &s->cell[i];
        const uint32_t lk = c->state & VTG_CELL_LK;
        const long double u = c->u, v = c->v, sg = (long double)c->sigma * ST_VEC_SIGMA;
        const long double need = sg * sg > px2 ? sg * sg : px2;
        const int moved = lk == 0 && u * u + v * v > need;
        const int lost = (lk & LK_RESIDUAL) != 0;
        const int changed = (long double)c->diff > limit;
        s->raw[(size_t)(i / cols + 1) * (size_t)wide + (size_t)(i % cols + 1)] = (uint8_t)(moved || lost || changed);
    }
    uint32_t tracked = 0, moving = 0, quiet = 0;
    long double su = 0.0L, sv = 0.0L, suu = 0.0L, counts = 0.0L;
    for (int r = 0; r < rows; r++)
        for (int q = 0; q < cols; q++) {
            const int i = r * cols + q;
            uint32_t votes = 0;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++)
                    votes += s->raw[(size_t)(r + 1 + dy) * (size_t)wide + (size_t)(q + 1 + dx)];
            const int any = votes >= ST_EVIDENCE || everything;
            const int own = s->raw[(size_t)(r + 1) * (size_t)wide + (size_t)(q + 1)];
            const uint32_t grown = (uint32_t)s->count[i] + 1u;
            const uint32_t k = any ? 1u : grown < (uint32_t)s->avg ? grown : (uint32_t)s->avg;
            s->count[i] = (uint16_t)k;
            s->weight[i] = (float)(1.0L / (long double)k);
            struct vtg_still_cell *c = &s->cell[i];
            const uint32_t lk = c->state & VTG_CELL_LK;
            c->state = lk | (own ? VTG_CELL_CHANGED : 0u) | (any ? VTG_CELL_MOVING : 0u) | (k << VTG_CELL_COUNT_SHIFT);
            tracked += lk == 0;
            moving += any != 0;
            counts += (long double)k;
            if (lk == 0 && !any) {
                const long double u = c->u, v = c->v;
                su += u;
                sv += v;
                suu += u * u + v * v;
                quiet++;
            }
        }
    s->tracked = tracked;
    s->moving = moving;
    s->count_mean = (float)(counts / (long double)n);
    const long double mu = quiet ? su / (long double)quiet : 0.0L, mv = quiet ? sv / (long double)quiet : 0.0L;
    s->drift_u = (float)mu;
    s->drift_v = (float)mv;
    const long double var = quiet ? suu / (long double)quiet - mu * mu - mv * mv : 0.0L;
    s->drift_rms = var > 0.0L ? (float)sqrtl(var) : 0.0f;
}

/* still_update (still.c:239-262): each 16-lane group of a row takes its cell's weight: a weight of 1 copies the
 * frame, otherwise mean + (frame - mean) * weight in one fused multiply-add; the rows' padding renewed. */
void vo_still_update(struct vo_still *s, const struct vt_plane *cur)
{
    for (int r = 0; r < s->rows; r++) {
        const int y0 = r * VTG_CELL_H, y1 = y0 + VTG_CELL_H < s->h ? y0 + VTG_CELL_H : s->h;
        for (int y = y0; y < y1; y++) {
            const float *in = plane_row(cur, y);
            float *m = plane_row(&s->mean, y);
            for (int x = 0; x < s->w; x += VTG_CELL_W) {
                const float wt = s->weight[r * s->cols + x / VTG_CELL_W];
                for (int i = 0; i < NL; i++)
                    m[x + i] = wt == 1.0f ? in[x + i] : fmaf(in[x + i] - m[x + i], wt, m[x + i]);
            }
        }
        vo_pad_rows(&s->mean, y0, y1);
    }
}

/* still_finish (still.c:264-270) */
void vo_still_finish(struct vo_still *s)
{
    vo_pad_tb(&s->mean);
    s->frames++;
}

/* still_hash (still.c:272-278) */
uint64_t vo_still_hash(const struct vo_still *s, uint64_t h)
{
    h = vo_fnv(h, s->cell, (size_t)s->ncells * sizeof(struct vtg_still_cell));
    for (int y = 0; y < s->h; y++)
        h = vo_fnv(h, plane_row(&s->mean, y), (size_t)s->w * sizeof(float));
    return h;
}

/* ------------------------------------------------------------------ stereo row scan (src/kern/stereo.c:1-80)
 * Pinned against the reference's own stereo_scan_vx from oracle/_ref (tests/test_oracle_stereo.py: every field bit
 * for bit over random features, disparity ranges and edge cases), and by closed forms (a right image that is the left
 * shifted by d: peak at d with correlation 1; a flat template: no correlation). */
#define SC_W 16
#define SC_H 11
#define SC_MAX 512

static long double sc_ncc(float st, float sj, float sjj, long double norm_t, long double n)
{
    const long double var = (long double)sjj - (long double)sj * (long double)sj / n;
    if (var <= 0 || norm_t <= 0)
        return 0;
    return (long double)st / sqrtl(norm_t * var);
}

void vo_stereo_scan(const struct vt_plane *left, const struct vt_plane *right, float x, float y, int dmin, int dmax,
                    struct vo_stereo_scan *out)
{
    const int x0 = (int)x - SC_W / 2, y0 = (int)y - SC_H / 2;
    float t[SC_H][SC_W], sum[SC_W] = {0};
    for (int r = 0; r < SC_H; r++)
        for (int k = 0; k < SC_W; k++) {
            t[r][k] = plane_row(left, y0 + r)[x0 + k];
            sum[k] += t[r][k];
        }
    const long double n = (long double)(SC_W * SC_H);
    const float mean = (float)((long double)lane_hsum(sum) / n);
    float stt[SC_W] = {0};
    for (int r = 0; r < SC_H; r++)
        for (int k = 0; k < SC_W; k++) {
            t[r][k] -= mean;
            stt[k] = fmaf(t[r][k], t[r][k], stt[k]);
        }
    const long double norm_t = lane_hsum(stt);
    long double curve[SC_MAX];
    int count = 0;
    for (int d = dmin; d <= dmax && count < SC_MAX; d++, count++) {
        const int xr = x0 - d;
        if (xr < 1 || xr + SC_W >= right->w - 1) {
            curve[count] = -1;
            continue;
        }
        float st[SC_W] = {0}, sj[SC_W] = {0}, sjj[SC_W] = {0};
        for (int r = 0; r < SC_H; r++)
            for (int k = 0; k < SC_W; k++) {
                const float j = plane_row(right, y0 + r)[xr + k];
                st[k] = fmaf(j, t[r][k], st[k]);
                sj[k] += j;
                sjj[k] = fmaf(j, j, sjj[k]);
            }
        curve[count] = sc_ncc(lane_hsum(st), lane_hsum(sj), lane_hsum(sjj), norm_t, n);
    }
    /* the first highest candidate, then the highest at least 3 candidates from it (stereo.c:17-34) */
    int bi = -1;
    long double bv = -2;
    for (int i = 0; i < count; i++)
        if (curve[i] > bv) {
            bv = curve[i];
            bi = i;
        }
    long double sv = -1;
    for (int i = 0; i < count; i++)
        if ((i < bi - 2 || i > bi + 2) && curve[i] > sv)
            sv = curve[i];
    out->disparity = bi < 0 ? -1.0f : (float)(dmin + bi);
    out->ncc = bi < 0 ? 0.0f : (float)bv;
    out->second = (float)sv;
}

/* ------------------------------------------------------------------ stereo geometry and match
 * (src/kern/stereo.c:124-141, src/kern/track.c:1859-1914). The geometry is pinned bit for bit against the reference's
 * own stereo_triangulate / stereo_reprojection from oracle/_ref; the match, against the reference's own scan, LK and
 * geometry composed in the tracker's order (tests/test_oracle_stereo.py), and by a closed form (a shifted copy: every
 * textured feature matched at the shift, its depth f B / d). */
int vo_stereo_triangulate(const struct vtg_stereo_geom *g, long double xl, long double yl, long double xr,
                          long double yr, long double *p)
{
    const long double d = xl - xr;
    if (d <= 0.0L)
        return 0;
    const long double z = g->fx * g->baseline / d;
    p[0] = (xl - g->cx) * z / g->fx;
    p[1] = (0.5L * (yl + yr) - g->cy) * z / g->fy;
    p[2] = z;
    return 1;
}

long double vo_stereo_reprojection(const struct vtg_stereo_geom *g, const long double *p, long double xl,
                                   long double yl, long double xr, long double yr)
{
    const long double ul = g->fx * p[0] / p[2] + g->cx, vl = g->fy * p[1] / p[2] + g->cy;
    const long double ur = g->fx * (p[0] - g->baseline) / p[2] + g->cx;
    const long double el = ul - xl, fl = vl - yl, er = ur - xr, fr = vl - yr;
    return sqrtl((el * el + fl * fl + er * er + fr * fr) * 0.5L);
}

/* The tracker's stereo step for n features (stereo_one, track.c:1859-1914): a feature without a prior disparity is
 * scanned at the stereo level (rejected below min_ncc, or marked ambiguous when the second peak is too close); LK (lane
 * tier) from the left to the right pyramid with the disparity as its x guess, from the stereo level (or level 0 with a
 * prior), the flowback from level 0; the match, its disparity, and the 3D point where the disparity is in range and it
 * reprojects within max_reproj. */
void vo_stereo_match_tier(int tier, const struct vt_plane *L, const struct vt_plane *R, const struct vtg_stereo_cfg *c,
                          int n, struct vtg_stereo_feature *f)
{
    for (int i = 0; i < n; i++) {
        struct vtg_stereo_feature *q = &f[i];
        uint32_t fl = q->flags & VTG_FF_NEW;
        q->xr = q->yr = 0.0f;
        q->z = 0.0f;
        float guess = q->disparity;
        int top = 0;
        if (guess <= 0.0f) {
            const int sl = c->stereo_level;
            const long double s = 1.0L / (long double)(1 << sl);
            const int dmin = (int)((long double)c->min_disp * s), dmax = (int)((long double)c->max_disp * s);
            struct vo_stereo_scan sc;
            vo_stereo_scan(&L[sl], &R[sl], (float)((long double)q->x * s), (float)((long double)q->y * s), dmin, dmax,
                           &sc);
            if (sc.disparity < 0.0f || sc.ncc < c->min_ncc) {
                q->flags = fl;
                q->disparity = -1.0f;
                continue;
            }
            if ((long double)sc.second > (long double)sc.ncc * (long double)c->ambiguity) {
                q->flags = fl | VTG_FF_AMBIGUOUS;
                q->disparity = -1.0f;
                continue;
            }
            guess = (float)((long double)sc.disparity * (long double)(1 << sl));
            top = sl;
        }
        struct lk_out m;
        vo_lk_track(tier, L, R, top, &c->lk, q->x, q->y, (float)(-(long double)guess), 0.0f, &m);
        if (c->fb_max > 0.0f)
            vo_lk_flowback(tier, L, R, 0, &c->lk, q->x, q->y, c->fb_max, &m);
        if (m.status) {
            q->flags = fl;
            q->disparity = -1.0f;
            continue;
        }
        q->xr = m.x;
        q->yr = m.y;
        const long double disparity = (long double)q->x - (long double)m.x;
        q->disparity = (float)disparity;
        fl |= VTG_FF_STEREO;
        if (disparity >= (long double)c->min_disp && disparity <= (long double)c->max_disp &&
            vo_stereo_triangulate(&c->geom, q->x, q->y, m.x, m.y, q->p) &&
            vo_stereo_reprojection(&c->geom, q->p, q->x, q->y, m.x, m.y) <= (long double)c->max_reproj) {
            fl |= VTG_FF_3D;
            q->z = (float)q->p[2];
        }
        q->flags = fl;
    }
}

void vo_stereo_match(const struct vt_plane *L, const struct vt_plane *R, const struct vtg_stereo_cfg *c, int n,
                     struct vtg_stereo_feature *f)
{
    vo_stereo_match_tier(VO_LANE, L, R, c, n, f);
}
