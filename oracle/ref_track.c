/* ref_track.c: TEST INFRASTRUCTURE. Compiles the reference's own track.c (src/kern/track.c, read from the reference
 * tree through the include path) into oracle/_ref/ and exports four of its static functions so the tests can pin the
 * oracle's restatement against them:
 *   store_appearance  (track.c:935-974)   the zero-mean template and its norm
 *   search_half       (track.c:1615-1647) the correlation at a corner in half pixels
 *   peak_offset       (track.c:1650-1657) the parabola vertex, clamped to half a step
 *   search_slot       (track.c:1661-1726) the cell map, top 8, half-pixel refinement, pick and second peak
 * Each runs on a tracker that holds only what the function reads (one appearance slot, one level of the current set,
 * the map's arrays, an executor that runs the job inline). Built only by `make ref`; nothing here ships. */
#include <stdlib.h>

#include "track.c"

static struct tracker *bare_tracker(void)
{
    struct tracker *t = aligned_alloc(64, (sizeof *t + 63) & ~(size_t)63);
    struct job *j = aligned_alloc(64, (sizeof *j + 63) & ~(size_t)63);
    if (!t || !j) {
        free(t);
        free(j);
        return NULL;
    }
    memset(t, 0, sizeof *t);
    memset(j, 0, sizeof *j);
    t->jt = j;
    j->t = t;
    t->prev = -1;
    return t;
}

static void bare_free(struct tracker *t)
{
    if (!t)
        return;
    free(t->jt);
    free(t);
}

static void run_inline(void *ctx, tk_fn fn, void *arg)
{
    (void)ctx;
    fn(arg, 0, 1);
}

/* store_appearance on level plane `pl` (as level 0 of a one-level pyramid) for a region centred at (cx, cy) of size
 * rw x rh: writes the 32 x 21 template to pix and returns 0, or -1 when the template would not fit (the reference then
 * leaves the slot unchanged). */
int ref_store_appearance(const struct vt_plane *pl, float cx, float cy, float rw, float rh, float *pix, float *norm)
{
    struct tracker *t = bare_tracker();
    float *buf = aligned_alloc(64, APP_W * APP_H * sizeof(float));
    if (!t || !buf) {
        bare_free(t);
        free(buf);
        return -2;
    }
    t->levels = 1;
    t->cur = 0;
    t->pyr[0][0].l[0] = *pl;
    t->cx = cx;
    t->cy = cy;
    t->rw = rw;
    t->rh = rh;
    t->app[0].pix = buf;
    store_appearance(t, 0);
    const int ok = t->app[0].valid;
    if (ok) {
        memcpy(pix, buf, APP_W * APP_H * sizeof(float));
        *norm = t->app[0].norm;
    }
    free(buf);
    bare_free(t);
    return ok ? 0 : -1;
}

long double ref_search_half(const struct vt_plane *pl, const float *pix, float norm, int hx, int hy)
{
    float *buf = aligned_alloc(64, APP_W * APP_H * sizeof(float));
    if (!buf)
        return -3.0L;
    memcpy(buf, pix, APP_W * APP_H * sizeof(float));
    const struct app_slot a = {.pix = buf, .norm = norm};
    const long double v = search_half(pl, &a, hx, hy);
    free(buf);
    return v;
}

long double ref_peak_offset(long double l, long double c, long double r) { return peak_offset(l, c, r); }

struct ref_found {
    int hx, hy;
    long double ncc, second;
};

/* search_slot for a template (pix, norm) at `level` of the current set, whose plane is pl. The map's arrays hold
 * cap cells (tk_create sizes them (w / 16 + 2) x (h / 16 + 2) for the full frame); *cells gets the map's size. */
int ref_search_slot(const struct vt_plane *pl, int level, const float *pix, float norm, int x0, int x1, int y0, int y1,
                    int step, struct ref_found *out, float *sc_v, int *sc_x, int *sc_y, int cap, int *cells)
{
    if (level < 0 || level >= VT_MAX_LEVELS)
        return -1;
    struct tracker *t = bare_tracker();
    float *buf = aligned_alloc(64, APP_W * APP_H * sizeof(float));
    if (!t || !buf) {
        bare_free(t);
        free(buf);
        return -2;
    }
    memcpy(buf, pix, APP_W * APP_H * sizeof(float));
    t->levels = level + 1;
    t->cur = 0;
    t->pyr[0][0].l[level] = *pl;
    t->app[0] = (struct app_slot){.pix = buf, .norm = norm, .level = level, .valid = 1};
    t->sc_v = sc_v;
    t->sc_x = sc_x;
    t->sc_y = sc_y;
    t->run_track = run_inline;
    /* the clip search_slot applies, to check the map fits the caller's arrays before it runs */
    const int hx = pl->w - APP_W - 2, hy = pl->h - APP_H - 2;
    const int cx0 = x0 < 1 ? 1 : x0, cx1 = x1 > hx ? hx : x1, cy0 = y0 < 1 ? 1 : y0, cy1 = y1 > hy ? hy : y1;
    const int n = (cx1 < cx0 || cy1 < cy0) ? 0 : ((cx1 - cx0) / SEARCH_CELL + 1) * ((cy1 - cy0) / SEARCH_CELL + 1);
    if (n > cap) {
        free(buf);
        bare_free(t);
        return -1;
    }
    struct found f;
    search_slot(t, 0, x0, x1, y0, y1, step, VT_ALG_MVEX, &f);
    out->hx = f.hx;
    out->hy = f.hy;
    out->ncc = f.ncc;
    out->second = f.second;
    *cells = n;
    free(buf);
    bare_free(t);
    return 0;
}
