/* tk_scene.c: a host program in the shape of the reference's replay host (src/host/vt_replay.c:16-26 host_alloc and
 * host_run, 125-264 the frame loop), compiled against include/ and linked with libvtgpu.so and libvtsynth.so. One
 * layered scene is rendered frame by frame (vts_scene_frame); frame k is built into pyramid set k mod the set count
 * (vtg_tk_build) and tracked (vtg_tk_track); every feature that is not new is scored as the reference's score.c scores
 * it (score.c:160-170): its last place, moved by the scene's true motion (vts_scene_map), against where the tracker
 * put it. The caller seeds features (the reference's region does): a grid over the object's true box at the first
 * frame, and again whenever fewer than half of them are left. One JSON line on stdout.
 *
 *     tk_scene SCENE W H FRAMES [SEED]      SCENE: translate camera shapes brightness noise occlusion */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vtg_track.h"
#include "vtsynth.h"

static void *host_alloc(void *ctx, size_t bytes)
{
    (void)ctx;
    return aligned_alloc(64, bytes);
}

static void host_run(void *ctx, tk_fn fn, void *arg)
{
    (void)ctx;
    fn(arg, 0, 1);
}

static int cmp_double(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y;
}

static double pct(const double *v, int n, double q)
{
    if (n == 0)
        return -1.0;
    int i = (int)(q * (n - 1) + 0.5);
    return v[i < 0 ? 0 : i >= n ? n - 1 : i];
}

/* A grid of up to nx x ny features over the object's box at time t, a margin of m pixels inside it (a disc: inside
 * its inscribed square). */
static int seed_grid(int scene, int w, int h, double t, int nx, int ny, double m, float *xs, float *ys)
{
    double cx, cy, hx, hy;
    vts_scene_object(scene, w, h, t, &cx, &cy, &hx, &hy);
    if (scene == VTS_SC_SHAPES)
        hx = hy = hx * 0.70710678;
    int n = 0;
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const double x = cx - hx + m + (2.0 * (hx - m)) * (i + 0.5) / nx;
            const double y = cy - hy + m + (2.0 * (hy - m)) * (j + 0.5) / ny;
            if (x < 16 || y < 16 || x > w - 17 || y > h - 17)
                continue;
            xs[n] = (float)x;
            ys[n] = (float)y;
            n++;
        }
    return n;
}

int main(int argc, char **argv)
{
    static const char *names[VTS_SC_COUNT] = {"translate", "camera", "shapes", "brightness", "noise", "occlusion"};
    if (argc < 5) {
        fprintf(stderr, "usage: %s SCENE W H FRAMES [SEED]\n", argv[0]);
        return 2;
    }
    int scene = -1;
    for (int i = 0; i < VTS_SC_COUNT; i++)
        if (!strcmp(argv[1], names[i]))
            scene = i;
    const int w = atoi(argv[2]), h = atoi(argv[3]), frames = atoi(argv[4]);
    const uint64_t seed = argc > 5 ? strtoull(argv[5], NULL, 0) : 1;
    if (scene < 0 || w < 64 || h < 64 || frames < 2) {
        fprintf(stderr, "%s: bad scene or size\n", argv[0]);
        return 2;
    }
    /* the dataset's tracker settings (src/host/dataset.c:132-140) */
    struct vtg_tk_config cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.width = (uint32_t)w;
    cfg.height = (uint32_t)h;
    cfg.levels = 4;
    cfg.patch_w = 21;
    cfg.patch_h = 21;
    cfg.max_iter = 10;
    cfg.eps = 0.01f;
    cfg.min_eig = 0.1f;
    cfg.max_residual = 25.0f;
    cfg.fb_max = 0.5f;
    cfg.max_features = 128;
    cfg.sets = 2;
    cfg.storage = VTG_STORE_Q6;
    cfg.device = 0;
    struct vtg_tracker *t = vtg_tk_create(&cfg, 1, 1, host_alloc, NULL, host_run, NULL, host_run, NULL);
    if (!t)
        return 1;
    const double fps = 30.0;
    uint8_t *frame = (uint8_t *)aligned_alloc(64, ((size_t)w * (size_t)h + 63) & ~(size_t)63);
    struct vtg_tk_result *res = (struct vtg_tk_result *)malloc(sizeof *res);
    double *epe = (double *)malloc(sizeof(double) * (size_t)frames * VTG_TK_MAX_FEATURES);
    float xs[VTG_TK_MAX_FEATURES], ys[VTG_TK_MAX_FEATURES];
    int nepe = 0, seeded = 0, reseeds = 0, dropped = 0, tracked = 0, off_layer = 0;
    for (int k = 0; k < frames; k++) {
        const double tk = k / fps;
        if (vts_scene_frame(scene, tk, frame, w, h, w, seed, seed * 7919 + (uint64_t)k, 8)) {
            fprintf(stderr, "render failed\n");
            return 1;
        }
        struct vtg_msg m;
        memset(&m, 0, sizeof m);
        m.magic = VTG_MSG_MAGIC;
        m.type = VTG_MSG_FRAME;
        m.dest = VTG_PIPE_PYR | VTG_PIPE_FLOW;
        m.instance = (uint64_t)k + 1;
        m.acq_ns = (uint64_t)llround(tk * 1e9);
        const int set = k % vtg_tk_sets(t);
        int rc = vtg_tk_build(t, set, frame, &m);
        if (!rc)
            rc = vtg_tk_track(t, set, &m, res);
        if (rc) {
            fprintf(stderr, "frame %d: %s\n", k, vtg_tk_error(t));
            return 1;
        }
        dropped += (int)res->dropped;
        tracked += (int)res->tracked;
        for (unsigned i = 0; i < res->nfeat; i++) {
            const struct vtg_feat_out *f = &res->feat[i];
            if (f->flags & VTG_FF_NEW)
                continue;
            double mx, my;
            const double px = (double)f->x - (double)f->u, py = (double)f->y - (double)f->v;
            if (vts_scene_map(scene, w, h, seed, (k - 1) / fps, px, py, tk, &mx, &my) < 0) {
                off_layer++;
                continue;
            }
            epe[nepe++] = hypot((double)f->x - mx, (double)f->y - my);
        }
        /* the caller's seeding: a grid over the object at the first frame, and whenever half the features are gone */
        if (k == 0 || (int)res->nfeat < seeded / 2) {
            const int n = seed_grid(scene, w, h, tk, 10, 8, 14.0, xs, ys);
            if (vtg_tk_set_features(t, n, xs, ys)) {
                fprintf(stderr, "seed: %s\n", vtg_tk_error(t));
                return 1;
            }
            seeded = n;
            reseeds += k > 0;
        }
    }
    qsort(epe, (size_t)nepe, sizeof(double), cmp_double);
    printf("{\"scene\": \"%s\", \"w\": %d, \"h\": %d, \"frames\": %d, \"epe\": {\"n\": %d, \"median\": %.5f, "
           "\"p95\": %.5f, \"max\": %.4f}, \"tracked\": %d, \"dropped\": %d, \"reseeds\": %d, \"off_layer\": %d}\n",
           names[scene], w, h, frames, nepe, pct(epe, nepe, 0.5), pct(epe, nepe, 0.95),
           nepe ? epe[nepe - 1] : -1.0, tracked, dropped, reseeds, off_layer);
    vtg_tk_destroy(t);
    free(epe);
    free(res);
    free(frame);
    return 0;
}
