/* synth.c: seeded synthetic frames for the benchmark workloads (see include/vtsynth.h). Pure C. */
#include "vtsynth.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

uint64_t vts_rand(uint64_t *s)
{
    uint64_t x = *s;
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    *s = x;
    return x * 0x2545F4914F6CDD1Dull;
}

double vts_uniform(uint64_t *s) { return (double)(vts_rand(s) >> 11) * (1.0 / 9007199254740992.0); }

static uint64_t seed_state(uint64_t seed)
{
    /* splitmix64 of the seed: nearby seeds give unrelated streams, and the state is never 0 */
    uint64_t z = seed + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return z ? z : 0x1234567ull;
}

static double gauss(uint64_t *s)
{
    /* Box-Muller, one value per call */
    double u1 = vts_uniform(s);
    const double u2 = vts_uniform(s);
    if (u1 < 1e-300)
        u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

struct blob { double cx, cy, inv2s2, amp, reach; };

static struct blob *make_blobs(int w, int h, uint64_t seed, const struct vts_scene *sc, const struct vts_motion *m)
{
    struct blob *b = (struct blob *)malloc(sizeof(struct blob) * (size_t)sc->nblobs);
    uint64_t s = seed_state(seed);
    const double c_x = 0.5 * (w - 1), c_y = 0.5 * (h - 1);
    const double th = m ? m->theta_deg * 3.14159265358979323846 / 180.0 : 0.0;
    const double ct = cos(th), st = sin(th);
    for (int k = 0; k < sc->nblobs; k++) {
        const double x = vts_uniform(&s) * (w - 1), y = vts_uniform(&s) * (h - 1);
        const double sig = sc->sig_lo + (sc->sig_hi - sc->sig_lo) * vts_uniform(&s);
        const double amp = sc->amp_lo + (sc->amp_hi - sc->amp_lo) * vts_uniform(&s);
        double px = x, py = y;
        if (m) {
            px = ct * (x - c_x) - st * (y - c_y) + c_x + m->dx;
            py = st * (x - c_x) + ct * (y - c_y) + c_y + m->dy;
        }
        b[k].cx = px;
        b[k].cy = py;
        b[k].inv2s2 = 1.0 / (2.0 * sig * sig);
        b[k].amp = amp;
        b[k].reach = 5.0 * sig; /* exp(-12.5) x 120 grey < 5e-4 grey beyond this */
    }
    return b;
}

struct render_job {
    uint8_t *f;
    int w, h, pitch, nblobs;
    const struct blob *b;
    double noise;
    uint64_t noise_seed;
    double flat_amp;
    const int *flat;
    int nflat;
    const struct vts_motion *m; /* to place flat rectangles in frame-1 coordinates */
};

static int in_flat(const struct render_job *j, double x, double y)
{
    for (int i = 0; i < j->nflat; i++) {
        const int *r = j->flat + 4 * i;
        if (x >= r[0] && x < r[2] && y >= r[1] && y < r[3])
            return 1;
    }
    return 0;
}

struct render_arg { struct render_job *j; int y0, y1; };

static void *render_rows(void *p)
{
    struct render_arg *a = (struct render_arg *)p;
    const struct render_job *j = a->j;
    double *acc = (double *)calloc((size_t)j->w, sizeof(double));
    const double c_x = 0.5 * (j->w - 1), c_y = 0.5 * (j->h - 1);
    const double th = j->m ? j->m->theta_deg * 3.14159265358979323846 / 180.0 : 0.0;
    const double ct = cos(th), st = sin(th);
    for (int y = a->y0; y < a->y1; y++) {
        memset(acc, 0, sizeof(double) * (size_t)j->w);
        for (int k = 0; k < j->nblobs; k++) {
            const struct blob *b = &j->b[k];
            const double dy = y - b->cy;
            if (fabs(dy) > b->reach)
                continue;
            int x0 = (int)floor(b->cx - b->reach), x1 = (int)ceil(b->cx + b->reach);
            x0 = x0 < 0 ? 0 : x0;
            x1 = x1 > j->w - 1 ? j->w - 1 : x1;
            const double ey = dy * dy;
            for (int x = x0; x <= x1; x++) {
                const double dx = x - b->cx;
                acc[x] += b->amp * exp(-(dx * dx + ey) * b->inv2s2);
            }
        }
        uint64_t ns = seed_state(j->noise_seed ^ (0x5851f42d4c957f2dull * (uint64_t)(y + 1)));
        for (int x = 0; x < j->w; x++) {
            double v = acc[x];
            if (j->nflat && j->flat_amp > 0.0) {
                /* the frame-0 place of this pixel decides whether it is in a flat rectangle */
                double qx = x, qy = y;
                if (j->m) {
                    const double rx = x - c_x - j->m->dx, ry = y - c_y - j->m->dy;
                    qx = ct * rx + st * ry + c_x;
                    qy = -st * rx + ct * ry + c_y;
                }
                if (in_flat(j, qx, qy)) {
                    /* a faint copy of the texture: at most flat_amp grey of contrast */
                    const double scale = j->flat_amp / 255.0;
                    v = 128.0 + scale * (v < 255.0 ? v : 255.0) - 0.5 * j->flat_amp;
                }
            }
            if (j->noise > 0.0)
                v += j->noise * gauss(&ns);
            v = floor(v + 0.5);
            j->f[(size_t)y * (size_t)j->pitch + (size_t)x] = (uint8_t)(v < 0.0 ? 0.0 : v > 255.0 ? 255.0 : v);
        }
    }
    free(acc);
    return NULL;
}

static void render(struct render_job *j, int threads)
{
    if (threads < 1)
        threads = 1;
    if (threads > 64)
        threads = 64;
    pthread_t th[64];
    struct render_arg a[64];
    for (int i = 0; i < threads; i++) {
        a[i].j = j;
        a[i].y0 = (int)((long)j->h * i / threads);
        a[i].y1 = (int)((long)j->h * (i + 1) / threads);
        if (i)
            pthread_create(&th[i], NULL, render_rows, &a[i]);
    }
    render_rows(&a[0]);
    for (int i = 1; i < threads; i++)
        pthread_join(th[i], NULL);
}

int vts_frame(uint8_t *f, int w, int h, int pitch, uint64_t seed, const struct vts_scene *scene,
              const struct vts_motion *motion, double noise_sigma, uint64_t noise_seed, int threads)
{
    struct blob *b = make_blobs(w, h, seed, scene, motion);
    struct render_job j = {f, w, h, pitch, scene->nblobs, b, noise_sigma, noise_seed, 0.0, NULL, 0, motion};
    render(&j, threads);
    free(b);
    return 0;
}

int vts_frame_pair(uint8_t *f0, uint8_t *f1, int w, int h, int pitch, uint64_t seed, const struct vts_scene *scene,
                   const struct vts_motion *motion, double noise_sigma, double flat_amp, const int *flat, int nflat,
                   int threads)
{
    struct blob *b0 = make_blobs(w, h, seed, scene, NULL);
    struct render_job j0 = {f0, w, h, pitch, scene->nblobs, b0, 0.0, 0, flat_amp, flat, nflat, NULL};
    render(&j0, threads);
    free(b0);
    struct blob *b1 = make_blobs(w, h, seed, scene, motion);
    struct render_job j1 = {f1, w, h, pitch, scene->nblobs, b1, noise_sigma, seed ^ 0xa0761d6478bd642full,
                            flat_amp, flat, nflat, motion};
    render(&j1, threads);
    free(b1);
    return 0;
}

void vts_draw_motion(uint64_t seed, double max_shift, double max_rot_deg, struct vts_motion *m)
{
    uint64_t s = seed_state(seed ^ 0x3c6ef372fe94f82bull);
    m->dx = (2.0 * vts_uniform(&s) - 1.0) * max_shift;
    m->dy = (2.0 * vts_uniform(&s) - 1.0) * max_shift;
    m->theta_deg = (2.0 * vts_uniform(&s) - 1.0) * max_rot_deg;
}

void vts_truth(const struct vts_motion *m, int w, int h, double x, double y, double *u, double *v)
{
    const double c_x = 0.5 * (w - 1), c_y = 0.5 * (h - 1);
    const double th = m->theta_deg * 3.14159265358979323846 / 180.0;
    const double ct = cos(th), st = sin(th);
    const double px = ct * (x - c_x) - st * (y - c_y) + c_x + m->dx;
    const double py = st * (x - c_x) + ct * (y - c_y) + c_y + m->dy;
    *u = px - x;
    *v = py - y;
}

void vts_grid(int nx, int ny, int w, int h, float *xs, float *ys)
{
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            xs[j * nx + i] = (float)((i + 0.5) * (double)w / (double)nx);
            ys[j * nx + i] = (float)((j + 0.5) * (double)h / (double)ny);
        }
}

void vts_points(uint64_t seed, int n, int w, int h, int margin, float *xs, float *ys)
{
    uint64_t s = seed_state(seed ^ 0x1b873593ull);
    for (int i = 0; i < n; i++) {
        xs[i] = (float)(margin + vts_uniform(&s) * (w - 1 - 2 * margin));
        ys[i] = (float)(margin + vts_uniform(&s) * (h - 1 - 2 * margin));
    }
}

/* ------------------------------------------------------------------------------------------------------------------
 * Layered scenes with analytic motion. A layer is a texture in its own coordinates, placed at an offset that is a
 * closed-form function of time, and seen through a mask (none, a rectangle or a disc) in those coordinates; layers are
 * drawn back to front, so a pixel shows the topmost layer whose mask holds it. A texture is fractal value noise: random
 * grey on an integer lattice, interpolated with smoothstep weights (a continuous slope, so sub-pixel motion moves the
 * picture smoothly), summed over three octaves of wavelength wl, wl / 2, wl / 4 with weights 0.5, 0.3, 0.2, and mapped
 * to base + gain x (0 .. 255). Sizes are given for a 640-wide frame and scale with the width. */

#define SC_LAYERS 4
struct sc_layer {
    int on;
    int mask;            /* 0 none, 1 rectangle, 2 disc */
    double hx, hy;       /* half sizes (radius in hx for a disc) */
    double base, gain;   /* grey = base + gain x noise (0 .. 255) */
    double wl;           /* the first octave's wavelength, pixels */
    uint64_t seed;
};

struct sc_def {
    int id, w, h;
    double s;            /* w / 640 */
    struct sc_layer L[SC_LAYERS];
};

/* The offset of layer k at time t (seconds), and the brightness transform and noise of the frame. */
static void sc_offset(const struct sc_def *d, int k, double t, double *ox, double *oy)
{
    const double s = d->s, cx = 0.5 * d->w, cy = 0.5 * d->h;
    *ox = 0.0;
    *oy = 0.0;
    switch (d->id) {
    case VTS_SC_TRANSLATE:
        if (k == 1)
            *ox = cx - 144.0 * s + 48.0 * s * t, *oy = cy - 32.0 * s + 17.6 * s * t;
        break;
    case VTS_SC_CAMERA: {
        /* the camera moves: every layer goes with it */
        const double mx = 128.0 * s * sin(0.8 * t), my = 48.0 * s * sin(1.3 * t);
        *ox = mx + (k == 1 ? cx + 32.0 * s : 0.0);
        *oy = my + (k == 1 ? cy : 0.0);
        break;
    }
    case VTS_SC_SHAPES:
        if (k == 1)
            *ox = cx + 144.0 * s * sin(0.9 * t), *oy = cy + 64.0 * s * sin(1.7 * t);
        else if (k == 2)
            *ox = cx + 144.0 * s * cos(0.5 * t), *oy = cy + 150.0 * s;
        break;
    case VTS_SC_BRIGHTNESS:
        if (k == 1)
            *ox = cx - 96.0 * s + 32.0 * s * t, *oy = cy;
        break;
    case VTS_SC_NOISE:
        if (k == 1)
            *ox = cx - 96.0 * s + 32.0 * s * t, *oy = cy + 24.0 * s * sin(1.1 * t);
        break;
    case VTS_SC_OCCLUSION:
        if (k == 1)
            *ox = cx - 64.0 * s + 12.8 * s * t, *oy = cy;
        else if (k == 3)
            *ox = cx - 256.0 * s + 88.0 * s * t, *oy = cy - 32.0 * s;
        break;
    }
}

static void sc_light(int id, double t, double *gain, double *offset, double *noise)
{
    *gain = 1.0;
    *offset = 0.0;
    *noise = 0.0;
    if (id == VTS_SC_BRIGHTNESS) {
        *gain = 1.0 + 0.25 * sin(1.3 * t);
        *offset = 16.0 * sin(0.7 * t);
    }
    if (id == VTS_SC_NOISE)
        *noise = 4.0;
}

/* The lattice's grey at (i, j) for a seed: 0 .. 255 from a splitmix64 mix of the three. */
static double sc_lattice(int64_t i, int64_t j, uint64_t seed)
{
    uint64_t z = (uint64_t)i * 0xd6e8feb86659fd93ull + (uint64_t)j * 0xa0761d6478bd642full + seed * 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z & 255u);
}

static double sc_vnoise(double x, double y, uint64_t seed)
{
    const double x0 = floor(x), y0 = floor(y);
    const int64_t i = (int64_t)x0, j = (int64_t)y0;
    const double fx = x - x0, fy = y - y0;
    const double sx = fx * fx * (3.0 - 2.0 * fx), sy = fy * fy * (3.0 - 2.0 * fy);
    const double a = sc_lattice(i, j, seed), b = sc_lattice(i + 1, j, seed);
    const double c = sc_lattice(i, j + 1, seed), d = sc_lattice(i + 1, j + 1, seed);
    const double top = a + (b - a) * sx, bottom = c + (d - c) * sx;
    return top + (bottom - top) * sy;
}

static int sc_make(int id, int w, int h, uint64_t seed, struct sc_def *d)
{
    if (id < 0 || id >= VTS_SC_COUNT || w < 64 || h < 64)
        return -1;
    memset(d, 0, sizeof *d);
    d->id = id;
    d->w = w;
    d->h = h;
    const double s = d->s = w / 640.0;
    /* the background: wider than the frame by the camera's reach */
    /* the background: a softer, lower-contrast texture */
    d->L[0].on = 1;
    d->L[0].base = 40.0;
    d->L[0].gain = 0.6;
    d->L[0].wl = 72.0 * s;
    d->L[0].seed = seed ^ 0x11;
    /* the object */
    d->L[1].on = 1;
    d->L[1].mask = id == VTS_SC_SHAPES ? 2 : 1;
    d->L[1].hx = (id == VTS_SC_SHAPES ? 64.0 : 80.0) * s;
    d->L[1].hy = 64.0 * s;
    d->L[1].gain = 1.0;
    d->L[1].wl = 16.0 * s;
    d->L[1].seed = seed ^ 0x22;
    if (id == VTS_SC_SHAPES) { /* a second mover */
        d->L[2].on = 1;
        d->L[2].mask = 1;
        d->L[2].hx = 40.0 * s;
        d->L[2].hy = 32.0 * s;
        d->L[2].gain = 1.0;
        d->L[2].wl = 14.0 * s;
        d->L[2].seed = seed ^ 0x33;
    }
    if (id == VTS_SC_OCCLUSION) { /* a bar as wide as half the object, over its upper half, nearer the camera */
        d->L[3].on = 1;
        d->L[3].mask = 1;
        d->L[3].hx = 40.0 * s;
        d->L[3].hy = 38.0 * s;
        d->L[3].gain = 1.0;
        d->L[3].wl = 12.0 * s;
        d->L[3].seed = seed ^ 0x44;
    }
    return 0;
}

static void sc_free(struct sc_def *d) { (void)d; }

static int sc_inside(const struct sc_layer *l, double lx, double ly)
{
    if (l->mask == 1)
        return fabs(lx) <= l->hx && fabs(ly) <= l->hy;
    if (l->mask == 2)
        return lx * lx + ly * ly <= l->hx * l->hx;
    return 1;
}

/* The topmost layer holding (x, y) at time t, or -1. */
static int sc_top(const struct sc_def *d, double t, double x, double y)
{
    for (int k = SC_LAYERS - 1; k >= 0; k--) {
        if (!d->L[k].on)
            continue;
        double ox, oy;
        sc_offset(d, k, t, &ox, &oy);
        if (sc_inside(&d->L[k], x - ox, y - oy))
            return k;
    }
    return -1;
}

static double sc_value(const struct sc_layer *l, double lx, double ly)
{
    const double n = 0.5 * sc_vnoise(lx / l->wl, ly / l->wl, l->seed) +
                     0.3 * sc_vnoise(lx / (0.5 * l->wl), ly / (0.5 * l->wl), l->seed + 1) +
                     0.2 * sc_vnoise(lx / (0.25 * l->wl), ly / (0.25 * l->wl), l->seed + 2);
    return l->base + l->gain * n;
}

struct sc_rows {
    const struct sc_def *d;
    double t;
    uint8_t *f;
    int pitch, y0, y1;
    uint64_t noise_seed;
};

static void *sc_render_rows(void *p)
{
    const struct sc_rows *a = (const struct sc_rows *)p;
    const struct sc_def *d = a->d;
    double ox[SC_LAYERS], oy[SC_LAYERS], gain, offset, noise;
    for (int k = 0; k < SC_LAYERS; k++)
        sc_offset(d, k, a->t, &ox[k], &oy[k]);
    sc_light(d->id, a->t, &gain, &offset, &noise);
    for (int y = a->y0; y < a->y1; y++) {
        uint64_t ns = seed_state(a->noise_seed ^ (0x5851f42d4c957f2dull * (uint64_t)(y + 1)));
        for (int x = 0; x < d->w; x++) {
            double v = 0.0;
            for (int k = SC_LAYERS - 1; k >= 0; k--) {
                if (!d->L[k].on || !sc_inside(&d->L[k], x - ox[k], y - oy[k]))
                    continue;
                v = sc_value(&d->L[k], x - ox[k], y - oy[k]);
                break;
            }
            v = gain * v + offset;
            if (noise > 0.0)
                v += noise * gauss(&ns);
            v = floor(v + 0.5);
            a->f[(size_t)y * (size_t)a->pitch + (size_t)x] = (uint8_t)(v < 0.0 ? 0.0 : v > 255.0 ? 255.0 : v);
        }
    }
    return NULL;
}

int vts_scene_frame(int scene, double t, uint8_t *f, int w, int h, int pitch, uint64_t seed, uint64_t noise_seed,
                    int threads)
{
    struct sc_def d;
    if (!f || pitch < w || sc_make(scene, w, h, seed, &d))
        return -1;
    if (threads < 1)
        threads = 1;
    if (threads > 64)
        threads = 64;
    pthread_t th[64];
    struct sc_rows a[64];
    for (int i = 0; i < threads; i++) {
        a[i].d = &d;
        a[i].t = t;
        a[i].f = f;
        a[i].pitch = pitch;
        a[i].noise_seed = noise_seed;
        a[i].y0 = (int)((long)h * i / threads);
        a[i].y1 = (int)((long)h * (i + 1) / threads);
        if (i)
            pthread_create(&th[i], NULL, sc_render_rows, &a[i]);
    }
    sc_render_rows(&a[0]);
    for (int i = 1; i < threads; i++)
        pthread_join(th[i], NULL);
    sc_free(&d);
    return 0;
}

int vts_scene_map(int scene, int w, int h, uint64_t seed, double t0, double x, double y, double t1, double *mx,
                  double *my)
{
    struct sc_def d;
    if (!mx || !my || sc_make(scene, w, h, seed, &d))
        return -1;
    const int k = sc_top(&d, t0, x, y);
    if (k >= 0) {
        double ax, ay, bx, by;
        sc_offset(&d, k, t0, &ax, &ay);
        sc_offset(&d, k, t1, &bx, &by);
        *mx = x - ax + bx;
        *my = y - ay + by;
    }
    sc_free(&d);
    return k;
}

int vts_scene_object(int scene, int w, int h, double t, double *cx, double *cy, double *hx, double *hy)
{
    struct sc_def d;
    if (!cx || !cy || !hx || !hy || sc_make(scene, w, h, 1, &d))
        return -1;
    sc_offset(&d, 1, t, cx, cy);
    *hx = d.L[1].hx;
    *hy = d.L[1].hy;
    sc_free(&d);
    return 0;
}
