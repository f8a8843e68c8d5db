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
