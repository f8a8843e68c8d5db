/* vtsynth.h: seeded synthetic frame pairs for the benchmark workloads. Pure C, no GPU.
 *
 * A frame is a sum of Gaussian blobs (centre uniform over the frame, sigma uniform in [sig_lo, sig_hi] px, amplitude
 * uniform in [amp_lo, amp_hi] grey levels), rounded and clipped to 0..255. The second frame of a pair holds the same
 * blobs moved by a rigid motion about the frame centre c = ((w-1)/2, (h-1)/2):
 *     p1 = R(theta) (p0 - c) + c + (dx, dy)
 * and rendered analytically (a rigid map keeps every blob a blob, so there is no resampling error), plus N(0, noise^2)
 * grey levels of noise. vts_truth gives the exact displacement of any point of frame 0. */
#ifndef VTSYNTH_H
#define VTSYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct vts_scene {
    int nblobs;             /* 200 in every workload */
    double sig_lo, sig_hi;  /* 4 .. 40 px */
    double amp_lo, amp_hi;  /* 20 .. 120 grey */
};

struct vts_motion {
    double dx, dy;     /* px */
    double theta_deg;  /* rotation about the frame centre */
};

/* Frame 0 and frame 1 (w x h bytes, rows `pitch` apart). The blobs come from `seed`; motion is given (draw it with
 * vts_draw_motion). noise_sigma is the grey-level noise added to frame 1 only (frame 0 is noiseless). flat_amp > 0
 * scales the blob amplitudes inside the rectangles flat[4*i .. 4*i+3] = (x0, y0, x1, y1) down to at most flat_amp
 * grey (near-flat regions); nflat rectangles. threads >= 1 splits rows. Returns 0. */
int vts_frame_pair(uint8_t *f0, uint8_t *f1, int w, int h, int pitch, uint64_t seed, const struct vts_scene *scene,
                   const struct vts_motion *motion, double noise_sigma, double flat_amp, const int *flat, int nflat,
                   int threads);

/* One frame of a sequence: the seed's blobs moved by `motion`, with noise from noise_seed (noise_sigma 0: none). */
int vts_frame(uint8_t *f, int w, int h, int pitch, uint64_t seed, const struct vts_scene *scene,
              const struct vts_motion *motion, double noise_sigma, uint64_t noise_seed, int threads);

/* dx, dy ~ U[-max_shift, max_shift], theta ~ U[-max_rot, max_rot] degrees, from seed. */
void vts_draw_motion(uint64_t seed, double max_shift, double max_rot_deg, struct vts_motion *m);

/* Exact displacement (u, v) of frame-0 point (x, y) under the motion. */
void vts_truth(const struct vts_motion *m, int w, int h, double x, double y, double *u, double *v);

/* nx x ny grid at ((i + 1/2) w / nx, (j + 1/2) h / ny), row-major: the tracking points of the plain workloads. */
void vts_grid(int nx, int ny, int w, int h, float *xs, float *ys);

/* n points uniform in [margin, w - 1 - margin] x [margin, h - 1 - margin], from seed. */
void vts_points(uint64_t seed, int n, int w, int h, int margin, float *xs, float *ys);

/* Layered scenes with analytic motion (the scene set of the reference's tracker tolerances, tests/scenes.md:11-39, as
 * image-plane motions): a still background and movers, each a fractal value-noise texture (wavelengths 72 px for the
 * background at 60 % contrast, 12-16 px for the movers) moving by a closed-form function of time (seconds), drawn back
 * to front, so a pixel shows the topmost layer holding it; sizes are for a 640-wide frame and scale with the width.
 *   TRANSLATE   an 160 x 128 object moving 48, 17.6 px/s over a still background
 *   CAMERA      the camera moves (128 sin 0.8t, 48 sin 1.3t) px: every layer goes with it
 *   SHAPES      a disc of radius 64 on a Lissajous path, and a second 80 x 64 mover
 *   BRIGHTNESS  the object at 32 px/s under a frame gain 1 + 0.25 sin 1.3t and offset 16 sin 0.7t grey
 *   NOISE       the object at 32 px/s, 24 sin 1.1t px, under N(0, 4^2) grey of noise
 *   OCCLUSION   the object at 12.8 px/s, a bar as wide as half of it sweeping over its upper half at 88 px/s, nearer */
enum { VTS_SC_TRANSLATE, VTS_SC_CAMERA, VTS_SC_SHAPES, VTS_SC_BRIGHTNESS, VTS_SC_NOISE, VTS_SC_OCCLUSION, VTS_SC_COUNT };
/* The scene's frame at time t (w x h bytes, rows `pitch` apart): layers from `seed`, noise (NOISE only) from noise_seed.
 * 0, or -1 for a bad argument. */
int vts_scene_frame(int scene, double t, uint8_t *f, int w, int h, int pitch, uint64_t seed, uint64_t noise_seed,
                    int threads);
/* Where the point (x, y) of the frame at time t0 is at time t1: moved with the topmost layer holding it at t0.
 * Returns that layer (0 background, 1 object, 2 second mover, 3 bar), or -1 for a bad argument. */
int vts_scene_map(int scene, int w, int h, uint64_t seed, double t0, double x, double y, double t1, double *mx,
                  double *my);
/* The object's centre and half sizes at time t (a disc: the radius in both). */
int vts_scene_object(int scene, int w, int h, double t, double *cx, double *cy, double *hx, double *hy);

/* The generator's xorshift64* stream, exposed so callers draw sizes and counts from the same source. */
uint64_t vts_rand(uint64_t *s);
double vts_uniform(uint64_t *s); /* [0, 1) */

#ifdef __cplusplus
}
#endif

#endif
