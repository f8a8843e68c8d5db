"""The planes and templates of the NCC refinement fixtures (tests/golden/ncc_refine.json): built the same way by
tests/golden/make_golden.py, which scores them with the reference's own search functions, and by the tests, which
score them with the oracle and the GPU. Deterministic: numpy's PCG64 from fixed seeds and the oracle's m2 renderer."""
import ctypes as C

import numpy as np

from oracle_ffi import HostPlane

TW, TH = 32, 21       # the reference's appearance template (track.c:21-22)
REGION = (40.0, 30.0)  # a region small enough that store_appearance keeps level 0


def smooth_texture(w, h, seed, box=3):
    """u8 texture: uniform noise box-filtered box x box (wrapping), stretched to 0..255."""
    rng = np.random.default_rng(seed)
    a = rng.random((h, w))
    for ax in (0, 1):
        a = sum(np.roll(a, s, axis=ax) for s in range(-(box // 2), box // 2 + 1))
    a -= a.min()
    return np.floor(a * (255.0 / a.max()) + 0.5).astype(np.uint8)


def planted(w, h, seed, at, to, size=(44, 33), gain=1.0, noise=0.0):
    """A texture with the patch at corner `at` copied to corner `to` (gain, then additive noise, clipped)."""
    img = smooth_texture(w, h, seed).astype(np.float64)
    (ax, ay), (tx, ty), (pw, ph) = at, to, size
    patch = img[ay:ay + ph, ax:ax + pw] * gain
    if noise:
        patch = patch + np.random.default_rng(seed + 1).normal(0.0, noise, patch.shape)
    img[ty:ty + ph, tx:tx + pw] = patch
    return np.clip(np.floor(img + 0.5), 0, 255).astype(np.uint8)


def half_shifted(w, h, seed):
    """A texture and its copy moved right by half a pixel (each pixel the mean of two neighbours, exact in Q6 when the
    pair sums are even): (original, shifted) u8 images."""
    img = smooth_texture(w, h, seed) & 0xFE  # even bytes: (a + b) / 2 is a whole byte
    sh = img.copy()
    sh[:, :-1] = ((img[:, :-1].astype(np.int32) + img[:, 1:]) // 2).astype(np.uint8)
    return img, sh


def host_plane(oracle, img):
    """A padded Q6 host plane (64 x the bytes, clamped borders) from a u8 image."""
    h, w = img.shape
    hp = HostPlane(w, h)
    hp.view()[:] = img.astype(np.float32) * 64
    oracle.lib.vo_pad_rows(C.byref(hp.c), 0, h)
    oracle.lib.vo_pad_tb(C.byref(hp.c))
    return hp


def cases(oracle):
    """(name, u8 image, region centre) for each fixture case."""
    out = []
    w, h = 320, 200
    out.append(("m2", oracle.m2_render(w, h, 1000, 700)[:, :w].copy(), (120.5, 90.25)))
    out.append(("planted", planted(288, 176, 7, (60, 50), (180, 100)), (82.0, 66.0)))
    out.append(("planted_dim", planted(288, 176, 8, (60, 50), (170, 40), gain=0.8, noise=3.0), (82.0, 66.0)))
    out.append(("repeat", np.tile(smooth_texture(24, 20, 9), (9, 16)), (130.0, 80.0)))
    return out


def windows(w, h, corner):
    """The search windows of a case: (x0, x1, y0, y1, step). The last three end at the template's own corner (x, y)
    and start step - 1 before it, so the only coarse corner is step - 1 pixels off in both axes and the refinement
    must reach 2 step - 2 half pixels to find the thing."""
    x, y = corner
    return [(1, w, 1, h, 1), (1, w, 1, h, 2), (1, w, 1, h, 3), (40, 150, 30, 120, 1), (-20, 90, -9, 70, 2),
            (w - 60, w + 40, h - 40, h + 30, 1), (100, 99, 10, 60, 1)] + \
        [(x - s + 1, x, y - s + 1, y, s) for s in (2, 3, 4)]
