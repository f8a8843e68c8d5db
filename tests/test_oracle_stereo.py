"""The oracle's stereo row scan (vo_stereo_scan, a restatement of src/kern/stereo.c:1-80) pinned: against the
reference's own stereo_scan_vx compiled from source into oracle/_ref (disparity, correlation and second peak bit for
bit over random features, ranges wider than the 512-candidate cap, candidates off the right image, an empty range),
and by closed forms (a right image that is the left shifted by d0 peaks at d0 with correlation 1; a flat template
correlates with nothing)."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle_ffi import REF_SO, StereoScan, bind_stereo
from test_oracle_still import q6_plane


def stereo_pair(o, w, h, d0, seed, noise=1.5):
    """A rectified pair: the right image is the left scene moved d0 px left (m2_render offsets are quarter pixels),
    each with its own noise."""
    rng = np.random.default_rng(seed)
    left = o.m2_render(w, h, 500, 300)[:, :w].astype(np.float64)
    right = o.m2_render(w, h, 500 + 4 * d0, 300)[:, :w].astype(np.float64)
    left += rng.normal(0.0, noise, left.shape)
    right += rng.normal(0.0, noise, right.shape)
    return (np.clip(left.round(), 0, 255).astype(np.uint8), np.clip(right.round(), 0, 255).astype(np.uint8))


def features(rng, w, h, n):
    return [(float(rng.uniform(12, w - 12)), float(rng.uniform(8, h - 8))) for _ in range(n)]


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref not built (no reference tree here)")
def test_stereo_scan_equals_reference(oracle):
    ref = C.CDLL(REF_SO)
    theirs_fn = bind_stereo(ref, "stereo_scan_vx")
    mine_fn = bind_stereo(oracle.lib, "vo_stereo_scan")
    w, h = 320, 200
    lft, rgt = stereo_pair(oracle, w, h, 13, 3)
    L, R = q6_plane(oracle, lft), q6_plane(oracle, rgt)
    rng = np.random.default_rng(8)
    ranges = [(0, 40), (-5, 30), (0, 700), (10, 10), (5, 4), (200, 260)]
    n = 0
    for (x, y) in features(rng, w, h, 120):
        for dmin, dmax in ranges:
            a, b = StereoScan(), StereoScan()
            mine_fn(C.byref(L.c), C.byref(R.c), x, y, dmin, dmax, C.byref(a))
            theirs_fn(C.byref(L.c), C.byref(R.c), x, y, dmin, dmax, C.byref(b))
            assert np.array(a.tup(), np.float32).tobytes() == np.array(b.tup(), np.float32).tobytes(), \
                (x, y, dmin, dmax, a.tup(), b.tup())
            n += 1
    assert n == 720


def test_stereo_scan_closed_forms(oracle):
    fn = bind_stereo(oracle.lib, "vo_stereo_scan")
    w, h = 320, 200
    lft, _ = stereo_pair(oracle, w, h, 0, 4, noise=0.0)
    d0 = 17
    rgt = np.zeros_like(lft)
    rgt[:, :w - d0] = lft[:, d0:]              # right x = left x - d0
    L, R = q6_plane(oracle, lft), q6_plane(oracle, rgt)
    rng = np.random.default_rng(5)
    hits = 0
    for (x, y) in features(rng, w - 40, h, 60):
        x += 30.0
        s = StereoScan()
        fn(C.byref(L.c), C.byref(R.c), x, y, 0, 40, C.byref(s))
        if s.disparity == d0:
            hits += 1
            assert abs(s.ncc - 1.0) < 1e-5
    assert hits >= 55
    flat = q6_plane(oracle, np.full((h, w), 77, np.uint8))
    s = StereoScan()
    fn(C.byref(flat.c), C.byref(R.c), 100.0, 100.0, 0, 40, C.byref(s))
    assert s.ncc == 0.0
    s = StereoScan()
    fn(C.byref(L.c), C.byref(R.c), 100.0, 100.0, 5, 4, C.byref(s))   # empty range
    assert (s.disparity, s.ncc, s.second) == (-1.0, 0.0, -1.0)
