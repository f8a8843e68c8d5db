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


def _match_inputs(oracle, w, h, d0, seed, n):
    from oracle_ffi import StereoFeature
    lft, rgt = stereo_pair(oracle, w, h, d0, seed)
    rng = np.random.default_rng(seed)
    feats = (StereoFeature * n)()
    for i in range(n):
        feats[i].x = float(rng.uniform(40, w - 20))
        feats[i].y = float(rng.uniform(12, h - 12))
        feats[i].disparity = float(d0 + rng.uniform(-0.6, 0.6)) if i % 4 == 0 else 0.0   # a prior on a quarter
        feats[i].flags = 8 if i % 3 == 0 else 0                                          # some NEW
    return lft, rgt, feats


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref not built (no reference tree here)")
def test_stereo_geometry_and_match_equal_reference(oracle):
    """vo_stereo_triangulate / _reprojection equal the reference's on random inputs (long double bits); vo_stereo_match
    equals the tracker's stereo step composed from the reference's own stereo_scan_vx, lk_track_vx, lk_flowback_vx and
    geometry in track.c:1859-1914's order, every field bit for bit."""
    from oracle_ffi import LK_OUT_DTYPE, LkCfg, Plane, StereoGeom, stereo_cfg
    ref = C.CDLL(REF_SO)
    LD = C.c_longdouble
    ref.stereo_triangulate.restype = C.c_int
    ref.stereo_triangulate.argtypes = [C.POINTER(StereoGeom), LD, LD, LD, LD, C.POINTER(LD)]
    ref.stereo_reprojection.restype = LD
    ref.stereo_reprojection.argtypes = [C.POINTER(StereoGeom), C.POINTER(LD), LD, LD, LD, LD]
    oracle.lib.vo_stereo_triangulate.restype = C.c_int
    oracle.lib.vo_stereo_triangulate.argtypes = ref.stereo_triangulate.argtypes
    oracle.lib.vo_stereo_reprojection.restype = LD
    oracle.lib.vo_stereo_reprojection.argtypes = ref.stereo_reprojection.argtypes
    g = StereoGeom(612.5, 608.25, 320.3, 200.7, 0.1234)
    rng = np.random.default_rng(1)
    for _ in range(500):
        xl, yl = rng.uniform(0, 640), rng.uniform(0, 400)
        xr, yr = xl - rng.uniform(-2, 80), yl + rng.normal(0, 0.5)
        pa, pb = (LD * 3)(), (LD * 3)()
        ra = oracle.lib.vo_stereo_triangulate(C.byref(g), xl, yl, xr, yr, pa)
        rb = ref.stereo_triangulate(C.byref(g), xl, yl, xr, yr, pb)
        assert ra == rb and bytes(pa) == bytes(pb)
        if ra:
            a = oracle.lib.vo_stereo_reprojection(C.byref(g), pa, xl + 0.1, yl, xr, yr - 0.2)
            b = ref.stereo_reprojection(C.byref(g), pb, xl + 0.1, yl, xr, yr - 0.2)
            assert bytes(LD(a)) == bytes(LD(b))
    # the match, composed from the reference's own functions
    w, h, d0 = 640, 400, 24
    lft, rgt, feats = _match_inputs(oracle, w, h, d0, 17, 160)
    Lp, L = oracle.pyramid(lft, w, h, 3)
    Rp, R = oracle.pyramid(rgt, w, h, 3)
    cfg = stereo_cfg(w, h)
    oracle.lib.vo_stereo_match.argtypes = [C.POINTER(Plane), C.POINTER(Plane), C.POINTER(type(cfg)), C.c_int,
                                           C.c_void_p]
    mine = type(feats).from_buffer_copy(feats)
    oracle.lib.vo_stereo_match(L, R, C.byref(cfg), len(mine), C.addressof(mine))
    scan = bind_stereo(ref, "stereo_scan_vx")
    ref.lk_tile_bytes.restype = C.c_size_t
    tile = np.zeros(ref.lk_tile_bytes() // 4 + 64, np.float32)
    tp = (tile.ctypes.data + 63) & ~63
    for fn in ("lk_track_vx",):
        getattr(ref, fn).argtypes = [C.POINTER(Plane), C.POINTER(Plane), C.c_int, C.POINTER(LkCfg), C.c_float,
                                     C.c_float, C.c_float, C.c_float, C.c_void_p, C.c_void_p]
    ref.lk_flowback_vx.argtypes = [C.POINTER(Plane), C.POINTER(Plane), C.c_int, C.POINTER(LkCfg), C.c_float,
                                   C.c_float, C.c_float, C.c_void_p, C.c_void_p]
    f32 = lambda v: float(np.float32(v))
    counts = {"3d": 0, "amb": 0, "none": 0}
    for i in range(len(feats)):
        q = feats[i]
        fl = q.flags & 8
        xr = yr = z = 0.0
        guess, top, disp, done, p = q.disparity, 0, None, False, None
        if guess <= 0:
            sl = cfg.stereo_level
            s = 1.0 / (1 << sl)
            sc = StereoScan()
            scan(C.byref(L[sl]), C.byref(R[sl]), f32(q.x * s), f32(q.y * s), int(cfg.min_disp * s),
                 int(cfg.max_disp * s), C.byref(sc))
            if sc.disparity < 0 or sc.ncc < np.float32(cfg.min_ncc):
                fl, disp, done = fl, -1.0, True
            elif sc.second > sc.ncc * np.longdouble(cfg.ambiguity):
                fl, disp, done = fl | 4, -1.0, True
            else:
                guess, top = f32(sc.disparity * (1 << sl)), sl
        if not done:
            m = np.zeros(1, LK_OUT_DTYPE)
            ref.lk_track_vx(L, R, top, C.byref(cfg.lk), q.x, q.y, f32(-guess), 0.0, m.ctypes.data, tp)
            ref.lk_flowback_vx(L, R, 0, C.byref(cfg.lk), q.x, q.y, cfg.fb_max, m.ctypes.data, tp)
            if m["status"][0]:
                disp = -1.0
            else:
                xr, yr = float(m["x"][0]), float(m["y"][0])
                d = np.longdouble(np.float32(q.x)) - np.longdouble(np.float32(xr))
                disp = float(np.float32(d))
                fl |= 1
                pb = (LD * 3)()
                if (d >= np.longdouble(cfg.min_disp) and d <= np.longdouble(cfg.max_disp)
                        and ref.stereo_triangulate(C.byref(cfg.geom), q.x, q.y, xr, yr, pb)):
                    p = bytes(pb)
                    if ref.stereo_reprojection(C.byref(cfg.geom), pb, q.x, q.y, xr, yr) <= np.longdouble(
                            np.float32(cfg.max_reproj)):
                        fl |= 2
                        z = float(np.float32(pb[2]))
        got = mine[i]
        assert (got.disparity, got.flags, got.xr, got.yr, got.z) == (f32(disp), fl, f32(xr), f32(yr), f32(z)), i
        if p is not None:
            assert bytes(got.p) == p
        counts["3d" if fl & 2 else "amb" if fl & 4 else "none" if disp < 0 else "3d"] += 1
    assert counts["3d"] > 100
