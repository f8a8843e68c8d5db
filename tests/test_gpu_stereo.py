"""The stereo row scan on the GPU (vtg_stereo_scan_many / _planes, src/kern/stereo.c:36-80) against the oracle's
restatement (itself pinned to the reference's stereo_scan_vx in test_oracle_stereo.py): for every feature the same
disparity, and the correlation and second peak within 1e-5 (the three sums are the reference's bits; the correlation
from them is double here, long double there). Features over the whole level, at its edges, at level 0 and 1; ranges
past the 512-candidate cap, off the right image, and empty; Q6, F32 and host-plane inputs agree."""
import ctypes as C

import numpy as np
import pytest

import vtgpu
from gpu_util import say
from oracle_ffi import StereoScan, bind_stereo
from test_oracle_stereo import stereo_pair

pytestmark = pytest.mark.gpu

RANGES = [(0, 40), (-5, 30), (0, 700), (10, 10), (5, 4), (150, 260)]


def _features(rng, w, h, n):
    xs = rng.uniform(0.0, w - 1.0, n).astype(np.float32)
    ys = rng.uniform(1.0, h - 2.0, n).astype(np.float32)
    xs[:8] = [0.0, 3.5, 7.9, w - 1.0, w - 4.2, 20.0, 40.0, 60.0]   # template off the left and right edges
    ys[:8] = [1.0, 2.5, 4.0, h - 2.0, h - 3.5, 1.0, h - 2.0, 6.0]
    return xs, ys


def _oracle_scans(oracle, L, R, xs, ys, dmin, dmax):
    fn = bind_stereo(oracle.lib, "vo_stereo_scan")
    out = np.zeros((len(xs), 3), np.float32)
    for i in range(len(xs)):
        s = StereoScan()
        fn(C.byref(L.c), C.byref(R.c), float(xs[i]), float(ys[i]), dmin, dmax, C.byref(s))
        out[i] = s.tup()
    return out


@pytest.mark.parametrize("lv", [0, 1])
def test_stereo_scan_matches_oracle(ctx, oracle, lv):
    W, H = 640, 400
    lft, rgt = stereo_pair(oracle, W, H, 24, 6)
    pl, pr = ctx.pyramid(W, H, lv + 1), ctx.pyramid(W, H, lv + 1)
    pl.build(lft)
    pr.build(rgt)
    Lo, _ = oracle.pyramid(lft, W, H, lv + 1)
    Ro, _ = oracle.pyramid(rgt, W, H, lv + 1)
    L, R = Lo[lv], Ro[lv]
    rng = np.random.default_rng(12 + lv)
    xs, ys = _features(rng, L.w, L.h, 400)
    xy = ctx.to_device(np.stack([xs, ys], 1).ravel())
    agree = 0
    for dmin, dmax in RANGES:
        say("stereo scan", lv, dmin, dmax)
        got = ctx.stereo_scan_many(pl, pr, lv, xy, dmin, dmax)
        ref = _oracle_scans(oracle, L, R, xs, ys, dmin, dmax)
        assert np.array_equal(got["disparity"], ref[:, 0]), (dmin, dmax)
        assert np.abs(got["ncc"] - ref[:, 1]).max() <= 1e-5
        assert np.abs(got["second"] - ref[:, 2]).max() <= 1e-5
        agree += int((got["ncc"] == ref[:, 1]).sum())
        if (dmin, dmax) == (0, 40):
            peak = got["ncc"] > 0.9
            assert peak.mean() > 0.5 and np.median(got["disparity"][peak]) == round(24 / 2 ** lv)
    say(f"stereo scan lv{lv}: {len(xs)} features x {len(RANGES)} ranges, correlation bit-equal on {agree}")
    xy.free()
    pl.destroy()
    pr.destroy()


def test_stereo_scan_storages_and_planes(ctx, oracle):
    """F32 pyramids and the reference's host planes give the Q6 pyramid's results bit for bit."""
    W, H = 320, 240
    lft, rgt = stereo_pair(oracle, W, H, 9, 7)
    q = [ctx.pyramid(W, H, 1) for _ in range(2)]
    f = [ctx.pyramid(W, H, 1, vtgpu.STORE_F32) for _ in range(2)]
    for p, img in zip(q + f, [lft, rgt, lft, rgt]):
        p.build(img)
    rng = np.random.default_rng(3)
    xs, ys = _features(rng, W, H, 200)
    xy = ctx.to_device(np.stack([xs, ys], 1).ravel())
    a = ctx.stereo_scan_many(q[0], q[1], 0, xy, -3, 30)
    b = ctx.stereo_scan_many(f[0], f[1], 0, xy, -3, 30)
    Lo, _ = oracle.pyramid(lft, W, H, 1)
    Ro, _ = oracle.pyramid(rgt, W, H, 1)
    c = ctx.stereo_scan_planes(Lo[0], Ro[0], xs, ys, -3, 30)
    assert a.tobytes() == b.tobytes() == c.tobytes()
    with pytest.raises(vtgpu.VtgError):
        ctx.stereo_scan_many(q[0], f[1], 0, xy, 0, 10)   # storages differ
    xy.free()
    for p in q + f:
        p.destroy()


@pytest.mark.parametrize("level,fb", [(1, 0.5), (0, 0.0)])
def test_stereo_match_matches_oracle(ctx, oracle, level, fb):
    """vtg_stereo_match against vo_stereo_match (pinned to the reference's functions in the tracker's order): every
    feature equal bit for bit (flags with a prior on a quarter of them and NEW kept, match, disparity, depth) to the
    oracle with its lane64 LK (the fp64 driver the GPU runs), which differs from the x87 lane tier on at most 1 in
    1,000 features; scan rejections and ambiguity decisions identical."""
    from oracle_ffi import Plane, stereo_cfg
    from test_oracle_stereo import _match_inputs
    W, H, d0 = 640, 400, 24
    lft, rgt, feats = _match_inputs(oracle, W, H, d0, 23, 300)
    cfg = stereo_cfg(W, H, level=level, max_disp=60.0 if level == 0 else 100.0, fb_max=fb)
    levels = level + 1
    pl, pr = ctx.pyramid(W, H, levels), ctx.pyramid(W, H, levels)
    pl.build(lft)
    pr.build(rgt)
    _, L = oracle.pyramid(lft, W, H, levels)
    _, R = oracle.pyramid(rgt, W, H, levels)
    mine = type(feats).from_buffer_copy(feats)
    theirs = type(feats).from_buffer_copy(feats)
    say("stereo match", level, fb)
    ctx.stereo_match(pl, pr, cfg, mine)
    oracle.lib.vo_stereo_match_tier.argtypes = [C.c_int, C.POINTER(Plane), C.POINTER(Plane), C.POINTER(type(cfg)),
                                                C.c_int, C.c_void_p]
    x87 = type(feats).from_buffer_copy(feats)
    oracle.lib.vo_stereo_match_tier(2, L, R, C.byref(cfg), len(theirs), C.addressof(theirs))
    oracle.lib.vo_stereo_match_tier(0, L, R, C.byref(cfg), len(x87), C.addressof(x87))
    apart = sum(bytes(theirs[i]) != bytes(x87[i]) for i in range(len(feats)))
    assert apart <= max(1, len(feats) // 1000), apart
    same = [mine[i].flags == theirs[i].flags for i in range(len(feats))]
    assert all(same), [i for i in range(len(feats)) if not same[i]][:8]
    worst, n3 = 0.0, 0
    for i in range(len(feats)):
        a, b = mine[i], theirs[i]
        if not same[i]:
            continue
        if a.flags & 1:
            worst = max(worst, abs(a.xr - b.xr), abs(a.yr - b.yr), abs(a.disparity - b.disparity))
        else:
            assert a.disparity == b.disparity == -1.0 and a.xr == b.xr == 0.0
        if a.flags & 2:
            n3 += 1
        assert bytes(a) == bytes(b), i
    assert worst == 0.0 and n3 > 150
    say(f"stereo match level {level}: {len(feats)} features, flags equal on {np.mean(same):.4f}, 3D {n3}, "
        f"max |d| {worst:.2e} px, x87 driver apart on {apart}")
    pl.destroy()
    pr.destroy()
