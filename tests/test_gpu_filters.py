"""The level-0 filters on the GPU (row a5: img_gauss5 / img_norm / img_minmax, src/kern/img_ref.c:42-96, and tk_build's
contrast / pre_gauss options, src/kern/track.c:600-623), bit for bit: the host-plane calls against the reference's own
m1 hash lines and the oracle, the filtered pyramids level by level against the oracle's restatement of tk_build, and
LK on a filtered pyramid against the oracle's LK on the same levels."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import vtgpu
from gpu_util import conditioning, lane_ref, say
from oracle_ffi import HostPlane, lk_cfg

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _pad(o, hp):
    o.lib.vo_pad_rows(C.byref(hp.c), 0, hp.h)
    o.lib.vo_pad_tb(C.byref(hp.c))


def _l0(o, img, w, h):
    a = HostPlane(w, h)
    o.lib.vo_u8_to_q6(C.byref(a.c), img.ctypes.data, img.strides[0], 0, h)
    _pad(o, a)
    return a


def _hash(o, hp):
    return "%016x" % o.hash_plane(hp)


@pytest.mark.parametrize("size", ["640x480", "1280x720", "1920x1080"])
def test_host_plane_filters_match_m1_lines(ctx, oracle, size):
    """vtg_img_norm(0.75, 512), vtg_img_minmax and vtg_img_gauss5 on the m1 input give the reference's hash lines
    (logs/m1/host/m1-host-<size>.txt)."""
    gold = json.load(open(os.path.join(GOLD, "m1_hashes.json")))[size]
    w, h = map(int, size.split("x"))
    a = _l0(oracle, oracle.m1_image(w, h), w, h)
    say("filters m1", size)
    c = HostPlane(w, h)
    go = np.array([0.75, 512.0], np.float32)
    ctx.check(vtgpu.lib().vtg_img_norm(ctx.h, C.byref(c.c), C.byref(a.c), go.ctypes.data, 0, h), "img_norm")
    assert _hash(oracle, c) == gold["norm"]
    mm = np.zeros(2, np.float32)
    ctx.check(vtgpu.lib().vtg_img_minmax(ctx.h, C.byref(a.c), mm.ctypes.data, 0, h), "img_minmax")
    assert list(mm) == gold["minmax"]
    ctx.check(vtgpu.lib().vtg_img_gauss5(ctx.h, C.byref(c.c), C.byref(a.c), 0, h), "img_gauss5")
    assert _hash(oracle, c) == gold["gauss5"]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_host_plane_norm_and_rows(ctx, oracle, seed):
    """Random gains and offsets (the long double sum's rounding exercised on every pixel), and a row range: the rows
    written (with their padding) equal the oracle's, the other rows are untouched."""
    rng = np.random.default_rng(seed)
    w, h = 517, 203
    f0, _, _ = vtgpu.frame_pair(w, h, 40 + seed)
    a = _l0(oracle, f0, w, h)
    go = np.array([rng.uniform(0.1, 40.0), rng.uniform(-20000.0, 5000.0)], np.float32)
    say("norm random", seed, go)
    ref, got = HostPlane(w, h), HostPlane(w, h)
    oracle.lib.vo_norm(C.byref(ref.c), C.byref(a.c), go.ctypes.data, 0, h)
    oracle.lib.vo_pad_rows(C.byref(ref.c), 0, h)
    ctx.check(vtgpu.lib().vtg_img_norm(ctx.h, C.byref(got.c), C.byref(a.c), go.ctypes.data, 0, h), "img_norm")
    assert np.array_equal(got.padded()[4:4 + h], ref.padded()[4:4 + h])
    y0, y1 = 37, 151
    part = HostPlane(w, h)
    part.padded()[:] = -7.0
    ctx.check(vtgpu.lib().vtg_img_gauss5(ctx.h, C.byref(part.c), C.byref(a.c), y0, y1), "img_gauss5")
    full = HostPlane(w, h)
    oracle.lib.vo_gauss5(C.byref(full.c), C.byref(a.c), 0, h)
    oracle.lib.vo_pad_rows(C.byref(full.c), 0, h)
    P, F = part.padded(), full.padded()
    assert np.array_equal(P[4 + y0:4 + y1], F[4 + y0:4 + y1])
    assert (P[:4 + y0] == -7.0).all() and (P[4 + y1:] == -7.0).all()


def oracle_filtered_pyramid(o, img, w, h, levels, contrast, gauss):
    """tk_build with the filters (track.c:600-629) from the oracle's primitives: u8 -> Q6 into the input plane; the
    contrast stretch from its min and max (gain and offset in long double, as track.c:609-614); the pre-blur into
    level 0; each level reduced from the one below; every plane padded after it is written."""
    inp = _l0(o, img, w, h)
    cur = inp
    if contrast:
        mm = np.zeros(2, np.float32)
        o.lib.vo_minmax(C.byref(inp.c), mm.ctypes.data, 0, h)
        span = np.longdouble(mm[1]) - np.longdouble(mm[0])
        gain = np.longdouble(16320.0) / span if span > 0 else np.longdouble(1.0)
        go = np.array([np.float32(gain), np.float32(-np.longdouble(mm[0]) * gain)], np.float32)
        nrm = HostPlane(w, h)
        o.lib.vo_norm(C.byref(nrm.c), C.byref(inp.c), go.ctypes.data, 0, h)
        _pad(o, nrm)
        cur = nrm
    if gauss:
        g = HostPlane(w, h)
        o.lib.vo_gauss5(C.byref(g.c), C.byref(cur.c), 0, h)
        _pad(o, g)
        cur = g
    planes = [cur]
    for _ in range(1, levels):
        lw, lh = (planes[-1].w + 1) // 2, (planes[-1].h + 1) // 2
        r = HostPlane(lw, lh)
        o.lib.vo_reduce(C.byref(r.c), C.byref(planes[-1].c), 0, lh)
        _pad(o, r)
        planes.append(r)
    from oracle_ffi import Plane
    return planes, (Plane * levels)(*[p.c for p in planes])


CASES = [(vtgpu.STORE_Q6W, vtgpu.BUILD_CONTRAST), (vtgpu.STORE_Q6W, vtgpu.BUILD_PRE_GAUSS),
         (vtgpu.STORE_Q6W, vtgpu.BUILD_CONTRAST | vtgpu.BUILD_PRE_GAUSS), (vtgpu.STORE_Q6W, 0),
         (vtgpu.STORE_F32, vtgpu.BUILD_CONTRAST | vtgpu.BUILD_PRE_GAUSS), (vtgpu.STORE_F32, vtgpu.BUILD_PRE_GAUSS)]


@pytest.mark.parametrize("storage,flags", CASES)
@pytest.mark.parametrize("w,h,seed", [(640, 480, 5), (1001, 333, 6)])
def test_filtered_pyramid_exact(ctx, oracle, storage, flags, w, h, seed):
    """Every level of a pyramid built with the filters equals the oracle's tk_build restatement, bit for bit; a dim
    frame (bytes 40 .. 90) makes the stretch do real work."""
    f0, _, _ = vtgpu.frame_pair(w, h, seed)
    f0 = (40 + (f0.astype(np.uint16) * 50) // 255).astype(np.uint8)
    L = 4
    say("filtered pyramid", storage, flags, w, h)
    p = ctx.pyramid(w, h, L, storage)
    p.build(f0, flags=flags)
    planes, _ = oracle_filtered_pyramid(oracle, f0, w, h, L, bool(flags & 1), bool(flags & 2))
    for lv in range(L):
        assert np.array_equal(p.read_level(lv).view(), planes[lv].view()), f"level {lv}"
    # the batched build gives the same levels
    q = ctx.pyramid(w, h, L, storage)
    d = ctx.to_device(f0)
    ctx.build_many([q], [d], flags=flags)
    for lv in range(L):
        assert np.array_equal(q.read_level(lv).view(), planes[lv].view()), f"batched level {lv}"
    p.destroy()
    q.destroy()
    d.free()


def test_filters_need_a_wide_level0(ctx):
    """A u8-level-0 pyramid refuses the filters with VTG_EINVAL (level 0 is the frame itself there)."""
    f0, _, _ = vtgpu.frame_pair(64, 64, 1)
    p = ctx.pyramid(64, 64, 2, vtgpu.STORE_Q6)
    rc = vtgpu.lib().vtg_pyr_build_ex(ctx.h, p.h, f0.ctypes.data, 64, 0, vtgpu.BUILD_CONTRAST)
    assert rc == -22
    p.destroy()


def test_lk_on_filtered_pyramids(ctx, oracle):
    """LK of the 640 x 480 grid on stretched + blurred pyramids (u16 level 0): the oracle's reference tier on the same
    levels gives the same status and flow within 0.01 px on every well-conditioned point."""
    w, h, L = 640, 480, 3
    f0, f1, _ = vtgpu.frame_pair(w, h, 11)
    flags = vtgpu.BUILD_CONTRAST | vtgpu.BUILD_PRE_GAUSS
    p0, p1 = ctx.pyramid(w, h, L, vtgpu.STORE_Q6W), ctx.pyramid(w, h, L, vtgpu.STORE_Q6W)
    p0.build(f0, flags=flags)
    p1.build(f1, flags=flags)
    k0, I = oracle_filtered_pyramid(oracle, f0, w, h, L, True, True)
    k1, J = oracle_filtered_pyramid(oracle, f1, w, h, L, True, True)
    xs, ys = vtgpu.grid(40, 30, w, h)
    cfg = lk_cfg(15, 15, 10, 0.01, 1e-4, 40.0)
    say("lk filtered")
    got = ctx.lk_track_host(p0, p1, L - 1, vtgpu.LkCfg(15, 15, 10, 0.01, 1e-4, 40.0, 0), xs, ys)
    ref = lane_ref(oracle, I, J, L - 1, cfg, xs, ys)
    well, _ = conditioning(oracle, I, J, L - 1, cfg, xs, ys, threads=8)
    assert (got["status"] == ref["status"])[well].all()
    ok = well & (ref["status"] == 0)
    d = np.maximum(np.abs(got["x"] - ref["x"]), np.abs(got["y"] - ref["y"]))[ok]
    print(f"[lk filtered] {ok.sum()} well-conditioned converged of {len(xs)}, max |d| {d.max():.2e} px")
    assert ok.sum() > len(xs) // 2 and d.max() <= 0.01
    p0.destroy()
    p1.destroy()
