"""GPU pyramids against the oracle: bit-exact for the exact storages (u8 / u16 Q6 and fp32), over the whole padded
plane, plus the reference's m1 hash lines; fp16 / bf16 within their rounding."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import vtgpu
from gpu_util import say
from oracle_ffi import HostPlane

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("size", ["640x480", "1280x720", "1920x1080"])
@pytest.mark.parametrize("storage", [vtgpu.STORE_Q6, vtgpu.STORE_F32, vtgpu.STORE_Q6F, vtgpu.STORE_F32_U8])
def test_m1_reduce_lines(ctx, oracle, size, storage):
    gold = json.load(open(os.path.join(GOLD, "m1_hashes.json")))[size]
    w, h = map(int, size.split("x"))
    img = oracle.m1_image(w, h)[:, :w]
    say("pyramid m1", size, "storage", storage)
    p = ctx.pyramid(w, h, 2, storage)
    p.build(img)
    l0, l1 = p.read_level(0), p.read_level(1)
    assert "%016x" % oracle.lib.vo_hash_plane(C.byref(l0.c)) == gold["u8_to_q6"]
    assert "%016x" % oracle.lib.vo_hash_plane(C.byref(l1.c)) == gold["reduce"]
    p.destroy()


@pytest.mark.parametrize("w,h,levels", [(640, 480, 3), (1920, 1080, 4), (641, 479, 5), (97, 61, 4), (3840, 2160, 5),
                                        (8, 8, 3)])
@pytest.mark.parametrize("storage", [vtgpu.STORE_Q6, vtgpu.STORE_F32, vtgpu.STORE_Q6F, vtgpu.STORE_F32_U8])
def test_pyramid_bit_exact(ctx, oracle, w, h, levels, storage):
    """Every level, padding included, equals the oracle's padded fp32 Q6 planes bit for bit."""
    f0, _, _ = vtgpu.frame_pair(w, h, w * 7 + h) if w >= 64 else (np.random.default_rng(1).integers(
        0, 256, (h, w), dtype=np.uint8), None, None)
    planes, _ = oracle.pyramid(f0, w, h, levels, threads=8)
    say("pyramid exact", w, h, levels, "storage", storage)
    p = ctx.pyramid(w, h, levels, storage)
    p.build(f0)
    for L in range(levels):
        got = p.read_level(L)
        assert got.padded().tobytes() == planes[L].buf[: got.padded().size].tobytes(), (w, h, L)
    p.destroy()


def test_pyramid_from_device_frame(ctx, oracle):
    """A frame already in HBM (pitch wider than the row) builds the same pyramid, and level 0 holds the frame."""
    w, h = 1280, 720
    f0, _, _ = vtgpu.frame_pair(w, h, 5)
    padded = np.zeros((h, 1408), np.uint8)
    padded[:, :w] = f0
    say("pyramid from device frame")
    d = ctx.to_device(padded)
    p = ctx.pyramid(w, h, 4)
    p.build(d, on_device=True, pitch=1408)
    planes, _ = oracle.pyramid(f0, w, h, 4)
    for L in range(4):
        assert np.array_equal(p.read_level(L).view(), planes[L].view())
    p.destroy()
    d.free()


@pytest.mark.parametrize("storage", [vtgpu.STORE_Q6, vtgpu.STORE_F32, vtgpu.STORE_F16, vtgpu.STORE_Q6F,
                                     vtgpu.STORE_F32_U8])
def test_build_many_uneven(ctx, oracle, storage):
    """Pyramids of uneven sizes and depths built together (one launch per level) equal the ones built alone."""
    sizes = [(640, 480, 3), (1280, 720, 5), (97, 61, 2), (1920, 1080, 4), (333, 250, 1)]
    frames = [vtgpu.frame_pair(w, h, 40 + i)[0] if w >= 64 else None for i, (w, h, _) in enumerate(sizes)]
    say("build_many", storage)
    pyrs = [ctx.pyramid(w, h, L, storage) for w, h, L in sizes]
    dev = [ctx.to_device(f) for f in frames]
    ctx.build_many(pyrs, dev)
    for (w, h, L), f, p in zip(sizes, frames, pyrs):
        q = ctx.pyramid(w, h, L, storage)
        q.build(f)
        for lv in range(L):
            assert p.read_level(lv).padded().tobytes() == q.read_level(lv).padded().tobytes(), (w, h, lv)
        if storage != vtgpu.STORE_F16:
            planes, _ = oracle.pyramid(f, w, h, L)
            assert np.array_equal(p.read_level(L - 1).view(), planes[L - 1].view())
        q.destroy()
        p.destroy()
    for d in dev:
        d.free()


@pytest.mark.parametrize("storage,tol", [(vtgpu.STORE_F16, 2.0 ** -11), (vtgpu.STORE_BF16, 2.0 ** -8)])
def test_pyramid_half_storage(ctx, oracle, storage, tol):
    """fp16 / bf16 levels hold grey = Q6 / 64 rounded once per level: within a few ulps of the exact pyramid."""
    w, h = 1920, 1080
    f0, _, _ = vtgpu.frame_pair(w, h, 3)
    planes, _ = oracle.pyramid(f0, w, h, 4, threads=8)
    say("pyramid half storage", storage)
    p = ctx.pyramid(w, h, 4, storage)
    p.build(f0)
    assert np.array_equal(p.read_level(0).view(), planes[0].view())
    for L in range(1, 4):
        got, ref = p.read_level(L).view(), planes[L].view()
        err = np.abs(got - ref) / 64.0
        # rounding of the store at each level plus the exact path's own (h + 128) >> 8 of up to 0.5 Q6
        assert err.max() <= 255 * tol * L + 0.5 / 64 * L + 1e-6, (L, err.max())


def test_reference_shaped_plane_calls(ctx, oracle):
    """vtg_img_u8_to_q6 / vtg_img_reduce over host planes, row ranges and padding as the reference fills them."""
    w, h = 333, 250
    img = oracle.m1_image(w, h)
    ref = HostPlane(w, h)
    oracle.lib.vo_u8_to_q6(C.byref(ref.c), img.ctypes.data, img.strides[0], 0, h)
    oracle.lib.vo_pad_rows(C.byref(ref.c), 0, h)
    got = vtgpu.HostPlane(w, h)
    L = vtgpu.lib()
    say("img_u8_to_q6 rows 10..200")
    ctx.check(L.vtg_img_u8_to_q6(ctx.h, C.byref(got.c), img.ctypes.data, img.strides[0], 10, 200))
    assert np.array_equal(got.padded()[4 + 10:4 + 200], ref.padded()[4 + 10:4 + 200])
    assert not got.padded()[4 + 200:].any()
    oracle.lib.vo_pad_tb(C.byref(ref.c))
    r_ref = HostPlane((w + 1) // 2, (h + 1) // 2)
    oracle.lib.vo_reduce(C.byref(r_ref.c), C.byref(ref.c), 0, r_ref.h)
    oracle.lib.vo_pad_rows(C.byref(r_ref.c), 0, r_ref.h)
    r_got = vtgpu.HostPlane((w + 1) // 2, (h + 1) // 2)
    src = vtgpu.HostPlane(w, h)
    src.buf[:] = ref.buf[: src.buf.size]
    say("img_reduce rows 3..120")
    ctx.check(L.vtg_img_reduce(ctx.h, C.byref(r_got.c), C.byref(src.c), 3, 120))
    assert np.array_equal(r_got.padded()[4 + 3:4 + 120], r_ref.padded()[4 + 3:4 + 120])


def test_bad_arguments(ctx):
    L = vtgpu.lib()
    p = C.c_void_p()
    assert L.vtg_pyr_create(ctx.h, 4, 4, 3, 0, C.byref(p)) == -34
    assert L.vtg_pyr_create(ctx.h, 64, 64, 7, 0, C.byref(p)) == -22
    assert L.vtg_pyr_create(ctx.h, 64, 64, 3, 9, C.byref(p)) == -22
    assert b"levels" in L.vtg_ctx_error(ctx.h) or b"storage" in L.vtg_ctx_error(ctx.h)
