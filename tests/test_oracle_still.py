"""The oracle's still-cell vectors (vo_still_*, a restatement of src/kern/still.c) pinned: against the reference's own
still.c compiled from source into oracle/_ref (every frame's still_hash over cells and mean, and the motion summary,
bit for bit), and by closed forms that a dropped term or a wrong index would break."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle_ffi import PAD_X, PAD_Y, REF_SO, HostPlane, Motion, OracleStill, Plane, StillCfg

FNV0 = 0xcbf29ce484222325


def still_frames(o, w, h, n, seed, square=True, jump=5):
    """n u8 frames: a fixed textured scene plus fresh noise each frame; from frame 4 a bright square walks 3 px a
    frame (so some cells move); frame 6 is `jump` grey levels brighter overall (12: past the usual noise limit, so
    every cell counts as moved)."""
    rng = np.random.default_rng(seed)
    base = o.m2_render(w, h, 300, 200)[:, :w].astype(np.int32)
    out = []
    for f in range(n):
        img = base + rng.normal(0.0, 2.0, base.shape).round().astype(np.int32)
        if square and f >= 4:
            x0 = 20 + 3 * (f - 4)
            img[30:50, x0:x0 + 20] = 230
        if f == 6:
            img += jump
        out.append(np.clip(img, 0, 255).astype(np.uint8))
    return out


def q6_plane(o, img):
    h, w = img.shape
    hp = HostPlane(w, h)
    img = np.ascontiguousarray(img)
    o.lib.vo_u8_to_q6(C.byref(hp.c), img.ctypes.data, img.strides[0], 0, h)
    o.lib.vo_pad_rows(C.byref(hp.c), 0, h)
    o.lib.vo_pad_tb(C.byref(hp.c))
    return hp


def run_oracle(o, cfg, planes):
    s = OracleStill(o, cfg)
    hashes, motions = [], []
    for cur in planes:
        if s.s.frames:
            s.vectors(cur)
        s.decide()
        s.update(cur)
        s.finish()
        hashes.append(s.hash(FNV0))
        motions.append(s.motion())
    return s, hashes, motions


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref not built (no reference tree here)")
@pytest.mark.parametrize("w,h,avg,mpx,mgrey,jump", [(160, 120, 16, 0.5, 4.0, 5), (150, 101, 4, 0.25, 2.0, 12)])
def test_still_equals_reference(oracle, w, h, avg, mpx, mgrey, jump):
    """The reference's still_create/vectors/decide/update/finish over 9 frames and the oracle's restatement give the same
    still_hash (cells + mean) and the same motion summary after every frame (ragged cells at the edges included)."""
    ref = C.CDLL(REF_SO)
    ALLOC = C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_size_t)
    keep = []

    def alloc(_ctx, n):
        a = np.zeros(n + 128, np.uint8)
        keep.append(a)
        return (a.ctypes.data + 63) & ~63

    alloc_cb = ALLOC(alloc)
    ref.still_create.restype = C.c_void_p
    ref.still_create.argtypes = [C.POINTER(StillCfg), ALLOC, C.c_void_p]
    for n in ("still_vectors", "still_update"):
        getattr(ref, n).argtypes = [C.c_void_p, C.POINTER(Plane), C.c_int, C.c_int]
    for n in ("still_decide", "still_finish"):
        getattr(ref, n).argtypes = [C.c_void_p]
    ref.still_first.argtypes = [C.c_void_p]
    ref.still_hash.restype = C.c_uint64
    ref.still_hash.argtypes = [C.c_void_p, C.c_uint64]
    motion_buf = np.zeros(1024, np.uint8)
    ref.still_motion.argtypes = [C.c_void_p, C.c_void_p]
    cfg = StillCfg(w, h, avg, mpx, mgrey, 1)
    rs = ref.still_create(C.byref(cfg), alloc_cb, None)
    assert rs
    planes = [q6_plane(oracle, f) for f in still_frames(oracle, w, h, 9, 5, jump=jump)]
    s, hashes, motions = run_oracle(oracle, cfg, planes)
    if jump == 12:
        assert motions[6]["moving"] == s.s.ncells   # everything moved
    moved = 0
    for k, cur in enumerate(planes):
        if not ref.still_first(rs):
            ref.still_vectors(rs, C.byref(cur.c), 0, 1)
        ref.still_decide(rs)
        ref.still_update(rs, C.byref(cur.c), 0, 1)
        ref.still_finish(rs)
        assert ref.still_hash(rs, FNV0) == hashes[k], f"frame {k}"
        ref.still_motion(rs, motion_buf.ctypes.data)
        m = Motion.from_buffer_copy(motion_buf[:C.sizeof(Motion)].tobytes())
        for key, val in motions[k].items():
            assert getattr(m, key) == val, (k, key)
        moved += motions[k]["moving"]
    assert moved > 0 and motions[-1]["tracked"] > 0
    s.free()


def test_still_closed_forms(oracle):
    """Identical frames: every vector and difference is exactly 0 and each cell's count grows to avg; a frame one grey
    level brighter everywhere: every cell's difference is exactly 1.0 (and the median 1.0); the mean after a weight-1
    frame is that frame exactly."""
    w, h = 96, 64
    img = still_frames(oracle, w, h, 1, 9, square=False)[0]
    a, b = q6_plane(oracle, img), q6_plane(oracle, np.clip(img.astype(np.int32) + 1, 0, 255).astype(np.uint8))
    assert (img < 255).all()
    cfg = StillCfg(w, h, 3, 0.5, 4.0, 1)
    s = OracleStill(oracle, cfg)
    for k in range(4):
        if s.s.frames:
            s.vectors(a)
            assert (s.cells["u"] == 0).all() and (s.cells["v"] == 0).all() and (s.cells["diff"] == 0).all()
        s.decide()
        s.update(a)
        s.finish()
        assert np.array_equal(s.mean(), a.view())
    assert (s.count == 3).all() and s.s.moving == 0
    s.vectors(b)
    assert (s.cells["diff"] == 1.0).all()
    s.decide()
    assert s.s.diff_median == 1.0 and s.s.diff_limit == 4.0
    s.free()
