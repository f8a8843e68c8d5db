"""The tracker-level layer (include/vtg_track.h) on the GPU: a host program compiled against include/ (tests/c/tk_scene,
the shape of the reference's replay host) holds six analytic scenes to the reference's EPE limits (tests/scenes.md:11-39,
score.c:283-401: translate, camera, shapes and occlusion at a median of 0.02 px, brightness and noise at 0.1 px), and
the layer's job ops equal the path's own calls and the oracle: tk_build's levels bit for bit, OP_LK's records (the
head-rate guess, forward-backward 0.5 px, lost features dropped) bit for bit against the oracle's lane64 tier, and the
armed search's pick against the oracle's search."""
import ctypes as C
import json
import os
import subprocess

import numpy as np
import pytest

import vtgpu
from gpu_util import lane_ref, say
from oracle_ffi import lk_cfg

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TK_SCENE = os.path.join(REPO, "tests", "c", "tk_scene")
LIMITS = {"translate": 0.02, "camera": 0.02, "shapes": 0.02, "occlusion": 0.02, "brightness": 0.1, "noise": 0.1}


@pytest.mark.parametrize("scene", list(LIMITS))
def test_scene_epe_within_the_reference_limit(scene):
    assert os.path.exists(TK_SCENE), "tests/c/tk_scene is built by __graft_entry__.build()"
    say("tk_scene", scene)
    r = subprocess.run([TK_SCENE, scene, "640", "480", "150"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    print(f"[tk] {res}")
    assert res["epe"]["n"] >= 2000
    assert res["epe"]["median"] <= LIMITS[scene], res


class TkConfig(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("levels", C.c_uint32), ("patch_w", C.c_uint32),
                ("patch_h", C.c_uint32), ("max_iter", C.c_uint32), ("eps", C.c_float), ("min_eig", C.c_float),
                ("max_residual", C.c_float), ("fb_max", C.c_float), ("max_features", C.c_uint32), ("sets", C.c_uint32),
                ("storage", C.c_int), ("device", C.c_int)]


class FeatOut(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("u", C.c_float), ("v", C.c_float), ("xr", C.c_float),
                ("yr", C.c_float), ("disparity", C.c_float), ("sigma", C.c_float), ("flags", C.c_uint32),
                ("z", C.c_float)]


class TkFound(C.Structure):
    _fields_ = [("armed", C.c_int), ("hx", C.c_int), ("hy", C.c_int), ("ncc", C.c_double), ("second", C.c_double),
                ("x", C.c_float), ("y", C.c_float)]


class TkResult(C.Structure):
    _fields_ = [("instance", C.c_uint64), ("nfeat", C.c_uint32), ("tracked", C.c_uint32), ("dropped", C.c_uint32),
                ("guess_u", C.c_float), ("guess_v", C.c_float), ("rate_u", C.c_double), ("rate_v", C.c_double),
                ("search", TkFound), ("feat", FeatOut * 512)]


ALLOC = C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_size_t)
FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int)
RUN = C.CFUNCTYPE(None, C.c_void_p, FN, C.c_void_p)


class Tracker:
    """ctypes wrapper: host memory from Python buffers (64 B aligned, kept alive), executors that run once."""

    def __init__(self, cfg):
        L = self.L = vtgpu.lib()
        L.vtg_tk_create.restype = C.c_void_p
        L.vtg_tk_create.argtypes = [C.POINTER(TkConfig), C.c_int, C.c_int, ALLOC, C.c_void_p, RUN, C.c_void_p, RUN,
                                    C.c_void_p]
        L.vtg_tk_error.restype = C.c_char_p
        L.vtg_tk_error.argtypes = [C.c_void_p]
        L.vtg_tk_level.restype = C.POINTER(vtgpu.Plane)
        for n, a in (("vtg_tk_destroy", [C.c_void_p]), ("vtg_tk_sets", [C.c_void_p]),
                     ("vtg_tk_build", [C.c_void_p, C.c_int, C.c_void_p, C.POINTER(vtgpu.Msg)]),
                     ("vtg_tk_track", [C.c_void_p, C.c_int, C.POINTER(vtgpu.Msg), C.POINTER(TkResult)]),
                     ("vtg_tk_set_features", [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
                     ("vtg_tk_add_features", [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
                     ("vtg_tk_set_rate", [C.c_void_p, C.c_double, C.c_double]),
                     ("vtg_tk_arm_search", [C.c_void_p] + [C.c_int] * 9),
                     ("vtg_tk_level", [C.c_void_p, C.c_int, C.c_int])):
            getattr(L, n).argtypes = a
        self.keep = []
        self.runs = 0

        def alloc(ctx, n):
            b = C.create_string_buffer(n + 64)
            self.keep.append(b)
            return (C.addressof(b) + 63) & ~63

        def run(ctx, fn, arg):
            self.runs += 1
            fn(arg, 0, 1)

        self.cbs = (ALLOC(alloc), RUN(run))
        self.cfg = cfg
        self.h = L.vtg_tk_create(C.byref(cfg), 1, 1, self.cbs[0], None, self.cbs[1], None, self.cbs[1], None)

    def check(self, rc):
        if rc:
            raise vtgpu.VtgError(f"{rc}: {self.L.vtg_tk_error(self.h).decode()}")

    def frame(self, k, img, flow=True):
        m = vtgpu.Msg()
        m.magic, m.type, m.dest, m.instance, m.acq_ns = 0x4d54, 1, 1 | (2 if flow else 0), k + 1, int(k * 1e9 / 30)
        s = k % self.L.vtg_tk_sets(self.h)
        img = np.ascontiguousarray(img)
        self.check(self.L.vtg_tk_build(self.h, s, img.ctypes.data, C.byref(m)))
        r = TkResult()
        self.check(self.L.vtg_tk_track(self.h, s, C.byref(m), C.byref(r)))
        return s, r

    def set_features(self, xs, ys):
        xs, ys = np.ascontiguousarray(xs, np.float32), np.ascontiguousarray(ys, np.float32)
        return self.L.vtg_tk_set_features(self.h, len(xs), xs.ctypes.data, ys.ctypes.data)

    def destroy(self):
        if self.h:
            self.L.vtg_tk_destroy(self.h)
            self.h = None


def _cfg(w, h, levels=4, sets=2, storage=vtgpu.STORE_Q6, fb=0.5):
    return TkConfig(w, h, levels, 21, 21, 10, 0.01, 0.1, 25.0, fb, 128, sets, storage, 0)


@pytest.mark.parametrize("storage", [vtgpu.STORE_Q6, vtgpu.STORE_F32_U8])
def test_job_ops_equal_the_oracle(oracle, storage):
    w, h, L = 640, 480, 4
    f0, f1, _ = vtgpu.frame_pair(w, h, 21)
    f2, _, _ = vtgpu.frame_pair(w, h, 22)
    t = Tracker(_cfg(w, h, L, storage=storage))
    assert t.h and t.L.vtg_tk_sets(t.h) == 2
    s0, r0 = t.frame(0, f0)
    assert r0.nfeat == 0 and r0.tracked == 0
    # tk_build: every level equals the oracle's pyramid (the reference's padded plane, visible part)
    k0, I = oracle.pyramid(f0, w, h, L, threads=8)
    for lv in range(L):
        p = t.L.vtg_tk_level(t.h, s0, lv)
        assert p, t.L.vtg_tk_error(t.h)
        pl = p.contents
        got = np.ctypeslib.as_array(C.cast(pl.p, C.POINTER(C.c_float)), shape=(pl.h * pl.stride,))
        got = np.lib.stride_tricks.as_strided(got, (pl.h, pl.w), (4 * pl.stride, 4)).copy()
        assert np.array_equal(got, k0[lv].view()), lv
    # OP_LK: 120 features, the head-rate guess (set: 1.5 px over a frame), forward-backward 0.5 px
    xs, ys = vtgpu.grid(12, 10, w, h)
    t.check(t.set_features(xs, ys))
    t.check(t.L.vtg_tk_set_rate(t.h, 45.0, -30.0))
    s1, r1 = t.frame(1, f1)
    assert abs(r1.guess_u - 1.5) < 1e-5 and abs(r1.guess_v + 1.0) < 1e-5
    k1, J = oracle.pyramid(f1, w, h, L, threads=8)
    gx = np.full(len(xs), np.float32(r1.guess_u), np.float32)
    gy = np.full(len(xs), np.float32(r1.guess_v), np.float32)
    ref = lane_ref(oracle, I, J, L - 1, lk_cfg(21, 21, 10, 0.01, 0.1, 25.0), xs, ys, gx=gx, gy=gy, fb_max=0.5,
                   threads=8)
    ok = ref["status"] == 0
    assert r1.tracked == ok.sum() and r1.dropped == (~ok).sum() and r1.nfeat == ok.sum()
    feats = r1.feat[:r1.nfeat]
    assert np.array_equal(np.array([f.x for f in feats], np.float32), ref["x"][ok])
    assert np.array_equal(np.array([f.y for f in feats], np.float32), ref["y"][ok])
    assert np.array_equal(np.array([f.u for f in feats], np.float32),
                          (ref["x"][ok].astype(np.float64) - xs[ok].astype(np.float64)).astype(np.float32))
    # the head's rate after the frame: the median displacement (the upper middle of an even count) over dt
    du = np.sort((ref["x"][ok].astype(np.float64) - xs[ok].astype(np.float64)).astype(np.float32))
    dt = (int(1 * 1e9 / 30) - 0) * 1e-9
    assert abs(r1.rate_u - float(du[len(du) // 2]) / dt) < 1e-9 * max(1.0, abs(r1.rate_u))
    # the armed search: a 32 x 21 appearance cut from frame 1's level 0 at (300, 200), searched around it in the next
    # frame, first the same picture again (found where it was cut, correlation 1), then an unrelated one (the oracle's
    # pick)
    t.check(t.L.vtg_tk_arm_search(t.h, s1, 0, 300, 200, 32, 21, 300, 200, 40))
    app, norm = oracle.ncc_template(k1[0], 300, 200, 32, 21)
    for k, img in ((2, f1), (3, f2)):
        s, r = t.frame(k, img)
        assert r.search.armed == 1
        kk, _ = oracle.pyramid(img, w, h, 1, threads=8)
        of, _ = oracle.ncc_search(kk[0], app, 32, 21, norm, 260, 340, 160, 240, 1)
        say("tk search", k, r.search.hx, r.search.hy, r.search.ncc, of.hx, of.hy, float(of.ncc))
        assert (r.search.hx, r.search.hy) == (of.hx, of.hy) and abs(r.search.ncc - float(of.ncc)) <= 1e-5
        if k == 2:
            assert (r.search.hx, r.search.hy) == (600, 400) and r.search.ncc > 0.9999
            assert abs(r.search.x - (300 + 15.5)) < 1e-4 and abs(r.search.y - (200 + 10)) < 1e-4
    assert t.runs >= 4  # the build and the per-feature work went through the caller's executors
    t.destroy()


def test_errors():
    w, h = 320, 240
    bad = Tracker(_cfg(w, h, sets=5))
    assert not bad.h
    t = Tracker(_cfg(w, h))
    f = np.zeros((h, w), np.uint8)
    m = vtgpu.Msg()
    assert t.L.vtg_tk_build(t.h, 2, f.ctypes.data, C.byref(m)) == -22
    r = TkResult()
    assert t.L.vtg_tk_track(t.h, 0, C.byref(m), C.byref(r)) == -22  # set 0 not built
    xs = np.zeros(129, np.float32)
    assert t.set_features(xs, xs) == -34
    assert t.set_features(xs[:10], xs[:10]) == 0
    assert t.L.vtg_tk_level(t.h, 1, 0) is None or not t.L.vtg_tk_level(t.h, 1, 0)
    t.destroy()
