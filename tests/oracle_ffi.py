"""ctypes binding of oracle/liboracle.so (the CPU oracle) and, when built, oracle/_ref/libvtref.so (the
reference's own path compiled from its sources). Test infrastructure only."""
import ctypes as C
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libvtref.so")

PAD_X, PAD_Y, Q6 = 32, 4, 64


try:
    # one ctypes class per C struct, shared with the GPU binding, so planes and configs pass to either library
    import sys as _sys
    _sys.path.insert(0, os.path.join(REPO, "intel-phi-visiontrack_amd"))
    from vtgpu import HostPlane as _VtgHostPlane, LkCfg, Plane  # noqa: E402
except ImportError:  # pragma: no cover
    _VtgHostPlane = None

    class Plane(C.Structure):
        _fields_ = [("p", C.POINTER(C.c_float)), ("w", C.c_int), ("h", C.c_int), ("stride", C.c_int)]

    class LkCfg(C.Structure):
        _fields_ = [("w", C.c_int), ("h", C.c_int), ("max_iter", C.c_int), ("eps", C.c_float),
                    ("min_eig", C.c_float), ("max_residual", C.c_float), ("prefetch", C.c_int)]


LK_OUT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("min_eig", "<f4"), ("residual", "<f4"),
                         ("cov_xx", "<f4"), ("cov_xy", "<f4"), ("cov_yy", "<f4"),
                         ("iters", "<u2"), ("status", "<u2")])
assert LK_OUT_DTYPE.itemsize == 32

CELL_DTYPE = np.dtype([("v", "<f4"), ("x", "<i4"), ("y", "<i4")])


class Found(C.Structure):
    _fields_ = [("hx", C.c_int), ("hy", C.c_int), ("ncc", C.c_longdouble), ("second", C.c_longdouble),
                ("cells", C.c_int)]


def plane_stride(w):
    return ((w + 15) & ~15) + 2 * PAD_X


HostPlane = _VtgHostPlane


def _ptr(a, t=C.c_void_p):
    return C.cast(a.ctypes.data, t) if a is not None else None


class Oracle:
    def __init__(self, path=ORACLE_SO):
        L = self.lib = C.CDLL(path)
        P, I, F, U8P = C.POINTER(Plane), C.c_int, C.c_float, C.c_void_p
        L.vo_hash_plane.restype = C.c_uint64
        L.vo_hash_plane.argtypes = [P]
        L.vo_hash_lk.restype = C.c_uint64
        L.vo_hash_lk.argtypes = [C.c_void_p, I]
        L.vo_m1_image.argtypes = [U8P, I, I, I]
        L.vo_m2_render.argtypes = [U8P, I, I, I, I, I]
        for n in ("vo_pad_rows",):
            getattr(L, n).argtypes = [P, I, I]
        L.vo_pad_tb.argtypes = [P]
        L.vo_u8_to_q6.argtypes = [P, U8P, I, I, I]
        L.vo_reduce.argtypes = [P, P, I, I]
        L.vo_gauss5.argtypes = [P, P, I, I]
        L.vo_sobel.argtypes = [P, P, P, I, I]
        L.vo_absdiff.argtypes = [P, P, P, I, I]
        L.vo_norm.argtypes = [P, P, C.c_void_p, I, I]
        L.vo_minmax.argtypes = [P, C.c_void_p, I, I]
        L.vo_pyr_build_mt.argtypes = [P, U8P, I, I, I]
        L.vo_lk_track.argtypes = [I, P, P, I, C.POINTER(LkCfg), F, F, F, F, C.c_void_p]
        L.vo_lk_flowback.argtypes = [I, P, P, I, C.POINTER(LkCfg), F, F, F, C.c_void_p]
        L.vo_lk_batch.argtypes = [I, P, P, I, C.POINTER(LkCfg), I, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, F, C.c_void_p, I]
        L.vo_ncc_template.restype = I
        L.vo_ncc_template.argtypes = [P, I, I, I, I, C.c_void_p, C.POINTER(C.c_float)]
        L.vo_ncc_at.restype = C.c_longdouble
        L.vo_ncc_at.argtypes = [P, C.c_void_p, I, I, C.c_longdouble, I, I]
        L.vo_ncc_half.restype = C.c_longdouble
        L.vo_ncc_half.argtypes = [P, C.c_void_p, I, I, C.c_longdouble, I, I]
        L.vo_peak_offset.restype = C.c_longdouble
        L.vo_peak_offset.argtypes = [C.c_longdouble] * 3
        L.vo_ncc_cells.argtypes = [P, C.c_void_p, I, I, F, I, I, I, I, I, C.c_void_p, I]
        L.vo_ncc_search.argtypes = [P, C.c_void_p, I, I, F, I, I, I, I, I, C.POINTER(Found), C.c_void_p, I]

    # --- inputs
    def m1_image(self, w, h):
        stride = (w + 15) & ~15
        img = np.zeros((h, stride), np.uint8)
        self.lib.vo_m1_image(img.ctypes.data, w, h, stride)
        return img

    def m2_render(self, w, h, ox, oy):
        stride = (w + 15) & ~15
        img = np.zeros((h, stride), np.uint8)
        self.lib.vo_m2_render(img.ctypes.data, w, h, stride, ox, oy)
        return img

    # --- pyramid
    def pyramid(self, img, w, h, levels, threads=1):
        img = np.ascontiguousarray(img)
        planes, ww, hh = [], w, h
        for _ in range(levels):
            planes.append(HostPlane(ww, hh))
            ww, hh = (ww + 1) // 2, (hh + 1) // 2
        arr = (Plane * levels)(*[p.c for p in planes])
        self.lib.vo_pyr_build_mt(arr, img.ctypes.data, img.strides[0], levels, threads)
        return planes, arr

    def hash_plane(self, hp):
        return self.lib.vo_hash_plane(C.byref(hp.c))

    # --- LK
    def lk_batch(self, tier, I_arr, J_arr, top, cfg, xs, ys, gx=None, gy=None, fb_max=0.0, threads=8):
        xs = np.ascontiguousarray(xs, np.float32)
        ys = np.ascontiguousarray(ys, np.float32)
        out = np.zeros(len(xs), LK_OUT_DTYPE)
        gx = None if gx is None else np.ascontiguousarray(gx, np.float32)
        gy = None if gy is None else np.ascontiguousarray(gy, np.float32)
        self.lib.vo_lk_batch(tier, I_arr, J_arr, top, C.byref(cfg), len(xs), xs.ctypes.data, ys.ctypes.data,
                             None if gx is None else gx.ctypes.data, None if gy is None else gy.ctypes.data,
                             fb_max, out.ctypes.data, threads)
        return out

    # --- NCC
    def ncc_template(self, hp, x0, y0, tw, th):
        app = np.zeros(tw * th, np.float32)
        norm = C.c_float()
        rc = self.lib.vo_ncc_template(C.byref(hp.c), x0, y0, tw, th, app.ctypes.data, C.byref(norm))
        if rc:
            raise ValueError("template does not fit")
        return app, norm.value

    def ncc_at(self, hp, app, tw, th, norm, x, y):
        return self.lib.vo_ncc_at(C.byref(hp.c), app.ctypes.data, tw, th, norm, x, y)

    def ncc_search(self, hp, app, tw, th, norm, x0, x1, y0, y1, step, threads=8):
        hi_x, hi_y = hp.w - tw - 2, hp.h - th - 2
        cx0, cx1, cy0, cy1 = max(x0, 1), min(x1, hi_x), max(y0, 1), min(y1, hi_y)
        ncell = max(0, ((cx1 - cx0) // 16 + 1) * ((cy1 - cy0) // 16 + 1)) if cx1 >= cx0 and cy1 >= cy0 else 0
        cells = np.zeros(max(ncell, 1), CELL_DTYPE)
        f = Found()
        self.lib.vo_ncc_search(C.byref(hp.c), app.ctypes.data, tw, th, norm, x0, x1, y0, y1, step, C.byref(f),
                               cells.ctypes.data, threads)
        return f, cells[:ncell]


def lk_cfg(w=21, h=21, max_iter=10, eps=0.01, min_eig=1e-4, max_residual=40.0):
    return LkCfg(w, h, max_iter, eps, min_eig, max_residual, 0)


# --- still-cell vectors (struct vtg_still_cfg / vtg_cell / vtg_motion in include/vt_types.h; struct vo_still)
class StillCfg(C.Structure):
    _fields_ = [("w", C.c_int), ("h", C.c_int), ("avg", C.c_int), ("motion_px", C.c_float),
                ("motion_grey", C.c_float), ("threads", C.c_int)]


CELL_STILL_DTYPE = np.dtype([("u", "<f4"), ("v", "<f4"), ("diff", "<f4"), ("state", "<u4"), ("sigma", "<f4"),
                             ("pad", "<f4")])
assert CELL_STILL_DTYPE.itemsize == 24

MOTION_FIELDS = ["cols", "rows", "cell_at", "tracked", "moving", "count_mean", "drift_u", "drift_v", "drift_rms",
                 "diff_median", "diff_limit", "pad", "cycles_vectors", "cycles_mean", "frames"]


class Motion(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in MOTION_FIELDS[:5]] + [(n, C.c_float) for n in MOTION_FIELDS[5:12]] + \
               [(n, C.c_uint64) for n in MOTION_FIELDS[12:]]


class VoStill(C.Structure):
    _fields_ = [("w", C.c_int), ("h", C.c_int), ("cols", C.c_int), ("rows", C.c_int), ("ncells", C.c_int),
                ("avg", C.c_int), ("motion_px", C.c_float), ("motion_grey", C.c_float), ("mean", Plane),
                ("cell", C.c_void_p), ("raw", C.c_void_p), ("count", C.c_void_p), ("weight", C.c_void_p),
                ("frames", C.c_uint64), ("tracked", C.c_uint32), ("moving", C.c_uint32),
                ("count_mean", C.c_float), ("drift_u", C.c_float), ("drift_v", C.c_float),
                ("drift_rms", C.c_float), ("diff_median", C.c_float), ("diff_limit", C.c_float),
                ("noise_usual", C.c_longdouble), ("lk_tier", C.c_int)]


class OracleStill:
    """The oracle's still-cell state (vo_still_*), with numpy views of its arrays."""

    def __init__(self, oracle, cfg):
        L = self.L = oracle.lib
        P = C.POINTER(VoStill)
        L.vo_still_new.restype = P
        L.vo_still_new.argtypes = [C.POINTER(StillCfg)]
        L.vo_still_free.argtypes = [P]
        L.vo_still_vectors.argtypes = [P, C.POINTER(Plane), C.c_int]
        L.vo_still_decide.argtypes = [P]
        L.vo_still_update.argtypes = [P, C.POINTER(Plane)]
        L.vo_still_finish.argtypes = [P]
        L.vo_still_hash.restype = C.c_uint64
        L.vo_still_hash.argtypes = [P, C.c_uint64]
        L.vo_still_centre.argtypes = [P, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        self.p = L.vo_still_new(C.byref(cfg))
        s = self.s = self.p.contents
        n = s.ncells
        self.cells = np.ctypeslib.as_array(C.cast(s.cell, C.POINTER(C.c_uint8)), (n * 24,)).view(CELL_STILL_DTYPE)
        self.count = np.ctypeslib.as_array(C.cast(s.count, C.POINTER(C.c_uint16)), (n,))
        self.weight = np.ctypeslib.as_array(C.cast(s.weight, C.POINTER(C.c_float)), (n,))
        st = s.mean.stride
        base = C.cast(s.mean.p, C.c_void_p).value - (PAD_Y * st + PAD_X) * 4
        self.mean_padded = np.ctypeslib.as_array(C.cast(base, C.POINTER(C.c_float)),
                                                 ((s.h + 2 * PAD_Y) * st,)).reshape(s.h + 2 * PAD_Y, st)

    def mean(self):
        s = self.s
        return self.mean_padded[PAD_Y:PAD_Y + s.h, PAD_X:PAD_X + s.w]

    def vectors(self, cur, threads=8):
        self.L.vo_still_vectors(self.p, C.byref(cur.c), threads)

    def decide(self):
        self.L.vo_still_decide(self.p)

    def update(self, cur):
        self.L.vo_still_update(self.p, C.byref(cur.c))

    def finish(self):
        self.L.vo_still_finish(self.p)

    def hash(self, h=0xcbf29ce484222325):
        return self.L.vo_still_hash(self.p, h)

    def motion(self):
        s = self.s
        return {k: getattr(s, k) for k in ("tracked", "moving", "count_mean", "drift_u", "drift_v", "drift_rms",
                                          "diff_median", "diff_limit", "frames")}

    def centres(self):
        n = self.s.ncells
        xs, ys = np.zeros(n, np.float32), np.zeros(n, np.float32)
        x, y = C.c_float(), C.c_float()
        for i in range(n):
            self.L.vo_still_centre(self.p, i, C.byref(x), C.byref(y))
            xs[i], ys[i] = x.value, y.value
        return xs, ys

    def free(self):
        self.L.vo_still_free(self.p)
        self.p = None


class StereoScan(C.Structure):
    """struct stereo_scan (src/kern/stereo.h:12-16) = struct vo_stereo_scan = struct vtg_stereo_scan."""
    _fields_ = [("disparity", C.c_float), ("ncc", C.c_float), ("second", C.c_float)]

    def tup(self):
        return (self.disparity, self.ncc, self.second)


def bind_stereo(lib, name):
    fn = getattr(lib, name)
    fn.argtypes = [C.POINTER(Plane), C.POINTER(Plane), C.c_float, C.c_float, C.c_int, C.c_int,
                   C.POINTER(StereoScan)]
    return fn


from vtgpu import StereoCfg, StereoFeature, StereoGeom  # noqa: E402  (one ctypes class per C struct)


def stereo_cfg(w, h, level=1, min_disp=2.0, max_disp=100.0, min_ncc=0.8, ambiguity=0.9, fb_max=0.5, max_reproj=1.0):
    lk = LkCfg(15, 15, 10, 0.01, 1e-4, 40.0, 0)
    return StereoCfg(level, min_disp, max_disp, min_ncc, ambiguity, fb_max, max_reproj, lk,
                     StereoGeom(500.0, 500.0, w / 2.0, h / 2.0, 0.12))
