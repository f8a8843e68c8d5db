"""vtgpu: Python host binding of the MI355X vision path's C-ABI (include/vtgpu.h, libvtgpu.so).

The mirror of the reference's operator surface for this path: pyramids (tk_build / img_reduce, src/kern/track.c:584-629,
src/kern/img.h:30-33), sparse pyramidal LK (lk_track / lk_flowback, src/kern/lk.h:45-53), NCC search (search_slot /
tk_correlate, src/kern/track.c:1661-1726, track.h:57-59). All compute runs in libvtgpu.so on the GPU; this module only
moves bytes and checks return codes. It raises if the library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_GPU = os.path.join(HERE, "libvtgpu.so")
LIB_SYNTH = os.path.join(HERE, "libvtsynth.so")

PAD_X, PAD_Y, Q6 = 32, 4, 64
STORE_Q6, STORE_F32, STORE_F16, STORE_BF16, STORE_Q6W, STORE_Q6F, STORE_F32_U8 = 0, 1, 2, 3, 4, 5, 6
BUILD_CONTRAST, BUILD_PRE_GAUSS = 1, 2
SOLVE_F64, SOLVE_MIXED = 0, 1
GRAD_SOBEL, GRAD_SCHARR = 0, 1
LK_OK, LK_OUT, LK_LOWEIG, LK_RESIDUAL, LK_FLOWBACK = 0, 1, 2, 4, 8

LK_OUT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("min_eig", "<f4"), ("residual", "<f4"),
                         ("cov_xx", "<f4"), ("cov_xy", "<f4"), ("cov_yy", "<f4"),
                         ("iters", "<u2"), ("status", "<u2")])
CELL_DTYPE = np.dtype([("v", "<f4"), ("x", "<i4"), ("y", "<i4")])


class Plane(C.Structure):
    _fields_ = [("p", C.POINTER(C.c_float)), ("w", C.c_int), ("h", C.c_int), ("stride", C.c_int)]


class LkCfg(C.Structure):
    _fields_ = [("w", C.c_int), ("h", C.c_int), ("max_iter", C.c_int), ("eps", C.c_float),
                ("min_eig", C.c_float), ("max_residual", C.c_float), ("prefetch", C.c_int)]


class Found(C.Structure):
    _fields_ = [("hx", C.c_int), ("hy", C.c_int), ("ncc", C.c_double), ("second", C.c_double),
                ("ox", C.c_double), ("oy", C.c_double), ("cells", C.c_int), ("best_x", C.c_int),
                ("best_y", C.c_int), ("best_v", C.c_float)]


class Tune(C.Structure):
    _fields_ = [("pyr_rows", C.c_int), ("lk_waves", C.c_int), ("pyr_us", C.c_float),
                ("lk_us", C.c_float), ("pyr_us_default", C.c_float), ("lk_us_default", C.c_float),
                ("candidates", C.c_int), ("lk_ppw", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class StillCfg(C.Structure):
    """struct vtg_still_cfg (= the reference's struct still_cfg)."""
    _fields_ = [("w", C.c_int), ("h", C.c_int), ("avg", C.c_int), ("motion_px", C.c_float),
                ("motion_grey", C.c_float), ("threads", C.c_int)]


STILL_CELL_DTYPE = np.dtype([("u", "<f4"), ("v", "<f4"), ("diff", "<f4"), ("state", "<u4"), ("sigma", "<f4"),
                             ("pad", "<f4")])

class StereoGeom(C.Structure):
    _fields_ = [(n, C.c_longdouble) for n in ("fx", "fy", "cx", "cy", "baseline")]


class StereoCfg(C.Structure):
    """struct vtg_stereo_cfg (the tracker's stereo settings)"""
    _fields_ = [("stereo_level", C.c_int), ("min_disp", C.c_float), ("max_disp", C.c_float),
                ("min_ncc", C.c_float), ("ambiguity", C.c_float), ("fb_max", C.c_float),
                ("max_reproj", C.c_float), ("lk", LkCfg), ("geom", StereoGeom)]


class StereoFeature(C.Structure):
    """struct vtg_stereo_feature (one feature of the stereo step, in and out)"""
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("disparity", C.c_float), ("flags", C.c_uint32),
                ("xr", C.c_float), ("yr", C.c_float), ("z", C.c_float), ("pad", C.c_float),
                ("p", C.c_longdouble * 3)]

    def key(self):
        return (self.x, self.y, self.disparity, self.flags, self.xr, self.yr, self.z)


class Msg(C.Structure):
    """struct vtg_msg (= the reference's struct vt_msg, 88 bytes)."""
    _fields_ = [("magic", C.c_uint16), ("type", C.c_uint16), ("dest", C.c_uint16), ("priority", C.c_uint16),
                ("algorithm", C.c_uint32), ("length", C.c_uint32), ("instance", C.c_uint64),
                ("acq_ns", C.c_uint64), ("pub_ns", C.c_uint64), ("slot", C.c_uint32), ("cmd", C.c_uint32),
                ("roi", C.c_float * 4), ("sc_key", C.c_uint32), ("sc_value", C.c_float), ("camera", C.c_uint32),
                ("generated", C.c_uint32), ("check", C.c_uint64)]


MSG_MAGIC, DESC_RING, MSG_FRAME, PIPE_PYR, PIPE_ALL, PRIO_ORDERED, PRIO_LIVE = 0x4d54, 256, 1, 1, 15, 0, 1
EXC_NONE, EXC_BAD_CHECK, EXC_BAD_TYPE, EXC_BAD_SLOT, EXC_BAD_LENGTH, EXC_INSTANCE_ORDER = range(6)


class DoorCfg(C.Structure):
    _fields_ = [("cameras", C.c_int), ("w", C.c_int * 3), ("h", C.c_int * 3), ("levels", C.c_int * 3),
                ("npoints", C.c_int * 3), ("max_points", C.c_int), ("slots", C.c_uint32),
                ("slot_bytes", C.c_uint64), ("dev_slots", C.c_int), ("stereo", C.c_int), ("max_inflight", C.c_int)]


class DoorState(C.Structure):
    _fields_ = [("frames_taken", C.c_uint64), ("exceptions", C.c_uint64), ("last_exception", C.c_uint32),
                ("pad", C.c_uint32), ("skipped", C.c_uint64), ("submitted", C.c_uint64), ("in_flight", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class DoorResult(C.Structure):
    _fields_ = [("instance", C.c_uint64), ("acq_ns", C.c_uint64), ("camera", C.c_uint32), ("slot", C.c_uint32),
                ("nrec", C.c_int), ("pad", C.c_int), ("latency_ms", C.c_double)]


STEREO_SCAN_DTYPE = np.dtype([("disparity", "<f4"), ("ncc", "<f4"), ("second", "<f4")])

STILL_MOTION_FIELDS = ["cols", "rows", "cell_at", "tracked", "moving", "count_mean", "drift_u", "drift_v",
                       "drift_rms", "diff_median", "diff_limit", "pad", "cycles_vectors", "cycles_mean", "frames"]


class StillMotion(C.Structure):
    """struct vtg_motion (= the leading fields of the reference's struct vt_motion)."""
    _fields_ = [(n, C.c_uint32) for n in STILL_MOTION_FIELDS[:5]] + \
               [(n, C.c_float) for n in STILL_MOTION_FIELDS[5:12]] + \
               [(n, C.c_uint64) for n in STILL_MOTION_FIELDS[12:]]

    def as_dict(self):
        return {k: getattr(self, k) for k in ("tracked", "moving", "count_mean", "drift_u", "drift_v", "drift_rms",
                                              "diff_median", "diff_limit", "frames")}


class SrvCfg(C.Structure):
    _fields_ = [("device", C.c_int), ("hbm_budget", C.c_size_t), ("lanes", C.c_int), ("max_batch", C.c_int),
                ("lk", LkCfg), ("solve_mode", C.c_int), ("fb_max", C.c_float), ("tmpl_w", C.c_int),
                ("tmpl_h", C.c_int)]


class SrvStats(C.Structure):
    _fields_ = [("submitted", C.c_long), ("completed", C.c_long), ("batches", C.c_long), ("lk_points", C.c_long),
                ("max_batch", C.c_int), ("max_inflight", C.c_int), ("hbm_budget", C.c_size_t),
                ("hbm_resident", C.c_size_t), ("hbm_in_use", C.c_size_t), ("hbm_peak", C.c_size_t)]


class VtgError(RuntimeError):
    pass


_lib = None
_synth = None


def lib():
    """libvtgpu.so, loaded once. Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_GPU):
        raise VtgError(f"{LIB_GPU} is missing: build it (python intel-phi-visiontrack_amd/build.py)")
    L = C.CDLL(LIB_GPU)
    P, I, F, V, S = C.c_void_p, C.c_int, C.c_float, C.c_void_p, C.c_size_t
    sig = {
        "vtg_ctx_create": [I, C.POINTER(V)], "vtg_ctx_destroy": [V], "vtg_ctx_sync": [V], "vtg_ctx_wait": [V, V],
        "vtg_alloc": [V, S, C.POINTER(V)], "vtg_free": [V, V], "vtg_upload": [V, V, V, S],
        "vtg_download": [V, V, V, S], "vtg_copy_d2d": [V, V, V, S], "vtg_download_async": [V, V, V, S],
        "vtg_upload_async": [V, V, V, S],
        "vtg_host_alloc": [V, S, C.POINTER(V)], "vtg_host_free": [V, V],
        "vtg_timing_enable": [V, I], "vtg_timing_read": [V, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_long)],
        "vtg_timing_reset": [V],
        "vtg_pyr_create": [V, I, I, I, I, C.POINTER(V)], "vtg_pyr_destroy": [V, V],
        "vtg_pyr_wrap": [V, I, I, V, I, C.POINTER(V)],
        "vtg_ctx_mark": [V, C.POINTER(C.c_uint64)], "vtg_ctx_mark_wait": [V, C.c_uint64],
        "vtg_pyr_build": [V, V, V, I, I], "vtg_pyr_build_many": [V, I, V, V, V, I],
        "vtg_pyr_build_ex": [V, V, V, I, I, I], "vtg_pyr_build_many_ex": [V, I, V, V, V, I, I],
        "vtg_img_gauss5": [V, C.POINTER(Plane), C.POINTER(Plane), I, I],
        "vtg_img_norm": [V, C.POINTER(Plane), C.POINTER(Plane), V, I, I],
        "vtg_img_minmax": [V, C.POINTER(Plane), V, I, I],
        "vtg_pyr_level": [V, I, C.POINTER(V), C.POINTER(I), C.POINTER(I), C.POINTER(S), C.POINTER(I)],
        "vtg_pyr_read_level": [V, V, I, C.POINTER(Plane)],
        "vtg_lk_track": [V, V, V, I, C.POINTER(LkCfg), I, V, V, F, V],
        "vtg_lk_track_ex": [V, V, V, I, C.POINTER(LkCfg), I, V, V, F, I, V],
        "vtg_lk_track_many": [V, I, V, V, V, V, C.POINTER(LkCfg), V, V, F, I, V], "vtg_lk_set_waves": [V, I],
        "vtg_lk_set_ppw": [V, I], "vtg_lk_set_order": [V, I], "vtg_lk_set_gradient": [V, I],
        "vtg_pyr_lk_fused": [V, I, V, V, V, V, V, V, V, V, C.POINTER(LkCfg), V, V, F, I, V],
        "vtg_ctx_device_error": [V, C.POINTER(I)],
        "vtg_ncc_template": [V, V, I, I, I, I, I, C.POINTER(V)], "vtg_ncc_template_destroy": [V, V],
        "vtg_ncc_template_update": [V, V, V, I, I, I], "vtg_ncc_template_update_many": [V, I, V, V, V, V, V],
        "vtg_ncc_template_read": [V, V, V, C.POINTER(F)],
        "vtg_ncc_search": [V, V, V, I, I, I, I, I, V, C.POINTER(Found)],
        "vtg_ncc_search_async": [V, V, V, I, I, I, I, I, V, V],
        "vtg_ncc_search_many": [V, I, V, V, V, I, V, V], "vtg_ncc_search_many_split": [V, V, I, V, V, V, I, V, V],
        "vtg_tune_pair": [V, V, V, V, V, I, I, C.POINTER(LkCfg), I, V, I, C.POINTER(Tune)],
        "vtg_tune_get": [V, V, I, C.POINTER(LkCfg), C.POINTER(Tune)], "vtg_tune_reset": [V],
        "vtg_srv_create": [C.POINTER(SrvCfg), C.POINTER(V)], "vtg_srv_destroy": [V],
        "vtg_srv_open": [V, I, I, I, I, C.POINTER(I)], "vtg_srv_close": [V, I],
        "vtg_srv_submit": [V, I, V, I, I, V, V, I, I, C.POINTER(C.c_uint64)],
        "vtg_srv_poll": [V, C.c_uint64, V, C.POINTER(I), C.POINTER(Found), C.POINTER(C.c_double)],
        "vtg_srv_wait": [V, C.c_uint64, V, C.POINTER(I), C.POINTER(Found), C.POINTER(C.c_double)],
        "vtg_srv_get_stats": [V, C.POINTER(SrvStats)],
        "vtg_graph_begin": [V], "vtg_graph_end": [V, C.POINTER(V)], "vtg_graph_launch": [V, V],
        "vtg_graph_destroy": [V, V],
        "vtg_ncc_score": [V, V, V, I, V, I, V],
        "vtg_img_u8_to_q6": [V, C.POINTER(Plane), V, I, I, I],
        "vtg_img_reduce": [V, C.POINTER(Plane), C.POINTER(Plane), I, I],
        "vtg_lk_track_planes": [V, V, V, I, C.POINTER(LkCfg), I, V, V, V, V, F, V],
        "vtg_correlate_planes": [V, C.POINTER(Plane), V, I, I, C.c_double, I, V, V, V],
        "vtg_still_create": [V, C.POINTER(StillCfg), C.POINTER(V)], "vtg_still_destroy": [V, V],
        "vtg_still_first": [V], "vtg_still_size": [V, C.POINTER(I), C.POINTER(I)],
        "vtg_still_vectors": [V, V, V, I], "vtg_still_vectors_plane": [V, V, C.POINTER(Plane)],
        "vtg_still_decide": [V, V], "vtg_still_update": [V, V, V, I],
        "vtg_still_update_plane": [V, V, C.POINTER(Plane)], "vtg_still_finish": [V, V],
        "vtg_still_frame": [V, V, V, I], "vtg_still_frame_many": [V, I, V, V, V], "vtg_still_cells": [V, V, V],
        "vtg_still_mean": [V, V, C.POINTER(Plane)], "vtg_still_motion": [V, C.POINTER(StillMotion)],
        "vtg_still_hash": [V, V, C.c_uint64, C.POINTER(C.c_uint64)],
        "vtg_stereo_scan_many": [V, V, V, I, I, V, I, I, V],
        "vtg_stereo_scan_planes": [V, C.POINTER(Plane), C.POINTER(Plane), I, V, V, I, I, V],
        "vtg_stereo_match": [V, V, V, C.POINTER(StereoCfg), I, V],
        "vtg_door_create": [V, C.POINTER(DoorCfg), V, C.POINTER(V)], "vtg_door_destroy": [V],
        "vtg_door_ingest": [V, V, C.c_uint64, V, C.POINTER(DoorState)],
        "vtg_door_result": [V, I, C.POINTER(DoorResult), V, I],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = I
    L.vtg_msg_check.argtypes = [C.POINTER(Msg)]
    L.vtg_msg_check.restype = C.c_uint64
    L.vtg_ctx_error.argtypes = [V]
    L.vtg_ctx_error.restype = C.c_char_p
    L.vtg_ctx_stream.argtypes = [V]
    L.vtg_ctx_stream.restype = V
    L.vtg_srv_error.argtypes = [V]
    L.vtg_srv_error.restype = C.c_char_p
    _lib = L
    return L


# every symbol include/vtgpu.h declares
EXPORTS = ["vtg_ctx_create", "vtg_ctx_destroy", "vtg_ctx_sync", "vtg_ctx_wait", "vtg_ctx_error", "vtg_ctx_stream", "vtg_alloc",
           "vtg_free", "vtg_upload", "vtg_download", "vtg_copy_d2d", "vtg_download_async", "vtg_upload_async", "vtg_host_alloc", "vtg_host_free", "vtg_timing_enable", "vtg_timing_read",
           "vtg_timing_reset", "vtg_pyr_create", "vtg_pyr_destroy", "vtg_pyr_wrap", "vtg_ctx_mark", "vtg_ctx_mark_wait", "vtg_pyr_build", "vtg_pyr_build_many", "vtg_pyr_build_ex",
           "vtg_pyr_build_many_ex", "vtg_img_gauss5", "vtg_img_norm", "vtg_img_minmax", "vtg_pyr_level",
           "vtg_pyr_read_level", "vtg_lk_track", "vtg_lk_track_ex", "vtg_lk_track_many", "vtg_lk_set_waves", "vtg_lk_set_ppw", "vtg_lk_set_order", "vtg_lk_set_gradient", "vtg_pyr_lk_fused", "vtg_ctx_device_error", "vtg_ncc_template", "vtg_ncc_template_destroy",
           "vtg_ncc_template_read", "vtg_ncc_template_update", "vtg_ncc_template_update_many", "vtg_ncc_search", "vtg_ncc_search_async", "vtg_ncc_search_fixed", "vtg_ncc_search_many", "vtg_ncc_search_many_split",
           "vtg_tune_pair", "vtg_tune_get", "vtg_tune_reset", "vtg_srv_create", "vtg_srv_destroy", "vtg_srv_error",
           "vtg_srv_open", "vtg_srv_close", "vtg_srv_submit", "vtg_srv_poll", "vtg_srv_wait", "vtg_srv_get_stats",
           "vtg_ncc_score",
           "vtg_graph_begin", "vtg_graph_end", "vtg_graph_launch", "vtg_graph_destroy", "vtg_img_u8_to_q6", "vtg_img_reduce",
           "vtg_lk_track_planes", "vtg_correlate_planes",
           "vtg_still_create", "vtg_still_destroy", "vtg_still_first", "vtg_still_size", "vtg_still_vectors",
           "vtg_still_vectors_plane", "vtg_still_decide", "vtg_still_update", "vtg_still_update_plane",
           "vtg_still_finish", "vtg_still_frame", "vtg_still_frame_many", "vtg_still_cells", "vtg_still_mean", "vtg_still_motion",
           "vtg_still_hash", "vtg_stereo_scan_many", "vtg_stereo_scan_planes",
           "vtg_stereo_match", "vtg_msg_check", "vtg_door_create", "vtg_door_destroy", "vtg_door_ingest",
           "vtg_door_result"]


def synth():
    """libvtsynth.so (the synthetic-frame generator)."""
    global _synth
    if _synth is not None:
        return _synth
    if not os.path.exists(LIB_SYNTH):
        raise VtgError(f"{LIB_SYNTH} is missing: build it")
    L = C.CDLL(LIB_SYNTH)
    D = C.c_double
    L.vts_frame_pair.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_void_p,
                                 C.c_void_p, D, D, C.c_void_p, C.c_int, C.c_int]
    L.vts_frame.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_void_p, C.c_void_p, D,
                            C.c_uint64, C.c_int]
    L.vts_draw_motion.argtypes = [C.c_uint64, D, D, C.c_void_p]
    L.vts_truth.argtypes = [C.c_void_p, C.c_int, C.c_int, D, D, C.POINTER(D), C.POINTER(D)]
    L.vts_grid.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    L.vts_points.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    L.vts_scene_frame.argtypes = [C.c_int, D, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint64, C.c_int]
    L.vts_scene_map.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, D, D, D, D, C.POINTER(D), C.POINTER(D)]
    L.vts_scene_object.argtypes = [C.c_int, C.c_int, C.c_int, D, C.POINTER(D), C.POINTER(D), C.POINTER(D),
                                   C.POINTER(D)]
    L.vts_rand.argtypes = [C.POINTER(C.c_uint64)]
    L.vts_rand.restype = C.c_uint64
    L.vts_uniform.argtypes = [C.POINTER(C.c_uint64)]
    L.vts_uniform.restype = D
    _synth = L
    return L


class Scene(C.Structure):
    _fields_ = [("nblobs", C.c_int), ("sig_lo", C.c_double), ("sig_hi", C.c_double), ("amp_lo", C.c_double),
                ("amp_hi", C.c_double)]


class Motion(C.Structure):
    _fields_ = [("dx", C.c_double), ("dy", C.c_double), ("theta_deg", C.c_double)]


BLOB_SCENE = Scene(200, 4.0, 40.0, 20.0, 120.0)


def frame_pair(w, h, seed, max_shift=8.0, max_rot=2.0, noise=2.0, motion=None, flat_amp=0.0, flat=None,
               threads=8):
    """(f0, f1, motion): the benchmark's seeded blob frames, f1 = f0 moved rigidly plus noise."""
    S = synth()
    if motion is None:
        motion = Motion()
        S.vts_draw_motion(seed, max_shift, max_rot, C.byref(motion))
    f0 = np.zeros((h, w), np.uint8)
    f1 = np.zeros((h, w), np.uint8)
    fl = None if flat is None else np.ascontiguousarray(flat, np.int32)
    S.vts_frame_pair(f0.ctypes.data, f1.ctypes.data, w, h, w, seed, C.byref(BLOB_SCENE), C.byref(motion), noise,
                     flat_amp, None if fl is None else fl.ctypes.data, 0 if fl is None else len(fl) // 4, threads)
    return f0, f1, motion


def frame(w, h, seed, dx, dy, theta_deg=0.0, noise=2.0, noise_seed=0, threads=8):
    """One frame of a sequence: the seed's blob scene moved by (dx, dy, theta) about the centre, plus noise."""
    f = np.zeros((h, w), np.uint8)
    m = Motion(dx, dy, theta_deg)
    synth().vts_frame(f.ctypes.data, w, h, w, seed, C.byref(BLOB_SCENE), C.byref(m), noise, noise_seed, threads)
    return f


SCENES = ["translate", "camera", "shapes", "brightness", "noise", "occlusion"]


def scene_frame(scene, t, w, h, seed=1, noise_seed=0, threads=8):
    """Frame of a layered scene (vts_scene_frame) at time t seconds."""
    f = np.zeros((h, w), np.uint8)
    rc = synth().vts_scene_frame(SCENES.index(scene), t, f.ctypes.data, w, h, w, seed, noise_seed, threads)
    if rc:
        raise VtgError(f"vts_scene_frame: {rc}")
    return f


def scene_map(scene, w, h, t0, x, y, t1, seed=1):
    """(layer, mx, my): where the point (x, y) of the frame at t0 is at t1."""
    mx, my = C.c_double(), C.c_double()
    k = synth().vts_scene_map(SCENES.index(scene), w, h, seed, t0, float(x), float(y), t1, C.byref(mx), C.byref(my))
    return k, mx.value, my.value


def scene_object(scene, w, h, t):
    v = [C.c_double() for _ in range(4)]
    synth().vts_scene_object(SCENES.index(scene), w, h, t, *[C.byref(x) for x in v])
    return tuple(x.value for x in v)


def truth(motion, w, h, xs, ys):
    S = synth()
    u, v = C.c_double(), C.c_double()
    out = np.zeros((len(xs), 2))
    for i, (x, y) in enumerate(zip(xs, ys)):
        S.vts_truth(C.byref(motion), w, h, float(x), float(y), C.byref(u), C.byref(v))
        out[i] = (u.value, v.value)
    return out


def grid(nx, ny, w, h):
    xs = np.zeros(nx * ny, np.float32)
    ys = np.zeros(nx * ny, np.float32)
    synth().vts_grid(nx, ny, w, h, xs.ctypes.data, ys.ctypes.data)
    return xs, ys


def points(seed, n, w, h, margin):
    xs = np.zeros(n, np.float32)
    ys = np.zeros(n, np.float32)
    synth().vts_points(seed, n, w, h, margin, xs.ctypes.data, ys.ctypes.data)
    return xs, ys


def lk_cfg(w=21, h=21, max_iter=10, eps=0.01, min_eig=1e-4, max_residual=40.0):
    return LkCfg(w, h, max_iter, eps, min_eig, max_residual, 0)


class Context:
    """A vtg_ctx: one device, one stream."""

    def __init__(self, device=0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.vtg_ctx_create(device, C.byref(h))
        if rc:
            raise VtgError(f"vtg_ctx_create({device}) = {rc}")
        self.h = h

    def check(self, rc, what=""):
        if rc:
            raise VtgError(f"{what} failed ({rc}): {self.L.vtg_ctx_error(self.h).decode()}")

    def close(self):
        if self.h:
            self.L.vtg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self.check(self.L.vtg_ctx_sync(self.h), "sync")

    def wait(self, other):
        """Later work of this context starts after everything submitted to `other` so far."""
        self.check(self.L.vtg_ctx_wait(self.h, other.h), "wait")

    def mark(self):
        """A point after the work submitted so far; mark_wait(m) returns once that work has completed."""
        m = C.c_uint64()
        self.check(self.L.vtg_ctx_mark(self.h, C.byref(m)), "mark")
        return m.value

    def mark_wait(self, m):
        self.check(self.L.vtg_ctx_mark_wait(self.h, m), "mark_wait")

    # memory
    def alloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.L.vtg_alloc(self.h, max(int(nbytes), 1), C.byref(p)), "alloc")
        return DevBuf(self, p, int(nbytes))

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        b = self.alloc(arr.nbytes)
        self.check(self.L.vtg_upload(self.h, b.ptr, arr.ctypes.data, arr.nbytes), "upload")
        b.dtype, b.shape = arr.dtype, arr.shape
        return b

    # timing
    def timing(self, on=True):
        self.check(self.L.vtg_timing_enable(self.h, 1 if on else 0))

    def timing_read(self, kernel):
        ms, n = C.c_double(), C.c_long()
        self.check(self.L.vtg_timing_read(self.h, kernel.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def timing_reset(self):
        self.check(self.L.vtg_timing_reset(self.h))

    # pyramids
    def pyramid(self, w, h, levels, storage=STORE_Q6):
        return Pyramid(self, w, h, levels, storage)

    def pyramid_wrap(self, frame, w, h, pitch=None):
        """A one-level Q6 pyramid whose level 0 is the device frame (DevBuf) itself, read in place (vtg_pyr_wrap)."""
        return Pyramid(self, w, h, 1, STORE_Q6, wrap=(frame, pitch or w))

    def stereo_scan_many(self, left, right, level, xy, dmin, dmax):
        """The row scan of n features (xy: DevBuf of 2 n floats) at `level`; returns (n,) STEREO_SCAN_DTYPE."""
        n = xy.nbytes // 8
        out = self.alloc(max(12 * n, 16))
        self.check(self.L.vtg_stereo_scan_many(self.h, left.h, right.h, level, n, xy.ptr, dmin, dmax, out.ptr),
                   "stereo_scan_many")
        res = out.get(np.zeros(n, STEREO_SCAN_DTYPE)).copy()
        out.free()
        return res

    def stereo_scan_planes(self, left, right, xs, ys, dmin, dmax):
        xs = np.ascontiguousarray(xs, np.float32)
        ys = np.ascontiguousarray(ys, np.float32)
        out = np.zeros(len(xs), STEREO_SCAN_DTYPE)
        self.check(self.L.vtg_stereo_scan_planes(self.h, C.byref(left.c), C.byref(right.c), len(xs), xs.ctypes.data,
                                                 ys.ctypes.data, dmin, dmax, out.ctypes.data), "stereo_scan_planes")
        return out

    def still_frame_many(self, stills, pyrs, levels):
        """One frame for each of several still cameras (vtg_still_frame_many)."""
        n = len(stills)
        S = (C.c_void_p * n)(*[x.h for x in stills])
        P = (C.c_void_p * n)(*[x.h for x in pyrs])
        Lv = (C.c_int * n)(*levels)
        self.check(self.L.vtg_still_frame_many(self.h, n, S, P, Lv), "still_frame_many")

    def stereo_match(self, left, right, cfg, feats):
        """The tracker's stereo step over `feats` (a ctypes array of StereoFeature, read and written)."""
        self.check(self.L.vtg_stereo_match(self.h, left.h, right.h, C.byref(cfg), len(feats), C.addressof(feats)),
                   "stereo_match")
        return feats

    def still(self, w, h, avg=16, motion_px=0.5, motion_grey=4.0):
        return Still(self, StillCfg(w, h, avg, motion_px, motion_grey, 1))

    def build_many(self, pyrs, frames, flags=0):
        """Build several pyramids in one launch per level: frames are DevBufs (pitch = width)."""
        n = len(pyrs)
        P = (C.c_void_p * n)(*[p.h for p in pyrs])
        F = (C.c_void_p * n)(*[f.ptr for f in frames])
        pitch = (C.c_int * n)(*[p.w for p in pyrs])
        self.check(self.L.vtg_pyr_build_many_ex(self.h, n, P, F, pitch, 1, flags), "pyr_build_many")

    def pyr_lk_fused(self, Is, Js, frames0, frames1, tops, offsets, cfg, xy, out, guess=None, fb_max=0.0,
                     solve=SOLVE_F64):
        """Build I[k] from frames0[k] and J[k] from frames1[k] (DevBufs, pitch = width) and track every pair's points
        in one persistent launch; out: DevBuf of lk_out records."""
        n = len(Is)
        A = lambda xs: (C.c_void_p * n)(*xs)
        pitch0 = (C.c_int * n)(*[p.w for p in Is])
        pitch1 = (C.c_int * n)(*[p.w for p in Js])
        T = (C.c_int * n)(*tops)
        O = (C.c_int * (n + 1))(*offsets)
        self.check(self.L.vtg_pyr_lk_fused(self.h, n, A([p.h for p in Is]), A([p.h for p in Js]),
                                           A([f.ptr for f in frames0]), A([f.ptr for f in frames1]), pitch0, pitch1,
                                           T, O, C.byref(cfg), xy.ptr, guess.ptr if guess is not None else None,
                                           fb_max, solve, out.ptr), "pyr_lk_fused")

    def device_error(self):
        code = C.c_int()
        self.check(self.L.vtg_ctx_device_error(self.h, C.byref(code)), "device_error")
        return code.value

    # LK
    def lk_track(self, I, J, top, cfg, xy, guess=None, fb_max=0.0, out=None, solve=SOLVE_F64):
        """xy, guess, out: DevBuf. Returns out."""
        n = xy.nbytes // 8
        if out is None:
            out = self.alloc(n * 32)
        self.check(self.L.vtg_lk_track_ex(self.h, I.h, J.h, top, C.byref(cfg), n, xy.ptr,
                                          guess.ptr if guess is not None else None, fb_max, solve, out.ptr),
                   "lk_track")
        return out

    def lk_track_many(self, Is, Js, tops, offsets, cfg, xy, guess=None, fb_max=0.0, out=None, solve=SOLVE_F64):
        """Points of several pairs in one launch: offsets (len = pairs + 1) slice the DevBuf xy."""
        n = len(Is)
        if out is None:
            out = self.alloc(32 * max(1, offsets[-1]))
        PI = (C.c_void_p * n)(*[p.h for p in Is])
        PJ = (C.c_void_p * n)(*[p.h for p in Js])
        T = (C.c_int * n)(*tops)
        O = (C.c_int * (n + 1))(*offsets)
        self.check(self.L.vtg_lk_track_many(self.h, n, PI, PJ, T, O, C.byref(cfg), xy.ptr,
                                            guess.ptr if guess is not None else None, fb_max, solve, out.ptr),
                   "lk_track_many")
        return out

    def lk_track_host(self, I, J, top, cfg, xs, ys, gx=None, gy=None, fb_max=0.0, solve=SOLVE_F64):
        xy = np.stack([np.asarray(xs, np.float32), np.asarray(ys, np.float32)], 1).ravel()
        dxy = self.to_device(xy)
        dg = None
        if gx is not None:
            dg = self.to_device(np.stack([np.asarray(gx, np.float32), np.asarray(gy, np.float32)], 1).ravel())
        out = self.lk_track(I, J, top, cfg, dxy, dg, fb_max, solve=solve)
        res = out.get(np.zeros(len(xs), LK_OUT_DTYPE))
        for b in (dxy, dg, out):
            if b is not None:
                b.free()
        return res

    # NCC
    def ncc_template(self, pyr, level, x0, y0, tw, th):
        return Template(self, pyr, level, x0, y0, tw, th)

    def ncc_templates_update(self, tmpls, pyrs, levels, corners):
        """Re-cut tmpls[k] from levels[k] of pyrs[k] at corners[k] = (x0, y0), one launch per run of up to 8."""
        n = len(tmpls)
        T = (C.c_void_p * n)(*[t.h for t in tmpls])
        P = (C.c_void_p * n)(*[p.h for p in pyrs])
        Lv = (C.c_int * n)(*levels)
        X = (C.c_int * n)(*[int(c[0]) for c in corners])
        Y = (C.c_int * n)(*[int(c[1]) for c in corners])
        self.check(self.L.vtg_ncc_template_update_many(self.h, n, T, P, Lv, X, Y), "ncc_template_update_many")
        for t, lv in zip(tmpls, levels):
            t.level = lv

    def ncc_search(self, pyr, tmpl, x0, x1, y0, y1, step=1, cells=None):
        f = Found()
        self.check(self.L.vtg_ncc_search(self.h, pyr.h, tmpl.h, x0, x1, y0, y1, step,
                                         cells.ptr if cells is not None else None, C.byref(f)), "ncc_search")
        return f

    def ncc_search_async(self, pyr, tmpl, x0, x1, y0, y1, step, cells, found_dev):
        self.check(self.L.vtg_ncc_search_async(self.h, pyr.h, tmpl.h, x0, x1, y0, y1, step,
                                               cells.ptr if cells is not None else None, found_dev.ptr),
                   "ncc_search_async")

    def ncc_search_many(self, pyrs, tmpls, regions, step, found_dev, cells=None):
        """Searches k over regions[k] = (x0, x1, y0, y1) of pyrs[k] with tmpls[k]; picks to the device records
        found_dev (a DevBuf of len(pyrs) Found records). Stream-ordered."""
        n = len(pyrs)
        P = (C.c_void_p * n)(*[p.h for p in pyrs])
        T = (C.c_void_p * n)(*[t.h for t in tmpls])
        R = (C.c_int * (4 * n))(*[int(v) for r in regions for v in r])
        CL = None if cells is None else (C.c_void_p * n)(*[c.ptr if c is not None else None for c in cells])
        self.check(self.L.vtg_ncc_search_many(self.h, n, P, T, R, step, CL, found_dev.ptr), "ncc_search_many")

    def ncc_search_many_split(self, finish, pyrs, tmpls, regions, step, found_dev, cells=None):
        """ncc_search_many with the cross terms on this context and the scoring and finish on context `finish`."""
        n = len(pyrs)
        P = (C.c_void_p * n)(*[p.h for p in pyrs])
        T = (C.c_void_p * n)(*[t.h for t in tmpls])
        R = (C.c_int * (4 * n))(*[int(v) for r in regions for v in r])
        CL = None if cells is None else (C.c_void_p * n)(*[c.ptr if c is not None else None for c in cells])
        self.check(self.L.vtg_ncc_search_many_split(self.h, finish.h, n, P, T, R, step, CL, found_dev.ptr),
                   "ncc_search_many_split")

    # autotuner
    def tune_pair(self, I, J, f0, f1, top, cfg, xy, solve=SOLVE_F64):
        """Time the schedules on this pair (device frames f0, f1: DevBufs, pitch = width) and points (DevBuf) and keep
        the fastest for its size class; returns what was measured."""
        t = Tune()
        self.check(self.L.vtg_tune_pair(self.h, I.h, J.h, f0.ptr, f1.ptr, I.w, top, C.byref(cfg), xy.nbytes // 8,
                                        xy.ptr, solve, C.byref(t)), "tune_pair")
        return t.as_dict()

    def tune_get(self, I, top, cfg):
        t = Tune()
        self.check(self.L.vtg_tune_get(self.h, I.h, top, C.byref(cfg), C.byref(t)), "tune_get")
        return t.as_dict()

    def tune_reset(self):
        self.check(self.L.vtg_tune_reset(self.h), "tune_reset")

    # graphs
    def graph_begin(self):
        self.check(self.L.vtg_graph_begin(self.h), "graph_begin")

    def graph_end(self):
        g = C.c_void_p()
        self.check(self.L.vtg_graph_end(self.h, C.byref(g)), "graph_end")
        return g

    def graph_launch(self, g):
        self.check(self.L.vtg_graph_launch(self.h, g), "graph_launch")

    def graph_destroy(self, g):
        self.check(self.L.vtg_graph_destroy(self.h, g), "graph_destroy")

    def ncc_score(self, pyr, tmpl, xy, half=False):
        xy = np.ascontiguousarray(xy, np.int32).ravel()
        n = len(xy) // 2
        d = self.to_device(xy)
        o = self.alloc(8 * n)
        self.check(self.L.vtg_ncc_score(self.h, pyr.h, tmpl.h, n, d.ptr, 1 if half else 0, o.ptr), "ncc_score")
        r = o.get(np.zeros(n, np.float64))
        d.free()
        o.free()
        return r


class PinnedArray:
    """A numpy array over pinned host memory (vtg_host_alloc); .arr is the array, free() releases it."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        p = C.c_void_p()
        ctx.check(ctx.L.vtg_host_alloc(ctx.h, nbytes, C.byref(p)), "host_alloc")
        self.ptr = p
        buf = (C.c_char * nbytes).from_address(p.value)
        self.arr = np.frombuffer(buf, dtype=dtype).reshape(shape)

    def free(self):
        if self.ptr:
            self.arr = None
            self.ctx.L.vtg_host_free(self.ctx.h, self.ptr)
            self.ptr = None


class DevBuf:
    def __init__(self, ctx, ptr, nbytes):
        self.ctx, self.ptr, self.nbytes = ctx, ptr, nbytes

    def get(self, arr):
        self.ctx.check(self.ctx.L.vtg_download(self.ctx.h, arr.ctypes.data, self.ptr, arr.nbytes), "download")
        return arr

    def get_async(self, ctx, arr):
        """Download into `arr` (a PinnedArray's array) on ctx's stream, no wait."""
        ctx.check(ctx.L.vtg_download_async(ctx.h, arr.ctypes.data, self.ptr, arr.nbytes), "download_async")
        return arr

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        self.ctx.check(self.ctx.L.vtg_upload(self.ctx.h, self.ptr, arr.ctypes.data, arr.nbytes), "upload")

    def put_async(self, ctx, arr):
        """Upload `arr` (a PinnedArray's array) on ctx's stream, no wait."""
        ctx.check(ctx.L.vtg_upload_async(ctx.h, self.ptr, arr.ctypes.data, arr.nbytes), "upload_async")

    def free(self):
        if self.ptr:
            self.ctx.L.vtg_free(self.ctx.h, self.ptr)
            self.ptr = None

    def view(self, offset, nbytes):
        """Bytes [offset, offset + nbytes) of this buffer (not owned: never freed through the view)."""
        assert 0 <= offset and offset + nbytes <= self.nbytes
        return _View(self.ctx, C.c_void_p(self.ptr.value + offset), nbytes)


class _View(DevBuf):
    def free(self):
        self.ptr = None


class Pyramid:
    def __init__(self, ctx, w, h, levels, storage, wrap=None):
        self.ctx, self.w, self.h, self.levels, self.storage = ctx, w, h, levels, storage
        p = C.c_void_p()
        if wrap is None:
            ctx.check(ctx.L.vtg_pyr_create(ctx.h, w, h, levels, storage, C.byref(p)), "pyr_create")
        else:
            self.frame = wrap[0]  # kept alive: the pyramid reads it in place
            ctx.check(ctx.L.vtg_pyr_wrap(ctx.h, w, h, wrap[0].ptr, wrap[1], C.byref(p)), "pyr_wrap")
        self.h = p

    def build(self, frame, on_device=False, pitch=None, flags=0):
        """frame: host uint8 (h, w) array, or a DevBuf with on_device=True. flags: BUILD_CONTRAST | BUILD_PRE_GAUSS."""
        if on_device:
            self.ctx.check(self.ctx.L.vtg_pyr_build_ex(self.ctx.h, self.h, frame.ptr, pitch or self.w, 1, flags),
                           "pyr_build")
        else:
            frame = np.ascontiguousarray(frame, np.uint8)
            self.ctx.check(self.ctx.L.vtg_pyr_build_ex(self.ctx.h, self.h, frame.ctypes.data, frame.strides[0], 0,
                                                       flags), "pyr_build")

    def level_info(self, level):
        p, w, h, pb, eb = C.c_void_p(), C.c_int(), C.c_int(), C.c_size_t(), C.c_int()
        self.ctx.check(self.ctx.L.vtg_pyr_level(self.h, level, C.byref(p), C.byref(w), C.byref(h), C.byref(pb),
                                                C.byref(eb)), "pyr_level")
        return p.value, w.value, h.value, pb.value, eb.value

    def read_level(self, level):
        """Level as a padded fp32 Q6 host plane (numpy buffer + struct), the reference's layout."""
        _, w, h, _, _ = self.level_info(level)
        hp = HostPlane(w, h)
        self.ctx.check(self.ctx.L.vtg_pyr_read_level(self.ctx.h, self.h, level, C.byref(hp.c)), "pyr_read_level")
        return hp

    def destroy(self):
        if self.h:
            self.ctx.L.vtg_pyr_destroy(self.ctx.h, self.h)
            self.h = None


class Template:
    def __init__(self, ctx, pyr, level, x0, y0, tw, th):
        self.ctx, self.tw, self.th, self.level = ctx, tw, th, level
        p = C.c_void_p()
        ctx.check(ctx.L.vtg_ncc_template(ctx.h, pyr.h, level, x0, y0, tw, th, C.byref(p)), "ncc_template")
        self.h = p

    def update(self, pyr, level, x0, y0):
        self.ctx.check(self.ctx.L.vtg_ncc_template_update(self.ctx.h, self.h, pyr.h, level, x0, y0), "template_update")
        self.level = level

    def read(self):
        app = np.zeros(self.tw * self.th, np.float32)
        norm = C.c_float()
        self.ctx.check(self.ctx.L.vtg_ncc_template_read(self.ctx.h, self.h, app.ctypes.data, C.byref(norm)))
        return app, norm.value

    def destroy(self):
        if self.h:
            self.ctx.L.vtg_ncc_template_destroy(self.ctx.h, self.h)
            self.h = None


def plane_stride(w):
    return ((w + 15) & ~15) + 2 * PAD_X


class Still:
    """A vtg_still: one camera's running mean of the level it reads and its 16 x 8 cells' vectors (still.h)."""

    def __init__(self, ctx, cfg):
        self.ctx, self.cfg, self.L = ctx, cfg, ctx.L
        h = C.c_void_p()
        ctx.check(self.L.vtg_still_create(ctx.h, C.byref(cfg), C.byref(h)), "still_create")
        self.h = h
        c, r = C.c_int(), C.c_int()
        ctx.check(self.L.vtg_still_size(h, C.byref(c), C.byref(r)), "still_size")
        self.cols, self.rows = c.value, r.value

    @property
    def ncells(self):
        return self.cols * self.rows

    def first(self):
        return bool(self.L.vtg_still_first(self.h))

    def vectors(self, pyr=None, level=0, plane=None):
        if plane is not None:
            self.ctx.check(self.L.vtg_still_vectors_plane(self.ctx.h, self.h, C.byref(plane.c)), "still_vectors")
        else:
            self.ctx.check(self.L.vtg_still_vectors(self.ctx.h, self.h, pyr.h, level), "still_vectors")

    def decide(self):
        self.ctx.check(self.L.vtg_still_decide(self.ctx.h, self.h), "still_decide")

    def update(self, pyr=None, level=0, plane=None):
        if plane is not None:
            self.ctx.check(self.L.vtg_still_update_plane(self.ctx.h, self.h, C.byref(plane.c)), "still_update")
        else:
            self.ctx.check(self.L.vtg_still_update(self.ctx.h, self.h, pyr.h, level), "still_update")

    def finish(self):
        self.ctx.check(self.L.vtg_still_finish(self.ctx.h, self.h), "still_finish")

    def frame(self, pyr, level=0):
        self.ctx.check(self.L.vtg_still_frame(self.ctx.h, self.h, pyr.h, level), "still_frame")

    def cells(self):
        out = np.zeros(self.ncells, STILL_CELL_DTYPE)
        self.ctx.check(self.L.vtg_still_cells(self.ctx.h, self.h, out.ctypes.data), "still_cells")
        return out

    def mean(self):
        hp = HostPlane(self.cfg.w, self.cfg.h)
        self.ctx.check(self.L.vtg_still_mean(self.ctx.h, self.h, C.byref(hp.c)), "still_mean")
        return hp

    def motion(self):
        m = StillMotion()
        self.ctx.check(self.L.vtg_still_motion(self.h, C.byref(m)), "still_motion")
        return m.as_dict()

    def hash(self, h=0xcbf29ce484222325):
        out = C.c_uint64()
        self.ctx.check(self.L.vtg_still_hash(self.ctx.h, self.h, h, C.byref(out)), "still_hash")
        return out.value

    def destroy(self):
        if self.h:
            self.L.vtg_still_destroy(self.ctx.h, self.h)
            self.h = None


class HostPlane:
    """A padded fp32 Q6 plane in host memory, laid out as struct vt_plane."""

    def __init__(self, w, h):
        self.w, self.h, self.stride = w, h, plane_stride(w)
        self.buf = np.zeros((h + 2 * PAD_Y) * self.stride + 16, dtype=np.float32)
        off = PAD_Y * self.stride + PAD_X
        self.c = Plane(self.buf[off:].ctypes.data_as(C.POINTER(C.c_float)), w, h, self.stride)

    def padded(self):
        return self.buf[: (self.h + 2 * PAD_Y) * self.stride].reshape(self.h + 2 * PAD_Y, self.stride)

    def view(self):
        return self.padded()[PAD_Y:PAD_Y + self.h, PAD_X:PAD_X + self.w]


EAGAIN = -11


class Door:
    """A vtg_door in front of a Server: descriptor rings in, answers in order (the reference's wire format)."""

    def __init__(self, server, cfg, xy):
        self.L, self.server, self.cfg = lib(), server, cfg
        arr = (C.c_void_p * 3)(*[(x.ctypes.data if x is not None else None) for x in (list(xy) + [None] * 3)[:3]])
        self._xy = xy
        h = C.c_void_p()
        rc = self.L.vtg_door_create(server.h, C.byref(cfg), arr, C.byref(h))
        if rc:
            raise VtgError(f"vtg_door_create = {rc}")
        self.h = h

    def ingest(self, desc, frames_pub, slots):
        st = DoorState()
        rc = self.L.vtg_door_ingest(self.h, C.addressof(desc), frames_pub, slots.ctypes.data, C.byref(st))
        if rc:
            raise VtgError(f"vtg_door_ingest = {rc}")
        return st.as_dict()

    def result(self, npoints, wait=True):
        r = DoorResult()
        out = np.zeros(max(npoints, 1), LK_OUT_DTYPE)
        rc = self.L.vtg_door_result(self.h, 1 if wait else 0, C.byref(r), out.ctypes.data, len(out))
        if rc == EAGAIN:
            return None
        if rc:
            raise VtgError(f"vtg_door_result = {rc}")
        return r, out[:r.nrec]

    def destroy(self):
        if self.h:
            self.L.vtg_door_destroy(self.h)
            self.h = None


def msg_check(m):
    return lib().vtg_msg_check(C.byref(m))


class Server:
    """A vtg_srv: streams of frames served by a scheduler thread in continuous batches under a device budget."""

    def __init__(self, budget, lanes=2, max_batch=16, cfg=None, solve=SOLVE_F64, fb_max=0.0, tmpl=(64, 64), device=0):
        self.L = lib()
        c = SrvCfg(device, int(budget), lanes, max_batch, cfg if cfg is not None else lk_cfg(), solve, fb_max,
                   tmpl[0], tmpl[1])
        h = C.c_void_p()
        rc = self.L.vtg_srv_create(C.byref(c), C.byref(h))
        if rc:
            raise VtgError(f"vtg_srv_create = {rc}")
        self.h, self.n = h, {}

    def check(self, rc, what):
        if rc:
            raise VtgError(f"{what} failed ({rc}): {self.L.vtg_srv_error(self.h).decode()}")

    def open(self, w, h, levels, max_points):
        sid = C.c_int()
        self.check(self.L.vtg_srv_open(self.h, w, h, levels, max_points, C.byref(sid)), "srv_open")
        return sid.value

    def close(self, stream):
        self.check(self.L.vtg_srv_close(self.h, stream), "srv_close")

    def submit(self, stream, frame, xy=None, guess=None, ncc=None):
        frame = np.ascontiguousarray(frame, np.uint8)
        n = 0 if xy is None else len(xy) // 2
        xyp = None if xy is None else np.ascontiguousarray(xy, np.float32)
        gp = None if guess is None else np.ascontiguousarray(guess, np.float32)
        x0, y0 = ncc if ncc is not None else (-1, -1)
        t = C.c_uint64()
        self.check(self.L.vtg_srv_submit(self.h, stream, frame.ctypes.data, frame.strides[0], n,
                                         None if xyp is None else xyp.ctypes.data,
                                         None if gp is None else gp.ctypes.data, x0, y0, C.byref(t)), "srv_submit")
        self.n[t.value] = n
        return t.value

    def _take(self, fn, ticket):
        n = self.n[ticket]
        out = np.zeros(n, LK_OUT_DTYPE)
        f = Found()
        lat = C.c_double()
        nrec = C.c_int()
        rc = fn(self.h, ticket, out.ctypes.data if n else None, C.byref(nrec), C.byref(f), C.byref(lat))
        if rc == EAGAIN:
            return None
        self.check(rc, "srv_take")
        del self.n[ticket]
        return out[:nrec.value], f, lat.value

    def poll(self, ticket):
        """(records, found, latency_ms) once done, else None."""
        return self._take(self.L.vtg_srv_poll, ticket)

    def wait(self, ticket):
        return self._take(self.L.vtg_srv_wait, ticket)

    def stats(self):
        st = SrvStats()
        self.check(self.L.vtg_srv_get_stats(self.h, C.byref(st)), "srv_stats")
        return {k: getattr(st, k) for k, _ in st._fields_}

    def destroy(self):
        if self.h:
            self.L.vtg_srv_destroy(self.h)
            self.h = None
