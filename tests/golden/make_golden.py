"""Writes the committed fixtures of tests/golden/ from the reference tree (run here, where /root/reference exists):
    m1_hashes.json   the reference-held hash lines logs/m1/host/m1-host-*.txt (data, parsed)
    m2_lines.json    logs/m2/host/640x480.txt: per displacement, point 0's raw bits and the FNV of all 330 lk_out
    ncc_ref.json     tk_correlate (src/kern/track.c:208) of oracle/_ref/libvtref.so on seeded 32 x 21 inputs
    lk_ref.json      lk_track_vx / lk_flowback_vx of oracle/_ref/libvtref.so on the 640 x 480 blob workload, and
                     lk_track_vx on the 1920 x 1080 one (two seeds)
The reference itself never leaves this container: only these numbers are committed."""
import ctypes as C
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("VT_REFERENCE", "/root/reference")
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "intel-phi-visiontrack_amd"))
from oracle_ffi import HostPlane, LkCfg, Oracle, Plane, REF_SO  # noqa: E402
import vtgpu  # noqa: E402


def m1():
    out = {}
    for size in ("640x480", "1280x720", "1920x1080"):
        for line in open(os.path.join(REF, "logs/m1/host/m1-host-%s.txt" % size)):
            f = line.split()
            if f[1] == "minmax":
                out.setdefault(size, {})["minmax"] = [float(f[3]), float(f[4])]
            else:
                out.setdefault(size, {})[f[1]] = f[3]
    return out


def m2():
    cases = []
    pat = re.compile(r"lk_first (-?\d+),(-?\d+) x (\w+) y (\w+) eig (\w+) res (\w+) cxx (\w+) cxy (\w+) cyy (\w+) it (\d+) st (\d+)")
    lines = open(os.path.join(REF, "logs/m2/host/640x480.txt")).read().splitlines()
    for a, b in zip(lines[0::2], lines[1::2]):
        m = pat.match(a)
        dx, dy = int(m.group(1)), int(m.group(2))
        cases.append({"dx": dx, "dy": dy, "bits": [m.group(i) for i in range(3, 10)], "iters": int(m.group(10)),
                      "status": int(m.group(11)), "hash": b.split()[2]})
    return {"w": 640, "h": 480, "levels": 4,
            "cfg": {"w": 21, "h": 21, "max_iter": 10, "eps": 0.005, "min_eig": 0.1, "max_residual": 40.0},
            "grid": {"start": 61, "step": 24}, "offset": 2048, "cases": cases}


def ncc_ref():
    ref = C.CDLL(REF_SO)
    ref.tk_correlate.restype = C.c_longdouble
    ref.tk_correlate.argtypes = [C.POINTER(Plane), C.c_void_p, C.c_longdouble, C.c_int, C.c_int]
    o = Oracle()
    rng = np.random.default_rng(20261203)
    cases = []
    # a smooth textured plane, a template cut from it, scored at the cut, near it, far away and at the border
    w, h = 160, 96
    img = o.m2_render(w, h, 1000, 700)[:, :w]
    hp = HostPlane(w, h)
    hp.view()[:] = img.astype(np.float32) * 64
    o.lib.vo_pad_rows(C.byref(hp.c), 0, h)
    o.lib.vo_pad_tb(C.byref(hp.c))
    for k in range(6):
        x0, y0 = int(rng.integers(2, w - 36)), int(rng.integers(2, h - 25))
        app, norm = o.ncc_template(hp, x0, y0, 32, 21)
        corners = [(x0, y0)] + [(int(rng.integers(-2, w - 30)), int(rng.integers(-2, h - 19))) for _ in range(40)]
        vals = [float(ref.tk_correlate(C.byref(hp.c), app.ctypes.data, norm, x, y)) for x, y in corners]
        cases.append({"plane": "m2_render", "w": w, "h": h, "ox": 1000, "oy": 700, "x0": x0, "y0": y0,
                      "norm": float(norm), "corners": corners, "ncc": vals})
    return {"tw": 32, "th": 21, "cases": cases}


def _ref_lk_records(ref, o, w, h, levels, seed, nx, ny, win, fb):
    """lk_track_vx (and lk_flowback_vx when fb > 0) of the reference library on the seeded blob pair, every grid point."""
    f0, f1, mo = vtgpu.frame_pair(w, h, seed)
    keep_i, I = o.pyramid(f0, w, h, levels, threads=8)  # the planes' buffers must outlive the calls
    keep_j, J = o.pyramid(f1, w, h, levels, threads=8)
    xs, ys = vtgpu.grid(nx, ny, w, h)
    cfg = LkCfg(win, win, 10, 0.01, 1e-4, 40.0, 0)
    tile = np.zeros(ref.lk_tile_bytes() // 4 + 64, np.float32)
    tp = (tile.ctypes.data + 63) & ~63
    out = np.zeros(len(xs), vtgpu.LK_OUT_DTYPE)
    outfb = np.zeros(len(xs), vtgpu.LK_OUT_DTYPE)
    for i in range(len(xs)):
        ref.lk_track_vx(I, J, levels - 1, C.byref(cfg), xs[i], ys[i], 0.0, 0.0, out[i:i + 1].ctypes.data, tp)
        if fb:
            outfb[i] = out[i]
            ref.lk_flowback_vx(I, J, levels - 1, C.byref(cfg), xs[i], ys[i], fb, outfb[i:i + 1].ctypes.data, tp)
    return mo, out, outfb


def lk_ref():
    ref = C.CDLL(REF_SO)
    ref.lk_track_vx.argtypes = [C.POINTER(Plane), C.POINTER(Plane), C.c_int, C.POINTER(LkCfg), C.c_float, C.c_float,
                                C.c_float, C.c_float, C.c_void_p, C.c_void_p]
    ref.lk_flowback_vx.argtypes = [C.POINTER(Plane), C.POINTER(Plane), C.c_int, C.POINTER(LkCfg), C.c_float,
                                   C.c_float, C.c_float, C.c_void_p, C.c_void_p]
    ref.lk_tile_bytes.restype = C.c_size_t
    o = Oracle()
    w, h, levels, seed = 640, 480, 3, 1
    mo, out, outfb = _ref_lk_records(ref, o, w, h, levels, seed, 40, 30, 15, 0.5)
    bits = out.view(np.uint32).reshape(-1, 8)
    res = {"w": w, "h": h, "levels": levels, "seed": seed, "grid": [40, 30],
           "motion": [mo.dx, mo.dy, mo.theta_deg],
           "cfg": {"w": 15, "h": 15, "max_iter": 10, "eps": 0.01, "min_eig": 1e-4, "max_residual": 40.0},
           "fnv": "%016x" % o.lib.vo_hash_lk(out.ctypes.data, len(out)),
           "fnv_fb05": "%016x" % o.lib.vo_hash_lk(outfb.ctypes.data, len(outfb)),
           "first_bits": ["%08x" % b for b in bits[0]],
           "status_counts": {str(s): int((out["status"] == s).sum()) for s in np.unique(out["status"])},
           "status_counts_fb05": {str(s): int((outfb["status"] == s).sum()) for s in np.unique(outfb["status"])}}
    # the native workload: 1920 x 1080, 4 levels, 180 x 90 grid, 21 x 21, two seeds
    res["blob1080"] = []
    for seed in (2, 3):
        mo, out, _ = _ref_lk_records(ref, o, 1920, 1080, 4, seed, 180, 90, 21, 0.0)
        res["blob1080"].append({
            "w": 1920, "h": 1080, "levels": 4, "seed": seed, "grid": [180, 90], "motion": [mo.dx, mo.dy, mo.theta_deg],
            "cfg": {"w": 21, "h": 21, "max_iter": 10, "eps": 0.01, "min_eig": 1e-4, "max_residual": 40.0},
            "fnv": "%016x" % o.lib.vo_hash_lk(out.ctypes.data, len(out)),
            "status_counts": {str(s): int((out["status"] == s).sum()) for s in np.unique(out["status"])},
            "iters_sum": int(out["iters"].astype(np.int64).sum())})
    return res


if __name__ == "__main__":
    json.dump(m1(), open(os.path.join(HERE, "m1_hashes.json"), "w"), indent=1)
    json.dump(m2(), open(os.path.join(HERE, "m2_lines.json"), "w"), indent=1)
    if os.path.exists(REF_SO):
        json.dump(ncc_ref(), open(os.path.join(HERE, "ncc_ref.json"), "w"))
        json.dump(lk_ref(), open(os.path.join(HERE, "lk_ref.json"), "w"), indent=1)
    print("fixtures written")
