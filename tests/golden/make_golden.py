"""Writes the committed fixtures of tests/golden/ from the reference tree (run here, where /root/reference exists):
    m1_hashes.json   the reference-held hash lines logs/m1/host/m1-host-*.txt (data, parsed)
    m2_lines.json    logs/m2/host/640x480.txt: per displacement, point 0's raw bits and the FNV of all 330 lk_out
    ncc_ref.json     tk_correlate (src/kern/track.c:208) of oracle/_ref/libvtref.so on seeded 32 x 21 inputs
    lk_ref.json      lk_track_vx / lk_flowback_vx of oracle/_ref/libvtref.so on the 640 x 480 blob workload, and
                     lk_track_vx on the 1920 x 1080 one (two seeds)
    ncc_refine.json  store_appearance, search_half, peak_offset and search_slot (src/kern/track.c:935-974,
                     1615-1726) of oracle/_ref/libvtsearch.so on the planes of tests/ncc_cases.py
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


SEARCH_SO = os.path.join(REPO, "oracle", "_ref", "libvtsearch.so")


class RefFound(C.Structure):
    _fields_ = [("hx", C.c_int), ("hy", C.c_int), ("ncc", C.c_longdouble), ("second", C.c_longdouble)]


def _ld_hex(v):
    """A long double returned through ctypes (a Python float: the value rounded to double) as a hex string."""
    return float(v).hex()


def ncc_refine():
    import hashlib
    import ncc_cases as nc
    ref = C.CDLL(SEARCH_SO)
    P = C.POINTER(Plane)
    ref.ref_store_appearance.argtypes = [P, C.c_float, C.c_float, C.c_float, C.c_float, C.c_void_p,
                                         C.POINTER(C.c_float)]
    ref.ref_search_half.restype = C.c_longdouble
    ref.ref_search_half.argtypes = [P, C.c_void_p, C.c_float, C.c_int, C.c_int]
    ref.ref_peak_offset.restype = C.c_longdouble
    ref.ref_peak_offset.argtypes = [C.c_longdouble] * 3
    ref.ref_search_slot.argtypes = [P, C.c_int, C.c_void_p, C.c_float] + [C.c_int] * 5 + \
        [C.POINTER(RefFound), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
    o = Oracle()
    rng = np.random.default_rng(20261204)
    out = {"tw": nc.TW, "th": nc.TH, "region": list(nc.REGION), "cases": []}
    for name, img, (cx, cy) in nc.cases(o):
        h, w = img.shape
        hp = nc.host_plane(o, img)
        pix = np.zeros(nc.TW * nc.TH, np.float32)
        norm = C.c_float()
        assert ref.ref_store_appearance(C.byref(hp.c), cx, cy, *nc.REGION, pix.ctypes.data, C.byref(norm)) == 0
        case = {"name": name, "w": w, "h": h, "centre": [cx, cy],
                "template_sha256": hashlib.sha256(pix.tobytes()).hexdigest(),
                "norm": float(norm.value).hex()}
        x0, y0 = int(cx) - nc.TW // 2, int(cy) - nc.TH // 2
        corners = [(2 * x0 + dx, 2 * y0 + dy) for dy in range(-3, 4) for dx in range(-3, 4)]
        corners += [(int(rng.integers(-4, 2 * w)), int(rng.integers(-4, 2 * h))) for _ in range(60)]
        corners += [(1, 1), (2, 2), (2 * (w - nc.TW - 3), 2), (2 * (w - nc.TW - 3) + 1, 3),
                    (2 * (w - nc.TW - 2), 4), (4, 2 * (h - nc.TH - 3) + 1), (4, 2 * (h - nc.TH - 2))]
        case["half"] = [[hx, hy, _ld_hex(ref.ref_search_half(C.byref(hp.c), pix.ctypes.data, norm, hx, hy))]
                        for hx, hy in corners]
        cap = (w // 16 + 2) * (h // 16 + 2)
        sv, sx, sy = np.zeros(cap, np.float32), np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        case["searches"] = []
        for (a, b, c, d, step) in nc.windows(w, h, (x0, y0)):
            f, n = RefFound(), C.c_int()
            assert ref.ref_search_slot(C.byref(hp.c), 0, pix.ctypes.data, norm, a, b, c, d, step, C.byref(f),
                                       sv.ctypes.data, sx.ctypes.data, sy.ctypes.data, cap, C.byref(n)) == 0
            m = n.value
            cells = np.concatenate([sv[:m].view(np.uint32), sx[:m].view(np.uint32), sy[:m].view(np.uint32)])
            case["searches"].append({"window": [a, b, c, d, step], "hx": f.hx, "hy": f.hy, "ncc": _ld_hex(f.ncc),
                                     "second": _ld_hex(f.second), "cells": m,
                                     "cells_sha256": hashlib.sha256(cells.tobytes()).hexdigest()})
        out["cases"].append(case)
    triples = [(float(rng.normal()), float(rng.normal()), float(rng.normal())) for _ in range(150)]
    triples += [(0.2, 1.0, 0.6), (0.5, 1.0, 0.5), (1.0, 1.0, 1.0), (0.0, 1.0, 0.99), (0.99, 1.0, 0.0),
                (0.3, 0.5, 0.9), (0.9, 0.5, 0.3), (-1.0, 0.0, 1.0), (0.75, 1.0, 0.25), (0.25, 1.0, 0.75)]
    out["peak_offset"] = [[l.hex(), c.hex(), r.hex(), _ld_hex(ref.ref_peak_offset(l, c, r))] for l, c, r in triples]
    return out


if __name__ == "__main__":
    json.dump(m1(), open(os.path.join(HERE, "m1_hashes.json"), "w"), indent=1)
    json.dump(m2(), open(os.path.join(HERE, "m2_lines.json"), "w"), indent=1)
    if os.path.exists(REF_SO):
        json.dump(ncc_ref(), open(os.path.join(HERE, "ncc_ref.json"), "w"))
        json.dump(lk_ref(), open(os.path.join(HERE, "lk_ref.json"), "w"), indent=1)
    if os.path.exists(SEARCH_SO):
        json.dump(ncc_refine(), open(os.path.join(HERE, "ncc_refine.json"), "w"))
    print("fixtures written")
