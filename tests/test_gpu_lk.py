"""GPU LK against the reference: every record (estimate, min eigenvalue, residual, covariance, iterations, status) equal
bit for bit to the reference's own lane-model LK (lk_track_vx / lk_flowback_vx, src/kern/lk_vx.c), pinned three ways:
the reference's committed hash lines (logs/m2/host/640x480.txt), fixtures written from the reference library
(tests/golden/lk_ref.json), and the oracle's lane tier on every other workload. The GPU's driver (the per-level and
per-iteration solve around the fp32 patch sums) runs in fp64 where the reference's runs in x87 long double; the
oracle's lane64 tier is that same patch arithmetic with an fp64 driver, so every record must equal lane64 bit for bit,
and lane64 itself may differ from the x87 lane tier only on the rare record where the two drivers round apart (held
here to at most 1 in 1,000 per workload). Each result's error against the FP64 oracle is reported beside it."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import vtgpu
from gpu_util import flow_diff, lane_ref as _lane_ref, say
from oracle_ffi import lk_cfg

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
LANE, F64, LANE64 = 0, 1, 2
FIELDS = ("x", "y", "min_eig", "residual", "cov_xx", "cov_xy", "cov_yy", "iters", "status")


def _gpu_pyr(ctx, img, w, h, levels, storage=vtgpu.STORE_Q6):
    p = ctx.pyramid(w, h, levels, storage)
    p.build(img)
    return p


def _fnv(oracle, out):
    out = np.ascontiguousarray(out)
    return "%016x" % oracle.lib.vo_hash_lk(out.ctypes.data, len(out))


def _assert_same(gpu, ref, label=""):
    """Every record equal bit for bit; on failure, which fields of how many points differ."""
    g = np.ascontiguousarray(gpu).view(np.uint32).reshape(-1, 8)
    r = np.ascontiguousarray(ref).view(np.uint32).reshape(-1, 8)
    bad = np.nonzero((g != r).any(1))[0]
    print(f"[lk] {label}: {len(ref)} records, {len(bad)} differ; status counts "
          f"{dict(zip(*np.unique(ref['status'], return_counts=True)))}")
    if len(bad):
        fields = {f: int((gpu[f][bad] != ref[f][bad]).sum()) for f in FIELDS}
        d = flow_diff(gpu, ref)
        raise AssertionError(f"{label}: {len(bad)} records differ (first {bad[:8]}), per field {fields}, "
                             f"status differs on {int((gpu['status'] != ref['status']).sum())}, max |d| {d.max():.3g}")


def _report_f64(got, f64, label):
    """Each result's error against the FP64 oracle over the points both converged (reported, and held to the 0.05 px
    band at the median)."""
    both = (got["status"] == 0) & (f64["status"] == 0)
    e = flow_diff(got, f64)
    print(f"[lk] {label} error vs FP64: {both.sum()} converged in both: median {np.median(e[both]):.2e} p99 "
          f"{np.percentile(e[both], 99):.2e} share <= 0.05 px {(e[both] <= 0.05).mean():.4f}; status equal "
          f"{(got['status'] == f64['status']).mean():.5f}")
    assert np.median(e[both]) <= 0.05


def _m2_setup():
    g = json.load(open(os.path.join(GOLD, "m2_lines.json")))
    c = g["cfg"]
    cfg = lk_cfg(c["w"], c["h"], c["max_iter"], c["eps"], c["min_eig"], c["max_residual"])
    m, step = g["grid"]["start"], g["grid"]["step"]
    pts = [(x, y) for y in range(m, g["h"] - m, step) for x in range(m, g["w"] - m, step)]
    xs = np.array([p[0] for p in pts], np.float32)
    ys = np.array([p[1] for p in pts], np.float32)
    return g, cfg, xs, ys


@pytest.mark.parametrize("storage", [vtgpu.STORE_Q6, vtgpu.STORE_F32, vtgpu.STORE_Q6F, vtgpu.STORE_F32_U8,
                                     vtgpu.STORE_Q6W])
def test_m2_reference_hash_lines(ctx, oracle, storage):
    """The reference's own m2 gate (m2_ref.c: 4 levels, 21 x 21, eps 0.005, min_eig 0.1, 330 points, ten exact
    displacements): the GPU's records hash to the committed lines of logs/m2/host/640x480.txt, and point 0's seven
    floats, iterations and status equal the logged bits, under every exact storage."""
    g, cfg, xs, ys = _m2_setup()
    w, h, L, off = g["w"], g["h"], g["levels"], g["offset"]
    i0 = oracle.m2_render(w, h, off, off)[:, :w]
    PI = _gpu_pyr(ctx, i0, w, h, L, storage)
    for case in g["cases"]:
        say("lk m2", case["dx"], case["dy"], "storage", storage)
        i1 = oracle.m2_render(w, h, off - case["dx"], off - case["dy"])[:, :w]
        PJ = _gpu_pyr(ctx, i1, w, h, L, storage)
        got = ctx.lk_track_host(PI, PJ, L - 1, cfg, xs, ys)
        bits = ["%08x" % b for b in got.view(np.uint32).reshape(-1, 8)[0][:7]]
        assert bits == case["bits"], (case, bits)
        assert int(got[0]["iters"]) == case["iters"] and int(got[0]["status"]) == case["status"]
        assert _fnv(oracle, got) == case["hash"], case
        PJ.destroy()
    PI.destroy()


@pytest.mark.parametrize("fb", [0.0, 0.5])
def test_blob_640_reference(ctx, oracle, fb):
    """The CPU-reference workload (640 x 480, 3 levels, 40 x 30 grid, 15 x 15, 10 iterations, eps 0.01, min_eig 1e-4),
    without and with the 0.5 px flowback: the GPU's 1,200 records hash to the reference library's
    (tests/golden/lk_ref.json, written from lk_track_vx / lk_flowback_vx) and equal the oracle's lane tier."""
    g = json.load(open(os.path.join(GOLD, "lk_ref.json")))
    w, h, L = g["w"], g["h"], g["levels"]
    f0, f1, mo = vtgpu.frame_pair(w, h, g["seed"])
    xs, ys = vtgpu.grid(*g["grid"], w, h)
    cfg = lk_cfg(15, 15, 10, 0.01, 1e-4, 40.0)
    kI, I = oracle.pyramid(f0, w, h, L)
    kJ, J = oracle.pyramid(f1, w, h, L)
    ref = _lane_ref(oracle, I, J, L - 1, cfg, xs, ys, fb_max=fb)
    f64 = oracle.lk_batch(F64, I, J, L - 1, cfg, xs, ys, fb_max=fb)
    say("lk blob 640 fb", fb)
    PI, PJ = _gpu_pyr(ctx, f0, w, h, L), _gpu_pyr(ctx, f1, w, h, L)
    got = ctx.lk_track_host(PI, PJ, L - 1, cfg, xs, ys, fb_max=fb)
    assert _fnv(oracle, got) == (g["fnv_fb05"] if fb else g["fnv"])
    _assert_same(got, ref, f"blob640 fb{fb}")
    _report_f64(got, f64, f"blob640 fb{fb}")
    PI.destroy()
    PJ.destroy()


@pytest.mark.parametrize("case", [0, 1])
def test_blob_1080_reference(ctx, oracle, case):
    """The native workload at full size (1920 x 1080, 4 levels, 180 x 90 grid = 16,200 points, 21 x 21): every record
    hashes to the reference library's lk_track_vx output (two seeds), statuses and all."""
    g = json.load(open(os.path.join(GOLD, "lk_ref.json")))["blob1080"][case]
    w, h, L = g["w"], g["h"], g["levels"]
    f0, f1, mo = vtgpu.frame_pair(w, h, g["seed"])
    xs, ys = vtgpu.grid(*g["grid"], w, h)
    c = g["cfg"]
    cfg = lk_cfg(c["w"], c["h"], c["max_iter"], c["eps"], c["min_eig"], c["max_residual"])
    say("lk blob 1080 seed", g["seed"])
    for storage in (vtgpu.STORE_Q6, vtgpu.STORE_F32_U8):
        PI, PJ = _gpu_pyr(ctx, f0, w, h, L, storage), _gpu_pyr(ctx, f1, w, h, L, storage)
        got = ctx.lk_track_host(PI, PJ, L - 1, cfg, xs, ys)
        counts = {str(s): int(n) for s, n in zip(*np.unique(got["status"], return_counts=True))}
        print(f"[lk] blob1080 seed {g['seed']} storage {storage}: status counts {counts} (reference "
              f"{g['status_counts']}), iterations {int(got['iters'].astype(np.int64).sum())} ({g['iters_sum']})")
        assert counts == g["status_counts"]
        assert int(got["iters"].astype(np.int64).sum()) == g["iters_sum"]
        assert _fnv(oracle, got) == g["fnv"]
        PI.destroy()
        PJ.destroy()
    if case == 0:
        # the analytic motion, for the points that converged, and the FP64 oracle beside it
        kI, I = oracle.pyramid(f0, w, h, L, threads=8)
        kJ, J = oracle.pyramid(f1, w, h, L, threads=8)
        f64 = oracle.lk_batch(F64, I, J, L - 1, cfg, xs, ys, threads=16)
        _report_f64(got, f64, "blob1080")
        tr = vtgpu.truth(mo, w, h, xs, ys)
        ok = got["status"] == 0
        epe = np.hypot(got["x"] - xs - tr[:, 0], got["y"] - ys - tr[:, 1])[ok]
        print(f"[lk] blob1080 EPE vs truth: median {np.median(epe):.3f} px over {ok.sum()} OK points")


def test_guess_and_plane_api(ctx, oracle):
    """A non-zero guess with the flowback, through the reference-shaped call over host planes (lk_track_vx's
    signature, src/kern/lk.h:45-53): every record equal to the lane tier's."""
    w, h, L = 640, 480, 3
    f0, f1, mo = vtgpu.frame_pair(w, h, 4)
    xs, ys = vtgpu.grid(20, 15, w, h)
    cfg = lk_cfg(21, 21, 10, 0.01, 1e-4, 40.0)
    kI, I = oracle.pyramid(f0, w, h, L)
    kJ, J = oracle.pyramid(f1, w, h, L)
    gx = np.full(len(xs), mo.dx * 0.5, np.float32)
    gy = np.full(len(xs), mo.dy * 0.5, np.float32)
    ref = _lane_ref(oracle, I, J, L - 1, cfg, xs, ys, gx, gy, fb_max=0.5)
    say("lk plane api")
    got = np.zeros(len(xs), vtgpu.LK_OUT_DTYPE)
    Lb = vtgpu.lib()
    Iv = (vtgpu.Plane * L)(*[vtgpu.Plane(p.p, p.w, p.h, p.stride) for p in I])
    Jv = (vtgpu.Plane * L)(*[vtgpu.Plane(p.p, p.w, p.h, p.stride) for p in J])
    ctx.check(Lb.vtg_lk_track_planes(ctx.h, Iv, Jv, L - 1, C.byref(vtgpu.LkCfg(21, 21, 10, 0.01, 1e-4, 40.0, 0)),
                                     len(xs), xs.ctypes.data, ys.ctypes.data, gx.ctypes.data, gy.ctypes.data, 0.5,
                                     got.ctypes.data))
    _assert_same(got, ref, "plane api + guess + fb")


@pytest.mark.parametrize("win", [15, 21])
def test_edges_and_flat(ctx, oracle, win):
    """Points at and beyond the border (LK_OUT); points whose windows converge just inside the fit limits at level 0
    and above (a zero-motion pair, so the estimate stays put: status 0 right at the edge); a flat frame (LK_LOWEIG);
    an empty batch and bad arguments. Every record equal to the lane tier's."""
    w, h, L = 320, 240, 3
    f0, f1, mo = vtgpu.frame_pair(w, h, 6)
    m = win // 2 + 1
    xs = np.array([0, 5, 10.5, 11, 11.01, 308.9, 309, 319, 160, 160, -3, 400], np.float32)
    ys = np.array([0, 5, 10.5, 11, 11.01, 228.9, 228, 239, 3, 236, 120, 120], np.float32)
    cfg = lk_cfg(win, win, 10, 0.01, 1e-4, 40.0)
    kI, I = oracle.pyramid(f0, w, h, L)
    kJ, J = oracle.pyramid(f1, w, h, L)
    ref = _lane_ref(oracle, I, J, L - 1, cfg, xs, ys)
    say("lk edges", win)
    PI, PJ = _gpu_pyr(ctx, f0, w, h, L), _gpu_pyr(ctx, f1, w, h, L)
    got = ctx.lk_track_host(PI, PJ, L - 1, cfg, xs, ys)
    _assert_same(got, ref, f"edges {win}")
    # the fit limits, p >= m and p < dim - m - 1 at level 0, on a textured frame (m2_render): the frame against itself
    # (zero motion: status 0 right at the limits, LK_OUT just past them) and against copies moved one pixel either way
    t0 = oracle.m2_render(w, h, 1000, 700)[:, :w]
    kT, T = oracle.pyramid(t0, w, h, L)
    PT = _gpu_pyr(ctx, t0, w, h, L)
    ex = np.array([m, m + 0.25, w - m - 1.01, w - m - 1.5, 2 * m + 0.5, w - 2 * m - 2.5, 100.0, 150.0, m - 0.01,
                   w - m - 1], np.float32)
    ey = np.array([120.0, 80.0, 100.0, 60.0, m + 0.1, h - m - 1.2, h - 2 * m - 2.6, 2 * m + 0.3, 100.0, 100.0],
                  np.float32)
    ref = _lane_ref(oracle, T, T, L - 1, cfg, ex, ey)
    got = ctx.lk_track_host(PT, PT, L - 1, cfg, ex, ey)
    assert ref["status"].tolist() == [0] * 8 + [vtgpu.LK_OUT] * 2
    _assert_same(got, ref, f"edge convergence {win}")
    for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
        t1 = oracle.m2_render(w, h, 1000 - dx, 700 - dy)[:, :w]
        k1, T1 = oracle.pyramid(t1, w, h, L)
        P1 = _gpu_pyr(ctx, t1, w, h, L)
        ref = _lane_ref(oracle, T, T1, L - 1, cfg, ex, ey, fb_max=0.5)
        got = ctx.lk_track_host(PT, P1, L - 1, cfg, ex, ey, fb_max=0.5)
        _assert_same(got, ref, f"edge shift {dx},{dy} win {win}")
        P1.destroy()
    PT.destroy()
    flat = np.full((h, w), 90, np.uint8)
    kF, F = oracle.pyramid(flat, w, h, L)
    PF = _gpu_pyr(ctx, flat, w, h, L)
    fx, fy = vtgpu.grid(8, 6, w, h)
    got = ctx.lk_track_host(PF, PF, L - 1, cfg, fx, fy)
    ref = _lane_ref(oracle, F, F, L - 1, cfg, fx, fy)
    _assert_same(got, ref, f"flat {win}")
    assert (got["status"] & vtgpu.LK_LOWEIG).all()
    empty = ctx.alloc(8)
    out = ctx.alloc(32)
    ctx.check(vtgpu.lib().vtg_lk_track(ctx.h, PI.h, PJ.h, L - 1, C.byref(cfg), 0, empty.ptr, None, 0.0, out.ptr))
    assert vtgpu.lib().vtg_lk_track(ctx.h, PI.h, PJ.h, L, C.byref(cfg), 1, empty.ptr, None, 0.0, out.ptr) == -22
    bad = lk_cfg(49, 21)
    assert vtgpu.lib().vtg_lk_track(ctx.h, PI.h, PJ.h, 1, C.byref(bad), 1, empty.ptr, None, 0.0, out.ptr) == -22
    for b in (empty, out):
        b.free()
    for p in (PI, PJ, PF):
        p.destroy()


@pytest.mark.parametrize("win", [(9, 7), (17, 13), (33, 25), (48, 31)])
def test_other_windows(ctx, oracle, win):
    """Windows other than the compiled exact shapes (one, two and three 16-column blocks, up to the reference's
    48 x 31 limit): every record equal to the lane tier's."""
    w, h, L = 640, 480, 3
    f0, f1, mo = vtgpu.frame_pair(w, h, 9)
    xs, ys = vtgpu.grid(24, 18, w, h)
    cfg = lk_cfg(win[0], win[1], 10, 0.01, 1e-4, 40.0)
    kI, I = oracle.pyramid(f0, w, h, L)
    kJ, J = oracle.pyramid(f1, w, h, L)
    ref = _lane_ref(oracle, I, J, L - 1, cfg, xs, ys, fb_max=0.5)
    say("lk window", win)
    PI, PJ = _gpu_pyr(ctx, f0, w, h, L), _gpu_pyr(ctx, f1, w, h, L)
    got = ctx.lk_track_host(PI, PJ, L - 1, cfg, xs, ys, fb_max=0.5)
    _assert_same(got, ref, f"window {win}")
    PI.destroy()
    PJ.destroy()


@pytest.mark.parametrize("storage,solve,tol", [(vtgpu.STORE_Q6, vtgpu.SOLVE_MIXED, 0.05),
                                               (vtgpu.STORE_F16, vtgpu.SOLVE_F64, 0.05),
                                               (vtgpu.STORE_F16, vtgpu.SOLVE_MIXED, 0.05),
                                               (vtgpu.STORE_BF16, vtgpu.SOLVE_F64, None)])
def test_precision_variants_vs_f64(ctx, oracle, storage, solve, tol):
    """fp16 / bf16 storage and the fp32 solve with fp64 fallback (not the reference's arithmetic): flow error against
    the FP64 oracle, reported per variant; fp16 and the mixed solve are held to 0.05 px at the median and their
    statuses to 99 % agreement (bf16 is reported only)."""
    w, h, L = 1920, 1080, 4
    f0, f1, mo = vtgpu.frame_pair(w, h, 2)
    xs, ys = vtgpu.grid(90, 45, w, h)
    cfg = lk_cfg(21, 21, 10, 0.01, 1e-4, 40.0)
    kI, I = oracle.pyramid(f0, w, h, L, threads=8)
    kJ, J = oracle.pyramid(f1, w, h, L, threads=8)
    f64 = oracle.lk_batch(F64, I, J, L - 1, cfg, xs, ys, threads=16)
    say("lk precision", storage, solve)
    PI, PJ = _gpu_pyr(ctx, f0, w, h, L, storage), _gpu_pyr(ctx, f1, w, h, L, storage)
    got = ctx.lk_track_host(PI, PJ, L - 1, cfg, xs, ys, solve=solve)
    agree = (got["status"] == f64["status"]).mean()
    both = (got["status"] == 0) & (f64["status"] == 0)
    e = flow_diff(got, f64)[both]
    print(f"[lk] storage {storage} solve {solve}: status agreement {agree:.4f}; error vs FP64 median {np.median(e):.2e}"
          f" p99 {np.percentile(e, 99):.2e}")
    if tol is not None:
        assert agree >= 0.99 and np.median(e) <= tol
    PI.destroy()
    PJ.destroy()


def test_schedules_bit_identical_and_batch(ctx, oracle):
    """Every LK schedule (3 or 4 points per wavefront, 1, 2, 4, 8 wavefronts per workgroup) gives the same bytes, and one
    batched launch over pairs of uneven sizes (point counts not multiples of 3 or 4) gives the same bytes as one launch
    per pair and as the lane tier."""
    sizes = [(640, 480, 3, 401), (1280, 720, 4, 902), (333, 250, 2, 151), (1920, 1080, 4, 1203)]
    cfg = lk_cfg(21, 21, 10, 0.01, 1e-4, 40.0)
    pairs, xys, offs, frames = [], [], [0], []
    for i, (w, h, L, n) in enumerate(sizes):
        f0, f1, _ = vtgpu.frame_pair(w, h, 70 + i)
        PI, PJ = _gpu_pyr(ctx, f0, w, h, L), _gpu_pyr(ctx, f1, w, h, L)
        xs, ys = vtgpu.points(70 + i, n, w, h, 8)
        pairs.append((PI, PJ, L - 1))
        frames.append((f0, f1, w, h, L))
        xys.append(np.stack([xs, ys], 1).ravel())
        offs.append(offs[-1] + n)
    xy = np.concatenate(xys).astype(np.float32)
    dxy = ctx.to_device(xy)
    Lb = vtgpu.lib()
    results = {}
    for ppw in (3, 4):
        ctx.check(Lb.vtg_lk_set_ppw(ctx.h, ppw))
        for waves in (1, 2, 4, 8):
            say("lk schedule ppw", ppw, "waves", waves)
            ctx.check(Lb.vtg_lk_set_waves(ctx.h, waves))
            out = ctx.lk_track_many([p[0] for p in pairs], [p[1] for p in pairs], [p[2] for p in pairs], offs, cfg,
                                    dxy, fb_max=0.5)
            results[ppw, waves] = out.get(np.zeros(offs[-1], vtgpu.LK_OUT_DTYPE))
            out.free()
    # the points in the caller's order instead of spatial order
    ctx.check(Lb.vtg_lk_set_order(ctx.h, 0))
    out = ctx.lk_track_many([p[0] for p in pairs], [p[1] for p in pairs], [p[2] for p in pairs], offs, cfg, dxy,
                            fb_max=0.5)
    results["unordered"] = out.get(np.zeros(offs[-1], vtgpu.LK_OUT_DTYPE))
    out.free()
    ctx.check(Lb.vtg_lk_set_order(ctx.h, 1))
    for key in results:
        assert results[key].tobytes() == results[3, 4].tobytes(), key
    assert Lb.vtg_lk_set_waves(ctx.h, 9) == -22
    assert Lb.vtg_lk_set_ppw(ctx.h, 5) == -22
    assert Lb.vtg_lk_set_order(ctx.h, 2) == -22
    ctx.check(Lb.vtg_lk_set_waves(ctx.h, 1))
    ctx.check(Lb.vtg_lk_set_ppw(ctx.h, 0))
    for k, (PI, PJ, top) in enumerate(pairs):
        one = ctx.lk_track_host(PI, PJ, top, cfg, xys[k][0::2], xys[k][1::2], fb_max=0.5)
        assert one.tobytes() == results[3, 4][offs[k]:offs[k + 1]].tobytes(), k
        f0, f1, w, h, L = frames[k]
        kI, I = oracle.pyramid(f0, w, h, L, threads=8)
        kJ, J = oracle.pyramid(f1, w, h, L, threads=8)
        ref = _lane_ref(oracle, I, J, top, cfg, xys[k][0::2], xys[k][1::2], fb_max=0.5)
        _assert_same(one, ref, f"batch pair {k}")
    dxy.free()
    for PI, PJ, _ in pairs:
        PI.destroy()
        PJ.destroy()
