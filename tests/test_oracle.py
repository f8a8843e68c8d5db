"""The CPU oracle pinned: against the reference's own committed hash lines, the reference built from its sources
(oracle/_ref, where present), closed forms and brute force. No GPU."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from oracle_ffi import CELL_DTYPE, HostPlane, LK_OUT_DTYPE, REF_SO, lk_cfg

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _l0(o, img, w, h):
    a = HostPlane(w, h)
    o.lib.vo_u8_to_q6(C.byref(a.c), img.ctypes.data, img.strides[0], 0, h)
    o.lib.vo_pad_rows(C.byref(a.c), 0, h)
    o.lib.vo_pad_tb(C.byref(a.c))
    return a


@pytest.mark.parametrize("size", ["640x480", "1280x720", "1920x1080"])
def test_m1_hash_lines(oracle, size):
    """Every image primitive against logs/m1/host/m1-host-<size>.txt (pure-integer m1 input)."""
    gold = json.load(open(os.path.join(GOLD, "m1_hashes.json")))[size]
    o = oracle
    w, h = map(int, size.split("x"))
    img = o.m1_image(w, h)
    a = _l0(o, img, w, h)
    assert "%016x" % o.hash_plane(a) == gold["u8_to_q6"]
    c = HostPlane(w, h)
    go = np.array([0.75, 512.0], np.float32)
    o.lib.vo_norm(C.byref(c.c), C.byref(a.c), go.ctypes.data, 0, h)
    assert "%016x" % o.hash_plane(c) == gold["norm"]
    mm = np.zeros(2, np.float32)
    o.lib.vo_minmax(C.byref(a.c), mm.ctypes.data, 0, h)
    assert list(mm) == gold["minmax"]
    o.lib.vo_gauss5(C.byref(c.c), C.byref(a.c), 0, h)
    assert "%016x" % o.hash_plane(c) == gold["gauss5"]
    r = HostPlane((w + 1) // 2, (h + 1) // 2)
    o.lib.vo_reduce(C.byref(r.c), C.byref(a.c), 0, r.h)
    assert "%016x" % o.hash_plane(r) == gold["reduce"]
    d = HostPlane(w, h)
    o.lib.vo_sobel(C.byref(c.c), C.byref(d.c), C.byref(a.c), 0, h)
    assert "%016x" % o.hash_plane(c) == gold["sobel_x"]
    assert "%016x" % o.hash_plane(d) == gold["sobel_y"]
    b = HostPlane(w, h)
    o.lib.vo_gauss5(C.byref(b.c), C.byref(a.c), 0, h)
    o.lib.vo_absdiff(C.byref(c.c), C.byref(a.c), C.byref(b.c), 0, h)
    assert "%016x" % o.hash_plane(c) == gold["absdiff"]


def _reduce_numpy(q6):
    """Textbook reduce: replicate-pad by 2, separable (1 4 6 4 1) in int64, even samples, (s + 128) >> 8."""
    k = np.array([1, 4, 6, 4, 1], np.int64)
    p = np.pad(q6.astype(np.int64), 2, mode="edge")
    h, w = q6.shape
    v = sum(k[j] * p[j:j + h, :] for j in range(5))
    s = sum(k[i] * v[:, i:i + w] for i in range(5))
    return ((s[::2, ::2] + 128) >> 8).astype(np.float32)


@pytest.mark.parametrize("w,h", [(8, 8), (9, 7), (33, 17), (64, 48)])
def test_reduce_brute_force(oracle, w, h):
    """Odd sizes and borders: the oracle's reduce equals an independent int64 numpy restatement."""
    rng = np.random.default_rng(w * 100 + h)
    img = rng.integers(0, 256, (h, w), dtype=np.uint8)
    a = _l0(oracle, img, w, h)
    r = HostPlane((w + 1) // 2, (h + 1) // 2)
    oracle.lib.vo_reduce(C.byref(r.c), C.byref(a.c), 0, r.h)
    assert np.array_equal(r.view(), _reduce_numpy(img.astype(np.int64) * 64))


def test_reduce_constant_plane(oracle):
    """A constant plane stays the same constant at every level (the kernel sums to 256)."""
    img = np.full((50, 70), 173, np.uint8)
    planes, _ = oracle.pyramid(img, 70, 50, 4)
    for p in planes:
        assert np.all(p.view() == 173 * 64)


def _m2_setup(oracle):
    g = json.load(open(os.path.join(GOLD, "m2_lines.json")))
    w, h = g["w"], g["h"]
    c = g["cfg"]
    cfg = lk_cfg(c["w"], c["h"], c["max_iter"], c["eps"], c["min_eig"], c["max_residual"])
    m = g["grid"]["start"]
    pts = [(x, y) for y in range(m, h - m, g["grid"]["step"]) for x in range(m, w - m, g["grid"]["step"])]
    xs = np.array([p[0] for p in pts], np.float32)
    ys = np.array([p[1] for p in pts], np.float32)
    return g, w, h, cfg, xs, ys


def test_m2_lk_lines(oracle):
    """The lane tier reproduces logs/m2/host/640x480.txt bit for bit: point 0's 7 floats, iters, status, and the
    FNV of all 330 records, for all ten displacements (a 4-level pyramid, 21 x 21)."""
    g, w, h, cfg, xs, ys = _m2_setup(oracle)
    off = g["offset"]
    keep0, I = oracle.pyramid(oracle.m2_render(w, h, off, off), w, h, g["levels"])
    for case in g["cases"]:
        keep1, J = oracle.pyramid(oracle.m2_render(w, h, off - case["dx"], off - case["dy"]), w, h, g["levels"])
        out = oracle.lk_batch(0, I, J, g["levels"] - 1, cfg, xs, ys)
        bits = ["%08x" % b for b in out.view(np.uint32).reshape(-1, 8)[0][:7]]
        assert bits == case["bits"], case
        assert int(out[0]["iters"]) == case["iters"] and int(out[0]["status"]) == case["status"]
        assert "%016x" % oracle.lib.vo_hash_lk(out.ctypes.data, len(out)) == case["hash"], case


def test_m2_f64_accuracy(oracle):
    """The FP64 tier on the m2 inputs, whose displacement is exact: EPE within the reference's own measured band
    (median <= 0.0115, p95 <= 0.0209 px for |d| <= 14.4 px, docs/measurements.md:93-103)."""
    g, w, h, cfg, xs, ys = _m2_setup(oracle)
    off = g["offset"]
    keep0, I = oracle.pyramid(oracle.m2_render(w, h, off, off), w, h, g["levels"])
    for case in g["cases"][:9]:
        keep1, J = oracle.pyramid(oracle.m2_render(w, h, off - case["dx"], off - case["dy"]), w, h, g["levels"])
        out = oracle.lk_batch(1, I, J, g["levels"] - 1, cfg, xs, ys)
        ok = out["status"] == 0
        epe = np.hypot(out["x"] - xs - case["dx"] / 4, out["y"] - ys - case["dy"] / 4)[ok]
        assert ok.mean() >= 0.9
        assert np.median(epe) <= 0.0115 and np.percentile(epe, 95) <= 0.0209, (case, np.median(epe))
        lane = oracle.lk_batch(0, I, J, g["levels"] - 1, cfg, xs, ys)
        both = ok & (lane["status"] == 0)
        assert np.max(np.abs(lane["x"] - out["x"])[both]) < 1e-3


def _eig_closed_form(px, py, cx, cy, half, m2):
    """G's smaller eigenvalue per window pixel for the plane f = X Y^2 + 3 X^2 Y (X = x - cx, Y = y - cy) at integer
    (px, py): a 3 x 3 operator with [-1 0 1] / 2 along and smoothing weights (sum 1, symmetric, second moment m2)
    across gives exactly d/dx = Y^2 + m2 + 6XY and d/dy = 2XY + 3 (X^2 + m2)."""
    X = np.arange(px - half, px + half + 1, dtype=np.float64)[None, :] - cx
    Y = np.arange(py - half, py + half + 1, dtype=np.float64)[:, None] - cy
    dx = Y * Y + m2 + 6 * X * Y
    dy = 2 * X * Y + 3 * (X * X + m2)
    a, b, c = (dx * dx).sum(), (dx * dy).sum(), (dy * dy).sum()
    return (a + c - np.sqrt((a - c) ** 2 + 4 * b * b)) / 2 / dx.size


def test_gradient_operators_closed_form(oracle):
    """Both gradient operators of every LK tier against the closed form on a polynomial plane: Sobel's smoothing
    [1 2 1] / 4 has second moment 1/2, Scharr's [3 10 3] / 16 has 3/8; a wrong weight, scale or axis moves the
    eigenvalue far outside the tolerance (checked: the two operators' closed forms differ by far more)."""
    n, cx, cy = 64, 32.0, 32.0
    hp = HostPlane(n, n)
    x = np.arange(n, dtype=np.float64)[None, :] - cx
    y = np.arange(n, dtype=np.float64)[:, None] - cy
    hp.view()[:] = (64.0 * (x * y * y + 3 * x * x * y)).astype(np.float32)  # Q6, exact integers below 2^24
    arr = (type(hp.c) * 1)(hp.c)
    cfg = lk_cfg(21, 21, 10, 0.01, 1e-4, 1e9)
    pts = [(30, 34), (35, 29), (32, 32), (27, 38)]
    xs = np.array([p[0] for p in pts], np.float32)
    ys = np.array([p[1] for p in pts], np.float32)
    for tier, m2, rtol in ((0, 0.5, 2e-5), (1, 0.5, 1e-6), (2, 0.5, 2e-5), (3, 0.375, 2e-5), (4, 0.375, 1e-6)):
        out = oracle.lk_batch(tier, arr, arr, 0, cfg, xs, ys, threads=1)
        for i, (px, py) in enumerate(pts):
            want = _eig_closed_form(px, py, cx, cy, 10, m2)
            other = _eig_closed_form(px, py, cx, cy, 10, 0.875 - m2)
            assert abs(other - want) > 20 * rtol * want, (px, py)
            assert abs(float(out["min_eig"][i]) - want) <= rtol * want, (tier, px, py, out["min_eig"][i], want)


def test_m2_scharr_accuracy(oracle):
    """The Scharr variant's tiers (no reference counterpart) on the m2 inputs, whose displacement is exact: the same
    EPE band as the Sobel tiers within 5 % (median <= 0.012, p95 <= 0.022 px; measured 0.0114 and 0.02094 at the
    worst displacement, Sobel's 0.0114 and 0.0209), the fp32 lane arithmetic within 1e-3 px of fp64, and different
    answers from Sobel."""
    g, w, h, cfg, xs, ys = _m2_setup(oracle)
    off = g["offset"]
    keep0, I = oracle.pyramid(oracle.m2_render(w, h, off, off), w, h, g["levels"])
    moved = 0
    for case in g["cases"][1:9]:
        keep1, J = oracle.pyramid(oracle.m2_render(w, h, off - case["dx"], off - case["dy"]), w, h, g["levels"])
        f64 = oracle.lk_batch(4, I, J, g["levels"] - 1, cfg, xs, ys)
        lane = oracle.lk_batch(3, I, J, g["levels"] - 1, cfg, xs, ys)
        sobel = oracle.lk_batch(2, I, J, g["levels"] - 1, cfg, xs, ys)
        ok = f64["status"] == 0
        epe = np.hypot(f64["x"] - xs - case["dx"] / 4, f64["y"] - ys - case["dy"] / 4)[ok]
        assert ok.mean() >= 0.9
        assert np.median(epe) <= 0.012 and np.percentile(epe, 95) <= 0.022, (case, np.median(epe))
        both = ok & (lane["status"] == 0)
        assert np.max(np.abs(lane["x"] - f64["x"])[both]) < 1e-3
        moved += int((lane["x"] != sobel["x"]).sum())
    assert moved > 0


def test_template_rounding_term(oracle):
    """The identity the GPU's fast NCC maps score with takes the stored template t_k = fl(64 u_k - mean) as exactly
    64 u_k - mean. On a u8 level with tw th a power of two it is exact: every e_k = t_k - (64 u_k - mean) is 0 (the
    mean has at most 6 + log2(tw th) fractional bits), on uniform, near-black and 4-grey templates. Otherwise it need not be,
    and the bound 4 * 64 * 255 * sum |e| / norm on what leaving it out moves a score (the GPU flags templates above 1e-6
    for the exact kernel) is below 1e-7 on a uniform texture and 1.8e-5 on a near-black template of 0s and 1s."""
    w, h = 200, 150
    rng = np.random.default_rng(5)

    def bound(img, tw, th):
        a = _l0(oracle, img, w, h)
        app, norm = oracle.ncc_template(a, 10, 10, tw, th)
        app = np.asarray(app, np.float32).reshape(th, tw)
        u = img[10:10 + th, 10:10 + tw].astype(np.float64)
        mean = np.float32(64.0 * u.sum() / (tw * th))
        assert np.array_equal(app, np.float32(64.0) * u.astype(np.float32) - mean)  # the oracle's own rounding
        e = app.astype(np.float64) - (64.0 * u - np.float64(mean))
        return np.abs(e).sum(), 4 * 64 * 255 * np.abs(e).sum() / float(norm)

    images = {"uniform": rng.integers(0, 256, (h, w)), "dark01": (rng.random((h, w)) < 0.15),
              "grey0-3": rng.integers(0, 4, (h, w))}
    images = {k: np.ascontiguousarray(v.astype(np.uint8)) for k, v in images.items()}
    for kind, img in images.items():
        for tw, th in ((64, 64), (32, 32), (16, 16), (64, 32), (32, 16), (16, 64)):
            assert bound(img, tw, th)[0] == 0.0, (kind, tw, th)
    assert bound(images["uniform"], 64, 40)[1] < 1e-7
    for tw, th in ((64, 48), (32, 24), (64, 21)):
        es, b = bound(images["dark01"], tw, th)
        assert es > 0 and 1e-5 < b < 3e-5, (tw, th, b)


def test_lk_reference_blob_workload(oracle):
    """The lane tier equals oracle/_ref's lk_track_vx on the 640 x 480 blob workload (fixture written from it)."""
    import vtgpu
    g = json.load(open(os.path.join(GOLD, "lk_ref.json")))
    w, h, L = g["w"], g["h"], g["levels"]
    f0, f1, mo = vtgpu.frame_pair(w, h, g["seed"])
    assert np.allclose([mo.dx, mo.dy, mo.theta_deg], g["motion"], rtol=0, atol=1e-12)
    k0, I = oracle.pyramid(f0, w, h, L)
    k1, J = oracle.pyramid(f1, w, h, L)
    xs, ys = vtgpu.grid(*g["grid"], w, h)
    c = g["cfg"]
    cfg = lk_cfg(c["w"], c["h"], c["max_iter"], c["eps"], c["min_eig"], c["max_residual"])
    out = oracle.lk_batch(0, I, J, L - 1, cfg, xs, ys)
    assert "%016x" % oracle.lib.vo_hash_lk(out.ctypes.data, len(out)) == g["fnv"]
    out = oracle.lk_batch(0, I, J, L - 1, cfg, xs, ys, fb_max=0.5)
    assert "%016x" % oracle.lib.vo_hash_lk(out.ctypes.data, len(out)) == g["fnv_fb05"]


@pytest.mark.parametrize("case", [0, 1])
def test_lk_reference_blob_1080(oracle, case):
    """The lane tier equals oracle/_ref's lk_track_vx on every record of the 1920 x 1080 blob workload (4 levels,
    180 x 90 grid, 21 x 21; fixture written from the reference library)."""
    import vtgpu
    g = json.load(open(os.path.join(GOLD, "lk_ref.json")))["blob1080"][case]
    w, h, L = g["w"], g["h"], g["levels"]
    f0, f1, mo = vtgpu.frame_pair(w, h, g["seed"])
    assert np.allclose([mo.dx, mo.dy, mo.theta_deg], g["motion"], rtol=0, atol=1e-12)
    k0, I = oracle.pyramid(f0, w, h, L, threads=8)
    k1, J = oracle.pyramid(f1, w, h, L, threads=8)
    xs, ys = vtgpu.grid(*g["grid"], w, h)
    c = g["cfg"]
    cfg = lk_cfg(c["w"], c["h"], c["max_iter"], c["eps"], c["min_eig"], c["max_residual"])
    out = oracle.lk_batch(0, I, J, L - 1, cfg, xs, ys, threads=8)
    assert "%016x" % oracle.lib.vo_hash_lk(out.ctypes.data, len(out)) == g["fnv"]
    assert int(out["iters"].astype(np.int64).sum()) == g["iters_sum"]


def _tiers_apart(oracle, I, J, top, cfg, xs, ys, fb=0.0, threads=8):
    """lane64 (fp64 driver) against the x87 lane tier: how many records differ, and the worst estimate gap among them
    where both converged."""
    a = oracle.lk_batch(2, I, J, top, cfg, xs, ys, fb_max=fb, threads=threads)
    b = oracle.lk_batch(0, I, J, top, cfg, xs, ys, fb_max=fb, threads=threads)
    diff = (a.view(np.uint32).reshape(-1, 8) != b.view(np.uint32).reshape(-1, 8)).any(1)
    both = diff & (a["status"] == 0) & (b["status"] == 0)
    gap = float(np.max(np.abs(np.concatenate([a["x"][both] - b["x"][both], a["y"][both] - b["y"][both]])),
                       initial=0.0))
    return int(diff.sum()), gap, len(a)


def test_lane64_tier(oracle):
    """The lane64 tier (the reference's fp32 lane sums and x87 covariance under an fp64 driver, the arithmetic the GPU
    implements) equals the x87 lane tier, itself pinned to the reference above, on every record of the m2 hash-line
    workload and the 640 x 480 blob workload with and without the flowback, and on all but at most 1 in 1,000 records
    of the 1920 x 1080 blob workload, where a differing record's estimate moves by less than 1e-3 px."""
    import vtgpu
    g, w, h, cfg, xs, ys = _m2_setup(oracle)
    off = g["offset"]
    keep0, I = oracle.pyramid(oracle.m2_render(w, h, off, off), w, h, g["levels"])
    for case in g["cases"]:
        keep1, J = oracle.pyramid(oracle.m2_render(w, h, off - case["dx"], off - case["dy"]), w, h, g["levels"])
        assert _tiers_apart(oracle, I, J, g["levels"] - 1, cfg, xs, ys)[0] == 0, case
    gb = json.load(open(os.path.join(GOLD, "lk_ref.json")))
    for gg in (gb, gb["blob1080"][0]):
        c = gg["cfg"]
        cfg = lk_cfg(c["w"], c["h"], c["max_iter"], c["eps"], c["min_eig"], c["max_residual"])
        w, h, L = gg["w"], gg["h"], gg["levels"]
        f0, f1, _ = vtgpu.frame_pair(w, h, gg["seed"])
        k0, I = oracle.pyramid(f0, w, h, L, threads=8)
        k1, J = oracle.pyramid(f1, w, h, L, threads=8)
        xs, ys = vtgpu.grid(*gg["grid"], w, h)
        for fb in ((0.0, 0.5) if gg is gb else (0.0,)):
            nd, gap, n = _tiers_apart(oracle, I, J, L - 1, cfg, xs, ys, fb)
            print(f"lane64 vs lane: {w}x{h} fb {fb}: {nd} of {n} records differ, worst gap {gap:.2e} px")
            assert (nd == 0) if gg is gb else (nd * 1000 <= n)
            assert gap < 1e-3


def _plane_from(img):
    h, w = img.shape
    hp = HostPlane(w, h)
    hp.view()[:] = img.astype(np.float32) * 64
    return hp


def _pad(o, hp):
    o.lib.vo_pad_rows(C.byref(hp.c), 0, hp.h)
    o.lib.vo_pad_tb(C.byref(hp.c))


def test_ncc_reference_vectors(oracle):
    """vo_ncc_at at 32 x 21 equals the reference's tk_correlate on the committed vectors (long double bits)."""
    g = json.load(open(os.path.join(GOLD, "ncc_ref.json")))
    for case in g["cases"]:
        img = oracle.m2_render(case["w"], case["h"], case["ox"], case["oy"])[:, :case["w"]]
        hp = _plane_from(img)
        _pad(oracle, hp)
        app, norm = oracle.ncc_template(hp, case["x0"], case["y0"], 32, 21)
        assert norm == np.float32(case["norm"])
        got = [oracle.ncc_at(hp, app, 32, 21, norm, x, y) for x, y in case["corners"]]
        assert got == case["ncc"]


def test_ncc_refine_reference(oracle):
    """The template, the half-pixel scores, the parabola and the whole search pinned to the reference's own functions
    (store_appearance, search_half, peak_offset and search_slot of src/kern/track.c, compiled from source by
    oracle/ref_track.c; fixture tests/golden/ncc_refine.json): vo_ncc_template's bits and norm, vo_ncc_half at 74
    corners a case (the 7 x 7 half-pixel neighbourhood of the template's own corner, random corners and both edges),
    vo_peak_offset on 160 triples, and vo_ncc_search's pick, score, second peak and whole cell map over 10 windows (whole
    plane at steps 1-3, inner, clipped at both corners, empty, and three whose one coarse corner is step - 1 pixels
    from the thing at steps 2-4) on four planes: textured, a planted copy, a dimmed noisy
    copy, and a repeating pattern."""
    import hashlib
    import ncc_cases as nc
    g = json.load(open(os.path.join(GOLD, "ncc_refine.json")))
    byname = {c["name"]: c for c in g["cases"]}
    for name, img, (cx, cy) in nc.cases(oracle):
        case = byname[name]
        h, w = img.shape
        hp = nc.host_plane(oracle, img)
        app, norm = oracle.ncc_template(hp, int(cx) - nc.TW // 2, int(cy) - nc.TH // 2, nc.TW, nc.TH)
        assert hashlib.sha256(app.tobytes()).hexdigest() == case["template_sha256"], name
        assert float(norm).hex() == case["norm"], name
        for hx, hy, want in case["half"]:
            got = oracle.lib.vo_ncc_half(C.byref(hp.c), app.ctypes.data, nc.TW, nc.TH, norm, hx, hy)
            assert float(got).hex() == want, (name, hx, hy)
        x0, y0 = int(cx) - nc.TW // 2, int(cy) - nc.TH // 2
        assert [tuple(s["window"]) for s in case["searches"]] == nc.windows(w, h, (x0, y0))
        for srch in case["searches"]:
            a, b, c, d, step = srch["window"]
            f, cells = oracle.ncc_search(hp, app, nc.TW, nc.TH, norm, a, b, c, d, step)
            assert (f.hx, f.hy, float(f.ncc).hex(), float(f.second).hex(), f.cells) == \
                (srch["hx"], srch["hy"], srch["ncc"], srch["second"], srch["cells"]), (name, srch["window"])
            bits = np.concatenate([np.ascontiguousarray(cells[k]).view(np.uint32) for k in ("v", "x", "y")])
            assert hashlib.sha256(bits.tobytes()).hexdigest() == srch["cells_sha256"], (name, srch["window"])
    for l, c, r, want in g["peak_offset"]:
        got = oracle.lib.vo_peak_offset(float.fromhex(l), float.fromhex(c), float.fromhex(r))
        assert float(got).hex() == want, (l, c, r)


def test_ncc_refine_closed_forms(oracle):
    """What the mathematics fixes about the refinement, independent of the reference:
    - the parabola through (-1, f(-1)), (0, f(0)), (1, f(1)) of a downward quadratic peaks at its vertex v, clamped to
      +-0.5; a flat or upward one gives 0;
    - a template cut from an image moved right by half a pixel scores 1 at the odd half-pixel corner (the 2-tap
      average restores the original exactly), and one cut moved down by half a pixel at the odd row;
    - a second, clear copy of the thing (more than half a template away) is the second peak, scoring what the plain
      score gives at that copy; a copy nearer than that is not counted;
    - corners whose half-pixel window leaves the plane score -1."""
    import ncc_cases as nc
    rng = np.random.default_rng(11)
    po = oracle.lib.vo_peak_offset
    for _ in range(200):
        a, v, c0 = -float(rng.uniform(0.1, 5)), float(rng.uniform(-0.9, 0.9)), float(rng.normal())
        f = [a * (x - v) ** 2 + c0 for x in (-1, 0, 1)]
        want = min(0.5, max(-0.5, v))
        assert abs(po(*f) - want) < 1e-12, (a, v)
        assert po(*[-y for y in f]) == 0.0 and po(c0, c0, c0) == 0.0
    # half-pixel copies
    img, sh = nc.half_shifted(200, 120, 3)
    hp, hs = nc.host_plane(oracle, img), nc.host_plane(oracle, sh)
    x0, y0 = 70, 40
    app, norm = oracle.ncc_template(hs, x0, y0, nc.TW, nc.TH)
    v = oracle.lib.vo_ncc_half(C.byref(hp.c), app.ctypes.data, nc.TW, nc.TH, norm, 2 * x0 + 1, 2 * y0)
    assert v > 1 - 1e-6
    assert oracle.lib.vo_ncc_half(C.byref(hp.c), app.ctypes.data, nc.TW, nc.TH, norm, 2 * x0, 2 * y0) < v
    f, _ = oracle.ncc_search(hp, app, nc.TW, nc.TH, norm, 1, 200, 1, 120, 1)
    assert (f.hx, f.hy) == (2 * x0 + 1, 2 * y0) and f.ncc > 1 - 1e-6
    imgT, shT = img.T.copy(), sh.T.copy()     # the same, down by half a pixel
    hpT, hsT = nc.host_plane(oracle, imgT), nc.host_plane(oracle, shT)
    app, norm = oracle.ncc_template(hsT, y0, x0, nc.TW, nc.TH)
    v = oracle.lib.vo_ncc_half(C.byref(hpT.c), app.ctypes.data, nc.TW, nc.TH, norm, 2 * y0, 2 * x0 + 1)
    assert v > 1 - 1e-6
    # a planted clear copy is the second peak; a near one is not
    far = nc.planted(288, 176, 21, (60, 50), (180, 100), size=(32, 21), gain=0.9, noise=4.0)
    hp = nc.host_plane(oracle, far)
    app, norm = oracle.ncc_template(hp, 60, 50, nc.TW, nc.TH)
    f, _ = oracle.ncc_search(hp, app, nc.TW, nc.TH, norm, 1, 288, 1, 176, 1)
    at_copy = max(oracle.lib.vo_ncc_half(C.byref(hp.c), app.ctypes.data, nc.TW, nc.TH, norm, 360 + dx, 200 + dy)
                  for dx in (-1, 0, 1) for dy in (-1, 0, 1))
    assert (f.hx, f.hy) == (120, 100) and f.ncc > 1 - 1e-6
    assert float(f.second) == float(at_copy) and 0.9 < f.second < 1 - 1e-4
    near = nc.planted(288, 176, 21, (60, 50), (60 + 12, 50), size=(32, 21))
    hp = nc.host_plane(oracle, near)
    app, norm = oracle.ncc_template(hp, 60, 50, nc.TW, nc.TH)
    f, _ = oracle.ncc_search(hp, app, nc.TW, nc.TH, norm, 1, 288, 1, 176, 1)
    assert f.second < 0.9, float(f.second)
    # a window whose only coarse corner is step - 1 pixels off in both axes still finds the exact copy
    for s in (2, 3, 4):
        f, _ = oracle.ncc_search(hp, app, nc.TW, nc.TH, norm, 60 - s + 1, 60, 50 - s + 1, 50, s)
        assert (f.hx, f.hy) == (120, 100) and f.ncc > 1 - 1e-6, s
    # off the plane
    for hx, hy in [(0, 10), (10, 0), (2 * (288 - nc.TW - 2), 10), (10, 2 * (176 - nc.TH - 2))]:
        assert oracle.lib.vo_ncc_half(C.byref(hp.c), app.ctypes.data, nc.TW, nc.TH, norm, hx, hy) == -1.0


def _textbook_zncc(win, tmpl):
    a = win.astype(np.float64) - win.mean()
    t = tmpl.astype(np.float64) - tmpl.mean()
    return float((a * t).sum() / np.sqrt((a * a).sum() * (t * t).sum()))


def test_ncc_closed_forms(oracle):
    """Affine copies of the template score +1 and -1; a blank window scores 0 (the energy floor); corners that do
    not fit score -1; small windows agree with the textbook fp64 ZNCC."""
    rng = np.random.default_rng(5)
    w, h = 120, 90
    img = rng.integers(40, 200, (h, w)).astype(np.uint8)
    hp = _plane_from(img)
    _pad(oracle, hp)
    app, norm = oracle.ncc_template(hp, 10, 12, 32, 21)
    assert abs(oracle.ncc_at(hp, app, 32, 21, norm, 10, 12) - 1.0) < 1e-6
    # affine copy and inverted copy elsewhere
    win = img[12:33, 10:42].astype(np.int64)
    img2 = img.copy()
    img2[50:71, 60:92] = np.clip(win // 2 + 30, 0, 255)
    img2[20:41, 70:102] = 255 - win
    hp2 = _plane_from(img2)
    _pad(oracle, hp2)
    assert oracle.ncc_at(hp2, app, 32, 21, norm, 60, 50) > 0.999
    assert oracle.ncc_at(hp2, app, 32, 21, norm, 70, 20) < -0.9999
    flat = _plane_from(np.full((h, w), 100, np.uint8))
    _pad(oracle, flat)
    assert oracle.ncc_at(flat, app, 32, 21, norm, 30, 30) == 0.0
    assert oracle.ncc_at(hp, app, 32, 21, norm, 0, 5) == -1.0
    assert oracle.ncc_at(hp, app, 32, 21, norm, w - 33, 5) == -1.0
    for (x, y) in [(3, 4), (40, 30), (77, 60)]:
        ref = _textbook_zncc(img[y:y + 21, x:x + 32], img[12:33, 10:42])
        assert abs(oracle.ncc_at(hp, app, 32, 21, norm, x, y) - ref) < 2e-6


def test_ncc_search_brute_force(oracle):
    """search_slot's map and pick against a Python brute force over the same scores: per-cell first maximum,
    top 8 by (score, cell order), half-pixel refinement."""
    rng = np.random.default_rng(11)
    img = oracle.m2_render(96, 80, 300, 500)[:, :96]
    hp = _plane_from(img)
    _pad(oracle, hp)
    tw, th = 16, 16
    app, norm = oracle.ncc_template(hp, 40, 30, tw, th)
    f, cells = oracle.ncc_search(hp, app, tw, th, norm, 1, 96, 1, 80, 1)
    x1, y1 = 96 - tw - 2, 80 - th - 2
    cols, rows = (x1 - 1) // 16 + 1, (y1 - 1) // 16 + 1
    assert len(cells) == cols * rows
    for ci in range(len(cells)):
        cy0, cx0 = 1 + (ci // cols) * 16, 1 + (ci % cols) * 16
        best = (-2.0, cx0, cy0)
        for y in range(cy0, min(cy0 + 16, y1 + 1)):
            for x in range(cx0, min(cx0 + 16, x1 + 1)):
                v = np.float32(oracle.ncc_at(hp, app, tw, th, norm, x, y))
                if v > best[0]:
                    best = (v, x, y)
        assert (cells[ci]["v"], cells[ci]["x"], cells[ci]["y"]) == best
    assert (f.hx, f.hy) == (80, 60) and abs(f.ncc - 1.0) < 1e-6


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref not built (no reference tree here)")
def test_live_reference_lk_and_ncc(oracle):
    """Where oracle/_ref exists: the lane tier equals the reference's lk_track_vx and lk_track_ref's distance stays
    in the published x87-vs-MVEX band on the m2 inputs; vo_ncc_at equals tk_correlate at random corners."""
    from oracle_ffi import LkCfg, Plane
    ref = C.CDLL(REF_SO)
    ref.lk_track_vx.argtypes = [C.POINTER(Plane), C.POINTER(Plane), C.c_int, C.POINTER(LkCfg), C.c_float, C.c_float,
                                C.c_float, C.c_float, C.c_void_p, C.c_void_p]
    ref.lk_tile_bytes.restype = C.c_size_t
    g, w, h, cfg, xs, ys = _m2_setup(oracle)
    tile = np.zeros(ref.lk_tile_bytes() // 4 + 64, np.float32)
    tp = (tile.ctypes.data + 63) & ~63
    k0, I = oracle.pyramid(oracle.m2_render(w, h, 2048, 2048), w, h, 4)
    k1, J = oracle.pyramid(oracle.m2_render(w, h, 2048 - 13, 2048 - 7), w, h, 4)
    mine = oracle.lk_batch(0, I, J, 3, cfg, xs, ys)
    theirs = np.zeros(len(xs), LK_OUT_DTYPE)
    for i in range(len(xs)):
        ref.lk_track_vx(I, J, 3, C.byref(cfg), xs[i], ys[i], 0.0, 0.0, theirs[i:i + 1].ctypes.data, tp)
    assert mine.tobytes() == theirs.tobytes()
    ref.tk_correlate.restype = C.c_longdouble
    ref.tk_correlate.argtypes = [C.POINTER(Plane), C.c_void_p, C.c_longdouble, C.c_int, C.c_int]
    rng = np.random.default_rng(3)
    img = oracle.m2_render(200, 150, 77, 99)[:, :200]
    hp = _plane_from(img)
    _pad(oracle, hp)
    app, norm = oracle.ncc_template(hp, 50, 60, 32, 21)
    for _ in range(200):
        x, y = int(rng.integers(-3, 175)), int(rng.integers(-3, 135))
        assert oracle.ncc_at(hp, app, 32, 21, norm, x, y) == ref.tk_correlate(C.byref(hp.c), app.ctypes.data, norm, x,
                                                                              y)
