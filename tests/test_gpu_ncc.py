"""GPU NCC search against the oracle: the template bit for bit, scores within 1e-5, the same per-cell argmax where
the oracle's winner is clear, and the same found corner and score; plus the full 1080p 64 x 64 exhaustive search."""
import numpy as np
import pytest

import vtgpu
from gpu_util import say

pytestmark = pytest.mark.gpu


def _setup(ctx, oracle, w, h, seed, levels=1):
    f0, f1, mo = vtgpu.frame_pair(w, h, seed)
    p0, p1 = ctx.pyramid(w, h, levels), ctx.pyramid(w, h, levels)
    p0.build(f0)
    p1.build(f1)
    k0, a0 = oracle.pyramid(f0, w, h, levels, threads=8)
    k1, a1 = oracle.pyramid(f1, w, h, levels, threads=8)
    return f0, f1, mo, p0, p1, k0, k1


@pytest.mark.parametrize("level,tw,th", [(0, 64, 64), (0, 32, 21), (1, 32, 21), (0, 16, 24)])
def test_template_exact(ctx, oracle, level, tw, th):
    f0, f1, mo, p0, p1, k0, k1 = _setup(ctx, oracle, 640, 480, 21, levels=2)
    say("ncc template", level, tw, th)
    t = ctx.ncc_template(p0, level, 101, 77, tw, th)
    app, norm = t.read()
    oapp, onorm = oracle.ncc_template(k0[level], 101, 77, tw, th)
    assert np.array_equal(app, oapp) and norm == np.float32(onorm)
    t.destroy()


@pytest.mark.parametrize("level,tw,th", [(0, 64, 64), (0, 32, 21), (1, 32, 21)])
def test_scores_whole_and_half(ctx, oracle, level, tw, th):
    """tk_correlate / search_half at random corners, the border (-1) and blank-ish windows: within 1e-5."""
    f0, f1, mo, p0, p1, k0, k1 = _setup(ctx, oracle, 640, 480, 22, levels=2)
    t = ctx.ncc_template(p0, level, 150, 100, tw, th)
    app, norm = t.read()
    rng = np.random.default_rng(level * 7 + tw)
    lw, lh = k1[level].w, k1[level].h
    xy = np.stack([rng.integers(-2, lw - tw + 2, 300), rng.integers(-2, lh - th + 2, 300)], 1).astype(np.int32)
    xy[0] = (150 + int(round(mo.dx)) >> level, 100 + int(round(mo.dy)) >> level)
    say("ncc scores", level, tw, th)
    got = ctx.ncc_score(p1, t, xy)
    ref = np.array([oracle.ncc_at(k1[level], app, tw, th, norm, int(x), int(y)) for x, y in xy])
    assert np.all((ref == -1) == (got == -1)) and np.all((ref == 0) == (got == 0))
    assert np.abs(got - ref).max() <= 1e-5, np.abs(got - ref).max()
    hxy = np.stack([rng.integers(-2, 2 * (lw - tw), 300), rng.integers(-2, 2 * (lh - th), 300)], 1).astype(np.int32)
    got = ctx.ncc_score(p1, t, hxy, half=True)
    ref = np.array([oracle.lib.vo_ncc_half(__import__("ctypes").byref(k1[level].c), app.ctypes.data, tw, th, norm,
                                           int(x), int(y)) for x, y in hxy])
    assert np.abs(got - ref).max() <= 1e-5
    t.destroy()


def test_refine_reference_fixture(ctx, oracle):
    """The GPU against the reference's own store_appearance, search_half and search_slot (tests/golden/ncc_refine.json,
    written from src/kern/track.c compiled from source): the template bit for bit, half-pixel scores within 1e-5 (-1
    exactly where the window leaves the plane), and every search's pick equal, its score and second peak within 1e-5,
    over the fixture's four planes and the windows at steps 1 and 2 (whole, inner, clipped, empty, one coarse corner
    a pixel off); the steps above 2, which the tracker never uses, are refused with VTG_EINVAL."""
    import hashlib
    import json
    import os
    import ncc_cases as nc
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ncc_refine.json")))
    byname = {c["name"]: c for c in g["cases"]}
    for name, img, (cx, cy) in nc.cases(oracle):
        case = byname[name]
        h, w = img.shape
        x0, y0 = int(cx) - nc.TW // 2, int(cy) - nc.TH // 2
        say("ncc refine fixture", name, w, h)
        p = ctx.pyramid(w, h, 1)
        p.build(np.ascontiguousarray(img))
        t = ctx.ncc_template(p, 0, x0, y0, nc.TW, nc.TH)
        app, norm = t.read()
        assert hashlib.sha256(app.tobytes()).hexdigest() == case["template_sha256"], name
        assert float(norm).hex() == case["norm"], name
        hxy = np.array([[a, b] for a, b, _ in case["half"]], np.int32)
        want = np.array([float.fromhex(v) for _, _, v in case["half"]])
        got = ctx.ncc_score(p, t, hxy, half=True)
        assert np.array_equal(got == -1, want == -1), name
        assert np.abs(got - want).max() <= 1e-5, (name, np.abs(got - want).max())
        for srch in case["searches"]:
            a, b, c, d, step = srch["window"]
            if step > 2:   # the tracker searches at steps 1 and 2 only (track.c:1796, 1813, 1836); others refused
                with pytest.raises(vtgpu.VtgError, match=r"\(-22\)"):
                    ctx.ncc_search(p, t, a, b, c, d, step)
                continue
            f = ctx.ncc_search(p, t, a, b, c, d, step)
            assert (f.hx, f.hy, f.cells) == (srch["hx"], srch["hy"], srch["cells"]), (name, srch["window"])
            assert abs(f.ncc - float.fromhex(srch["ncc"])) <= 1e-5, (name, srch["window"])
            assert abs(f.second - float.fromhex(srch["second"])) <= 1e-5, (name, srch["window"])
        t.destroy()
        p.destroy()


def _textured_corner(img, tw, th, seed):
    """A corner whose tw x th window has strong texture (the searches are only interesting there)."""
    rng = np.random.default_rng(seed)
    h, w = img.shape
    best, at = -1.0, (2, 2)
    for _ in range(200):
        x, y = int(rng.integers(4, w - tw - 4)), int(rng.integers(4, h - th - 4))
        sd = img[y:y + th, x:x + tw].std()
        if sd > best:
            best, at = sd, (x, y)
    return at


def _check_cells(gcells, ocells, score_at):
    """Scores within 1e-5; the same corner unless the oracle's best is within 1e-5 of the GPU's pick."""
    assert len(gcells) == len(ocells)
    dv = np.abs(gcells["v"].astype(np.float64) - ocells["v"])
    assert dv.max() <= 1e-5, dv.max()
    same = (gcells["x"] == ocells["x"]) & (gcells["y"] == ocells["y"])
    for i in np.nonzero(~same)[0]:
        assert abs(score_at(gcells["x"][i], gcells["y"][i]) - ocells["v"][i]) <= 1e-5
    return same.mean()


@pytest.mark.parametrize("w,h,tw,th,step", [(400, 300, 64, 64, 1), (400, 300, 32, 24, 1), (400, 300, 16, 16, 2),
                                            (400, 300, 32, 21, 1), (400, 300, 64, 64, 2), (400, 300, 64, 40, 1),
                                            (333, 257, 64, 21, 2), (517, 203, 64, 7, 1)])
def test_search_matches_oracle(ctx, oracle, w, h, tw, th, step):
    f0, f1, mo, p0, p1, k0, k1 = _setup(ctx, oracle, w, h, 31 + tw + step)
    x0, y0 = _textured_corner(f0, tw, th, tw + th)
    t = ctx.ncc_template(p0, 0, x0, y0, tw, th)
    app, norm = t.read()
    ncell = ((w - tw - 2 - 1) // 16 + 1) * ((h - th - 2 - 1) // 16 + 1)
    cells = ctx.alloc(12 * ncell)
    say("ncc search", w, h, tw, th, step)
    f = ctx.ncc_search(p1, t, 1, w, 1, h, step, cells)
    g = cells.get(np.zeros(ncell, vtgpu.CELL_DTYPE))
    of, oc = oracle.ncc_search(k1[0], app, tw, th, norm, 1, w, 1, h, step)
    frac = _check_cells(g, oc, lambda x, y: oracle.ncc_at(k1[0], app, tw, th, norm, int(x), int(y)))
    print(f"[ncc] {w}x{h} {tw}x{th} step {step}: cells {ncell}, same corner {frac:.4f}, found ({f.hx},{f.hy}) "
          f"{f.ncc:.6f} vs ({of.hx},{of.hy}) {float(of.ncc):.6f}")
    assert (f.hx, f.hy) == (of.hx, of.hy)
    assert abs(f.ncc - float(of.ncc)) <= 1e-5 and abs(f.second - float(of.second)) <= 1e-5
    assert f.ncc > 0.5
    hp = k1[0]
    import ctypes as C
    l = oracle.lib.vo_ncc_half(C.byref(hp.c), app.ctypes.data, tw, th, norm, of.hx - 1, of.hy)
    r = oracle.lib.vo_ncc_half(C.byref(hp.c), app.ctypes.data, tw, th, norm, of.hx + 1, of.hy)
    assert abs(f.ox - float(oracle.lib.vo_peak_offset(l, of.ncc, r))) <= 1e-3
    cells.free()
    t.destroy()


def test_search_1080_full(ctx, oracle):
    """The native workload: a 64 x 64 template cut from frame 0 at a seeded corner, exhaustive over frame 1."""
    w, h = 1920, 1080
    f0, f1, mo, p0, p1, k0, k1 = _setup(ctx, oracle, w, h, 2)
    rng = np.random.default_rng(2)
    x0, y0 = int(rng.integers(200, w - 264)), int(rng.integers(200, h - 264))
    t = ctx.ncc_template(p0, 0, x0, y0, 64, 64)
    app, norm = t.read()
    ncell = ((w - 66 - 1) // 16 + 1) * ((h - 66 - 1) // 16 + 1)
    cells = ctx.alloc(12 * ncell)
    say("ncc search 1080 full")
    f = ctx.ncc_search(p1, t, 1, w, 1, h, 1, cells)
    g = cells.get(np.zeros(ncell, vtgpu.CELL_DTYPE))
    of, oc = oracle.ncc_search(k1[0], app, 64, 64, norm, 1, w, 1, h, 1, threads=16)
    frac = _check_cells(g, oc, lambda x, y: oracle.ncc_at(k1[0], app, 64, 64, norm, int(x), int(y)))
    tr = vtgpu.truth(mo, w, h, [x0 + 32], [y0 + 32])[0]
    print(f"[ncc] 1080p: same corner {frac:.4f}; found ({f.hx / 2},{f.hy / 2}) {f.ncc:.6f}; the template moved to "
          f"({x0 + tr[0]:.2f},{y0 + tr[1]:.2f})")
    assert (f.best_x, f.best_y) == (int(oc[np.argmax(oc['v'])]["x"]), int(oc[np.argmax(oc['v'])]["y"]))
    assert (f.hx, f.hy) == (of.hx, of.hy) and abs(f.ncc - float(of.ncc)) <= 1e-5
    # the rotated, noisy copy peaks within a few pixels of where the template's centre moved (smooth blobs)
    assert abs(f.hx / 2 - (x0 + tr[0])) <= 5 and abs(f.hy / 2 - (y0 + tr[1])) <= 5
    cells.free()
    t.destroy()


def test_search_large_map(ctx, oracle):
    """A map of more than 32,768 cells (4200 x 2200, a 16 x 16 template: 262 x 137 cells), where the top-8 selection
    reads the map from memory each round instead of keeping it in registers: the pick, its score and the second peak
    equal the oracle's."""
    w, h, tw, th = 4200, 2200, 16, 16
    f0, f1, mo, p0, p1, k0, k1 = _setup(ctx, oracle, w, h, 77)
    x0, y0 = _textured_corner(f0, tw, th, tw + th)
    t = ctx.ncc_template(p0, 0, x0, y0, tw, th)
    app, norm = t.read()
    ncell = ((w - tw - 2 - 1) // 16 + 1) * ((h - th - 2 - 1) // 16 + 1)
    assert ncell > 32768
    say("ncc search large map", ncell)
    f = ctx.ncc_search(p1, t, 1, w, 1, h, 1)
    of, oc = oracle.ncc_search(k1[0], app, tw, th, norm, 1, w, 1, h, 1, threads=16)
    assert (f.hx, f.hy) == (of.hx, of.hy)
    assert abs(f.ncc - float(of.ncc)) <= 1e-5 and abs(f.second - float(of.second)) <= 1e-5
    t.destroy()


@pytest.mark.parametrize("tw,th,level,step", [(64, 64, 0, 1), (32, 21, 1, 2), (32, 24, 0, 2)])
def test_search_many_matches_single(ctx, tw, th, level, step):
    """Several searches in one call (frames of different sizes, partial windows, one window that clips to nothing and
    more searches than one launch holds) give the records of the single-search call, field for field."""
    import ctypes as C
    sizes = [(640, 480), (800, 600), (1280, 720), (640, 480), (960, 540), (640, 480), (720, 400), (1024, 768),
             (640, 480), (800, 600)]
    pyrs, tms, regions = [], [], []
    for k, (w, h) in enumerate(sizes):
        f0, f1, _ = vtgpu.frame_pair(w, h, 500 + k)
        p0, p1 = ctx.pyramid(w, h, 2), ctx.pyramid(w, h, 2)
        p0.build(f0)
        p1.build(f1)
        lw, lh = (w, h) if level == 0 else ((w + 1) // 2, (h + 1) // 2)
        tms.append(ctx.ncc_template(p0, level, lw // 3, lh // 4, tw, th))
        pyrs.append((p0, p1))
        regions.append((1, lw, 1, lh) if k % 3 else (lw // 4, lw // 2 + 7, lh // 5, lh // 2 + 3))
    regions[4] = (lw * 4, lw * 5, 1, 10)  # clips to nothing
    fs = C.sizeof(vtgpu.Found)
    found = ctx.alloc(fs * len(sizes))
    say("ncc search_many", tw, th, level, step)
    ctx.ncc_search_many([p[1] for p in pyrs], tms, regions, step, found)
    ctx.sync()
    got = (vtgpu.Found * len(sizes))()
    ctx.check(vtgpu.lib().vtg_download(ctx.h, C.cast(got, C.c_void_p), found.ptr, fs * len(sizes)))
    for k in range(len(sizes)):
        x0, x1, y0, y1 = regions[k]
        one = ctx.ncc_search(pyrs[k][1], tms[k], x0, x1, y0, y1, step)
        g = got[k]
        assert bytes(one) == bytes(g), (k, (one.hx, one.hy, one.ncc), (g.hx, g.hy, g.ncc))
        if k == 4:
            assert g.cells == 0 and g.ncc == -2.0
        else:
            assert g.cells > 0
    found.free()
    for t in tms:
        t.destroy()
    for p0, p1 in pyrs:
        p0.destroy()
        p1.destroy()


def test_template_update_many_matches_single(ctx):
    """Re-cutting several templates in one call (mixed levels, so the u8 and u16 runs launch apart, and more than one
    launch holds) gives each template the bits of its own single cut."""
    sizes = [(640, 480), (800, 600), (1280, 720), (640, 480), (960, 540), (720, 400), (1024, 768), (640, 480),
             (800, 600), (640, 480), (1280, 720)]
    pyrs, levels, corners = [], [], []
    for k, (w, h) in enumerate(sizes):
        f0, _, _ = vtgpu.frame_pair(w, h, 700 + k)
        p = ctx.pyramid(w, h, 2)
        p.build(f0)
        pyrs.append(p)
        lv = 1 if k in (3, 4, 9) else 0
        levels.append(lv)
        lw, lh = (w, h) if lv == 0 else ((w + 1) // 2, (h + 1) // 2)
        corners.append((lw // 3 + k, lh // 4 + 2 * k))
    tw, th = 32, 21
    many = [ctx.ncc_template(pyrs[k], levels[k], 5, 5, tw, th) for k in range(len(sizes))]
    say("ncc template update_many")
    ctx.ncc_templates_update(many, pyrs, levels, corners)
    for k in range(len(sizes)):
        one = ctx.ncc_template(pyrs[k], levels[k], corners[k][0], corners[k][1], tw, th)
        a1, n1 = one.read()
        a2, n2 = many[k].read()
        assert np.array_equal(a1, a2) and n1 == n2, k
        one.destroy()
    for t in many:
        t.destroy()
    for p in pyrs:
        p.destroy()


@pytest.mark.parametrize("tw,th", [(64, 64), (32, 21)])
def test_search_split_matches_single_context(ctx, tw, th):
    """The search run on another context (ordered after this one's work so far, beside whatever this one does next)
    gives the records of the one-context call, field for field."""
    import ctypes as C
    other = vtgpu.Context(0)
    sizes = [(1920, 1080), (640, 480), (1280, 720), (800, 600), (1920, 1080), (640, 480), (960, 540), (1024, 768),
             (640, 480)]
    pyrs, tms, regions = [], [], []
    for k, (w, h) in enumerate(sizes):
        f0, f1, _ = vtgpu.frame_pair(w, h, 900 + k)
        p0, p1 = ctx.pyramid(w, h, 1), ctx.pyramid(w, h, 1)
        p0.build(f0)
        p1.build(f1)
        tms.append(ctx.ncc_template(p0, 0, w // 3, h // 4, tw, th))
        pyrs.append((p0, p1))
        regions.append((1, w, 1, h) if k % 2 else (w // 5, w // 2, h // 6, h // 2))
    fs = C.sizeof(vtgpu.Found)
    a, b = ctx.alloc(fs * len(sizes)), ctx.alloc(fs * len(sizes))
    say("ncc search_many_split", tw, th)
    ctx.ncc_search_many([p[1] for p in pyrs], tms, regions, 1, a)
    ctx.ncc_search_many_split(other, [p[1] for p in pyrs], tms, regions, 1, b)
    ctx.wait(other)
    ctx.sync()
    ga, gb = (vtgpu.Found * len(sizes))(), (vtgpu.Found * len(sizes))()
    ctx.check(vtgpu.lib().vtg_download(ctx.h, C.cast(ga, C.c_void_p), a.ptr, fs * len(sizes)))
    ctx.check(vtgpu.lib().vtg_download(ctx.h, C.cast(gb, C.c_void_p), b.ptr, fs * len(sizes)))
    for k in range(len(sizes)):
        assert bytes(ga[k]) == bytes(gb[k]), k
        assert ga[k].cells > 0 and ga[k].ncc > -1.0
    a.free()
    b.free()
    for t in tms:
        t.destroy()
    for p0, p1 in pyrs:
        p0.destroy()
        p1.destroy()
    other.close()


def test_split_searches_in_a_row(ctx):
    """Split searches back to back with no join between them (the second over smaller windows: both use the finish
    context's scratch, in its stream order), joined once, then one-context searches on the map context, the first
    needing more scratch than before: every record equals the one-context call's."""
    import ctypes as C
    other = vtgpu.Context(0)
    sizes = [(1920, 1080), (640, 480), (1280, 720)]
    pyrs, tms = [], []
    for k, (w, h) in enumerate(sizes):
        f0, f1, _ = vtgpu.frame_pair(w, h, 950 + k)
        p0, p1 = ctx.pyramid(w, h, 1), ctx.pyramid(w, h, 1)
        p0.build(f0)
        p1.build(f1)
        tms.append(ctx.ncc_template(p0, 0, w // 3, h // 4, 64, 64))
        pyrs.append((p0, p1))
    big = [(1, w, 1, h) for (w, h) in sizes]
    small = [(w // 4, w // 2, h // 4, h // 2) for (w, h) in sizes]
    fs, n = C.sizeof(vtgpu.Found), len(sizes)
    bufs = [ctx.alloc(fs * n) for _ in range(5)]
    say("ncc split searches in a row")
    ctx.ncc_search_many_split(other, [p[1] for p in pyrs], tms, big, 1, bufs[0])
    ctx.ncc_search_many_split(other, [p[1] for p in pyrs], tms, small, 1, bufs[1])
    ctx.wait(other)
    many = [p[1] for p in pyrs] * 3
    grow = ctx.alloc(fs * 3 * n)
    ctx.ncc_search_many(many, tms * 3, big * 3, 1, grow)  # grows the scratch
    ctx.ncc_search_many([p[1] for p in pyrs], tms, big, 1, bufs[2])
    ctx.ncc_search_many([p[1] for p in pyrs], tms, small, 1, bufs[3])
    ctx.sync()
    got = []
    for b in bufs[:4]:
        g = (vtgpu.Found * n)()
        ctx.check(vtgpu.lib().vtg_download(ctx.h, C.cast(g, C.c_void_p), b.ptr, fs * n))
        got.append(g)
    for k in range(n):
        assert bytes(got[0][k]) == bytes(got[2][k]), ("big", k)
        assert bytes(got[1][k]) == bytes(got[3][k]), ("small", k)
        assert got[1][k].cells < got[0][k].cells
    for b in bufs + [grow]:
        b.free()
    for t in tms:
        t.destroy()
    for p0, p1 in pyrs:
        p0.destroy()
        p1.destroy()
    other.close()


def test_f32u8_kept_frame_build_many(ctx):
    """F32_U8 pyramids of uneven sizes and depths built together keep each frame beside level 0 (written by the level-1
    reduce from its tile, or copied for a one-level pyramid): a level-0 search over each gives the record the same
    template gives over a Q6 pyramid of the frame built alone (the matrix-core path for 64 x 64, the dot4 one for 32 x 21)."""
    sizes = [(640, 480, 3), (1280, 720, 5), (97, 61, 2), (333, 250, 1)]
    frames = [vtgpu.frame_pair(w, h, 70 + i) for i, (w, h, _) in enumerate(sizes)]
    pyrs = [ctx.pyramid(w, h, L, vtgpu.STORE_F32_U8) for w, h, L in sizes]
    dev = [ctx.to_device(f1) for _, f1, _ in frames]
    ctx.build_many(pyrs, dev)
    for (w, h, L), (f0, f1, _), p in zip(sizes, frames, pyrs):
        tw, th = (64, 64) if w >= 200 else (32, 21)
        say("f32_u8 kept frame", w, h, L, tw, th)
        a, b = ctx.pyramid(w, h, L), ctx.pyramid(w, h, L)
        a.build(f0)
        b.build(f1)
        t = ctx.ncc_template(a, 0, w // 3, h // 4, tw, th)
        ref = ctx.ncc_search(b, t, 1, w, 1, h, 1)
        got = ctx.ncc_search(p, t, 1, w, 1, h, 1)
        assert bytes(got) == bytes(ref), (w, h, L)
        assert ref.ncc > -1.0   # a pick was made
        t.destroy()
        a.destroy()
        b.destroy()
        p.destroy()
    for d in dev:
        d.free()


@pytest.mark.parametrize("tw,th,level", [(64, 64, 0), (32, 21, 1), (32, 24, 0)])
def test_search_q6f_equals_q6(ctx, tw, th, level):
    """Q6F pyramids (u8 level 0, fp32 Q6 above) and F32_U8 pyramids (fp32 Q6 levels, the u8 frame kept beside level 0)
    hold the Q6 pyramid's values: their templates and searches give the Q6 pyramid's records, field for field, on the
    matrix-core path (level 0, 64 x 64, the kept frame for F32_U8) and the generic one (level 1)."""
    w, h = 800, 600
    f0, f1, _ = vtgpu.frame_pair(w, h, 612)
    out = []
    for st in (vtgpu.STORE_Q6, vtgpu.STORE_Q6F, vtgpu.STORE_F32_U8):
        p0, p1 = ctx.pyramid(w, h, 2, st), ctx.pyramid(w, h, 2, st)
        p0.build(f0)
        p1.build(f1)
        lw, lh = (w, h) if level == 0 else ((w + 1) // 2, (h + 1) // 2)
        t = ctx.ncc_template(p0, level, lw // 3, lh // 4, tw, th)
        say("ncc q6f", st, tw, th, level)
        out.append(bytes(ctx.ncc_search(p1, t, 1, lw, 1, lh, 1)))
        t.destroy()
        p0.destroy()
        p1.destroy()
    assert out[0] == out[1] == out[2]


def _rounding_bound(img, x0, y0, tw, th, app, norm):
    """4 * 64 * 255 * sum |e_k| / norm for the u8 template at (x0, y0): e_k = t_k - (64 u_k - mean), the stored
    template's fp32 rounding (the template equals the oracle's bit for bit, test_template_exact)."""
    u = img[y0:y0 + th, x0:x0 + tw].astype(np.float64)
    mean = np.float32(64.0 * u.sum() / (tw * th))
    app = np.asarray(app, np.float32).reshape(th, tw)
    assert np.array_equal(app, np.float32(64.0) * u.astype(np.float32) - mean)
    e = app.astype(np.float64) - (64.0 * u - np.float64(mean))
    return 4 * 64 * 255 * np.abs(e).sum() / float(norm)


@pytest.mark.parametrize("kind,tw,th", [("dark01", 64, 48), ("dark01", 32, 24), ("dark01", 64, 64),
                                        ("uniform", 64, 40)])
def test_search_template_rounding(ctx, oracle, kind, tw, th):
    """The fast maps (matrix cores for 64-wide templates, dot4 for 16/32-wide ones) take the stored template as
    64 u - mean exactly. Where tw th is not a power of two and the rounding term's bound exceeds 1e-6 (a near-black
    template of 0s and 1s: 1.8e-5), the template kernel flags it and the search's map is the reference's arithmetic:
    the u8 level's map equals the fp32 level's (ncc_cells_generic) bit for bit. Below the bound (a power-of-two size,
    where every e_k is 0, or a uniform texture) the fast map runs, within 1e-5 of the oracle. Every search picks what
    the oracle picks."""
    w, h = 400, 300
    rng = np.random.default_rng(11 + tw + th)
    if kind == "dark01":
        f0 = (rng.random((h, w)) < 0.15).astype(np.uint8)
    else:
        f0 = rng.integers(0, 256, (h, w)).astype(np.uint8)
    f1 = np.roll(f0, (4, 7), axis=(0, 1))
    noise = rng.random((h, w)) < 0.02
    f1[noise] = (1 - f1[noise]) if kind == "dark01" else rng.integers(0, 256, int(noise.sum()))
    f1 = np.ascontiguousarray(f1)
    p0, p1 = ctx.pyramid(w, h, 1), ctx.pyramid(w, h, 1)
    q1 = ctx.pyramid(w, h, 1, vtgpu.STORE_F32)
    p0.build(f0)
    p1.build(f1)
    q1.build(f1)
    want = kind == "dark01" and (tw * th) & (tw * th - 1) != 0  # a template to flag: one whose mean rounds
    y0 = 120
    for x0 in range(150, 190):
        t = ctx.ncc_template(p0, 0, x0, y0, tw, th)
        app, norm = t.read()
        bound = _rounding_bound(f0, x0, y0, tw, th, app, norm)
        if bound > 0 or not want:
            break
        t.destroy()
    flagged = bound > 1e-6
    assert flagged == want, bound
    ncell = ((w - tw - 2 - 1) // 16 + 1) * ((h - th - 2 - 1) // 16 + 1)
    cq, cf = ctx.alloc(12 * ncell), ctx.alloc(12 * ncell)
    say("ncc template rounding", kind, tw, th, "bound", bound)
    fq = ctx.ncc_search(p1, t, 1, w, 1, h, 1, cq)
    ff = ctx.ncc_search(q1, t, 1, w, 1, h, 1, cf)
    gq = cq.get(np.zeros(ncell, vtgpu.CELL_DTYPE))
    gf = cf.get(np.zeros(ncell, vtgpu.CELL_DTYPE))
    k1, a1 = oracle.pyramid(f1, w, h, 1)
    of, oc = oracle.ncc_search(k1[0], app, tw, th, norm, 1, w, 1, h, 1)
    _check_cells(gq, oc, lambda x, y: oracle.ncc_at(k1[0], app, tw, th, norm, int(x), int(y)))
    if flagged:
        assert gq.tobytes() == gf.tobytes()
    print(f"[ncc] rounding {kind} {tw}x{th}: bound {bound:.3g}, flagged {flagged}, found ({fq.hx},{fq.hy}) "
          f"{fq.ncc:.6f} vs ({of.hx},{of.hy}) {float(of.ncc):.6f}")
    for f in (fq, ff):
        assert (f.hx, f.hy) == (of.hx, of.hy)
        assert abs(f.ncc - float(of.ncc)) <= 1e-5 and abs(f.second - float(of.second)) <= 1e-5
    for x in (cq, cf):
        x.free()
    t.destroy()
    for p in (p0, p1, q1):
        p.destroy()


def test_wrapped_frame_search_equals_pyramid(ctx):
    """vtg_pyr_wrap: a one-level pyramid over the device frame itself. The template cut from it and the exhaustive
    search of it equal those of the built pyramid, record and cell map bit for bit; building it, or wrapping a frame
    whose pitch is not a multiple of 4, is refused."""
    import ctypes as C
    w, h = 640, 480
    f0, f1, _ = vtgpu.frame_pair(w, h, 515)
    d0, d1 = ctx.to_device(np.ascontiguousarray(f0)), ctx.to_device(np.ascontiguousarray(f1))
    p0, p1 = ctx.pyramid(w, h, 3), ctx.pyramid(w, h, 3)
    p0.build(d0, on_device=True)
    p1.build(d1, on_device=True)
    q0, q1 = ctx.pyramid_wrap(d0, w, h), ctx.pyramid_wrap(d1, w, h)
    say("ncc search of a wrapped frame")
    ta, tb = ctx.ncc_template(p0, 0, 200, 150, 64, 64), ctx.ncc_template(q0, 0, 200, 150, 64, 64)
    aa, na = ta.read()
    ab, nb = tb.read()
    assert np.array_equal(aa, ab) and na == nb
    ncell = ((w - 66 - 1) // 16 + 1) * ((h - 66 - 1) // 16 + 1)
    ca, cb = ctx.alloc(12 * ncell), ctx.alloc(12 * ncell)
    fa = ctx.ncc_search(p1, ta, 1, w, 1, h, 1, ca)
    fb = ctx.ncc_search(q1, tb, 1, w, 1, h, 1, cb)
    assert bytes(fa) == bytes(fb) and fa.ncc > 0.9
    assert ca.get(np.zeros(3 * ncell, np.int32)).tobytes() == cb.get(np.zeros(3 * ncell, np.int32)).tobytes()
    with pytest.raises(vtgpu.VtgError, match=r"\(-22\)"):
        q1.build(d1, on_device=True)
    L, out = vtgpu.lib(), C.c_void_p()
    assert L.vtg_pyr_wrap(ctx.h, w, h, d0.ptr, w + 2, C.byref(out)) == -22
    assert L.vtg_pyr_wrap(ctx.h, w, h, d0.ptr, w - 4, C.byref(out)) == -22
    for t in (ta, tb):
        t.destroy()
    for p in (p0, p1, q0, q1):
        p.destroy()
    for x in (ca, cb, d0, d1):
        x.free()
