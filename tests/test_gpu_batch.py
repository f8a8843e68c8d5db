"""Batched and captured forms give the eager single-pair results: pyramids of several pairs built together, LK of all
pairs in one launch, searches on a second context, the whole step captured as a HIP graph and replayed several times
with other eager work between replays; every replay is checked against the eager result."""
import ctypes as C

import numpy as np
import pytest

import vtgpu
from gpu_util import lane_ref, say

pytestmark = pytest.mark.gpu


def _step_objects(ctx, sctx, sizes, seed):
    frames, pyrs, xys, offs, tm = [], [], [], [0], []
    for i, (w, h, L, n) in enumerate(sizes):
        f0, f1, _ = vtgpu.frame_pair(w, h, seed + i)
        frames.append((ctx.to_device(f0), ctx.to_device(f1)))
        pyrs.append((ctx.pyramid(w, h, L), ctx.pyramid(w, h, L)))
        xs, ys = vtgpu.points(seed + i, n, w, h, 8)
        xys.append(np.stack([xs, ys], 1).ravel())
        offs.append(offs[-1] + n)
    return frames, pyrs, np.concatenate(xys).astype(np.float32), offs


def test_graph_replays_match_eager(ctx):
    sctx = vtgpu.Context(0)
    sizes = [(640, 480, 3, 300), (1280, 720, 4, 500), (1920, 1080, 4, 700)]
    frames, pyrs, xy, offs = _step_objects(ctx, sctx, sizes, 300)
    cfg = vtgpu.lk_cfg(21, 21, 10, 0.01, 1e-4, 40.0)
    dxy = ctx.to_device(xy)
    out = ctx.alloc(32 * offs[-1])
    fs = C.sizeof(vtgpu.Found)
    found = sctx.alloc(fs * len(sizes))
    # the header's bound for corners [1, w] x [1, h]: ((x1 - x0) / 16 + 1) * ((y1 - y0) / 16 + 1) records
    cells = [sctx.alloc(12 * (((w - 1) // 16 + 1) * ((h - 1) // 16 + 1))) for w, h, _, _ in sizes]
    tm = []
    L = vtgpu.lib()

    def enqueue():
        ctx.build_many([p for pr in pyrs for p in pr], [f for fr in frames for f in fr])
        sctx.wait(ctx)
        ctx.lk_track_many([p[0] for p in pyrs], [p[1] for p in pyrs], [s[2] - 1 for s in sizes], offs, cfg, dxy,
                          fb_max=0.5, out=out)
        for k, (w, h, _, _) in enumerate(sizes):
            if len(tm) <= k:
                tm.append(sctx.ncc_template(pyrs[k][0], 0, w // 3, h // 3, 32, 32))
            else:
                tm[k].update(pyrs[k][0], 0, w // 3, h // 3)
        # the first two searches in one call, the last alone (both forms in the captured graph)
        sctx.ncc_search_many([p[1] for p in pyrs[:2]], tm[:2], [(1, w, 1, h) for w, h, _, _ in sizes[:2]], 1, found,
                             cells=cells[:2])
        w, h = sizes[2][:2]
        sctx.check(L.vtg_ncc_search_async(sctx.h, pyrs[2][1].h, tm[2].h, 1, w, 1, h, 1, cells[2].ptr,
                                          C.c_void_p(found.ptr.value + fs * 2)))
        ctx.wait(sctx)

    def results():
        ctx.sync()
        o = out.get(np.zeros(offs[-1], vtgpu.LK_OUT_DTYPE))
        f = (vtgpu.Found * len(sizes))()
        ctx.check(L.vtg_download(ctx.h, C.cast(f, C.c_void_p), found.ptr, fs * len(sizes)))
        return o.tobytes(), bytes(f)

    say("batch eager")
    enqueue()
    ref = results()
    # the eager batched result equals one pair at a time
    for k, (w, h, Lv, n) in enumerate(sizes):
        one = ctx.lk_track_host(pyrs[k][0], pyrs[k][1], Lv - 1, cfg, xy[2 * offs[k]:2 * offs[k + 1]:2],
                                xy[2 * offs[k] + 1:2 * offs[k + 1]:2], fb_max=0.5)
        assert one.tobytes() == ref[0][32 * offs[k]:32 * offs[k + 1]]
        f = ctx.ncc_search(pyrs[k][1], tm[k], 1, w, 1, h, 1)
        g = vtgpu.Found.from_buffer_copy(ref[1][fs * k:fs * (k + 1)])
        assert (f.hx, f.hy, f.ncc, f.second, f.ox, f.oy) == (g.hx, g.hy, g.ncc, g.second, g.ox, g.oy)
    say("batch capture")
    ctx.graph_begin()
    enqueue()
    graph = ctx.graph_end()
    other = ctx.pyramid(800, 600, 3)
    for r in range(4):
        say("batch replay", r)
        # eager work in between: another pyramid and an LK call that reuses the output buffer
        f0, _, _ = vtgpu.frame_pair(800, 600, 900 + r)
        other.build(f0)
        ctx.lk_track(other, other, 2, cfg, dxy, None, 0.0, out)
        ctx.sync()
        ctx.graph_launch(graph)
        assert results() == ref, r
    ctx.graph_destroy(graph)
    for t in tm:
        t.destroy()
    other.destroy()
    for p0, p1 in pyrs:
        p0.destroy()
        p1.destroy()
    sctx.close()


def test_uneven_batch(ctx, oracle):
    """Pairs of four frame sizes with 3-5 levels and +-24 px motion, their pyramids built together and all their points
    in one LK launch: every pair's records equal its own single-pair call, and two pairs match the oracle (pyramids bit
    for bit, every LK record equal to the lane tier's). Then the autotuner picks a schedule per size class, and the
    batch under the picks gives the same bytes."""
    import numpy as np
    from oracle_ffi import lk_cfg
    from test_gpu_lk import _assert_same
    from workloads import Batch, batch_spec
    spec = batch_spec(11, npairs=12, nmin=500, nmax=3000)
    sizes = {w for (w, _, _, _, _) in spec}
    assert len(sizes) >= 3 and {L for (_, _, L, _, _) in spec} >= {3, 4}
    b = Batch(vtgpu, ctx, spec)
    cfg = vtgpu.lk_cfg(21, 21, 10, 0.01, 1e-4, 40.0)
    say("uneven batch", len(spec), "pairs", b.npoints, "points")
    b.enqueue(cfg)
    ctx.sync()
    got = b.out.get(np.zeros(b.npoints, vtgpu.LK_OUT_DTYPE))
    for k, (w, h, L, n, _) in enumerate(spec):
        xy = b.xy_host[2 * b.offsets[k]:2 * b.offsets[k + 1]]
        one = ctx.lk_track_host(b.pyrs[k][0], b.pyrs[k][1], L - 1, cfg, xy[0::2], xy[1::2])
        assert one.tobytes() == got[b.offsets[k]:b.offsets[k + 1]].tobytes(), k
    for k in (min(range(len(spec)), key=lambda i: spec[i][0] * spec[i][1]),
              max(range(len(spec)), key=lambda i: spec[i][0] * spec[i][1])):
        w, h, L, n, _ = spec[k]
        say("uneven batch oracle pair", k, w, h, L)
        f0, f1 = b.host[k]
        k0, I = oracle.pyramid(f0, w, h, L, threads=16)
        k1, J = oracle.pyramid(f1, w, h, L, threads=16)
        for lv in range(L):
            assert np.array_equal(b.pyrs[k][1].read_level(lv).view(), k1[lv].view()), (k, lv)
        xy = b.xy_host[2 * b.offsets[k]:2 * b.offsets[k + 1]]
        xs, ys = xy[0::2], xy[1::2]
        ref = lane_ref(oracle, I, J, L - 1, lk_cfg(21, 21, 10, 0.01, 1e-4, 40.0), xs, ys, threads=16)
        _assert_same(got[b.offsets[k]:b.offsets[k + 1]], ref, f"uneven batch pair {k}")
    # the autotuner: one pair per size class, then the whole batch under the picks
    for (w, h, L), k in b.classes().items():
        say("autotune class", w, h, L)
        n = b.offsets[k + 1] - b.offsets[k]
        r = ctx.tune_pair(b.pyrs[k][0], b.pyrs[k][1], b.frames[k][0], b.frames[k][1], L - 1, cfg,
                          b.xy.view(8 * b.offsets[k], 8 * n))
        assert r["candidates"] == 3 + 8 and r["pyr_rows"] in (8, 16, 32) and r["lk_waves"] in (1, 2, 4, 8)
        assert r["lk_ppw"] in (3, 4)
        got_pick = ctx.tune_get(b.pyrs[k][0], L - 1, cfg)
        assert (got_pick["lk_waves"], got_pick["lk_ppw"]) == (r["lk_waves"], r["lk_ppw"])
    b.enqueue(cfg)
    ctx.sync()
    tuned = b.out.get(np.zeros(b.npoints, vtgpu.LK_OUT_DTYPE))
    assert tuned.tobytes() == got.tobytes()
    for k in range(len(spec)):
        w, h, L, _, _ = spec[k]
        f0, f1 = b.host[k]
        if k < 3:
            # the picked tile height builds the same levels
            assert np.array_equal(b.pyrs[k][0].read_level(L - 1).view(),
                                  oracle.pyramid(f0, w, h, L, threads=16)[0][L - 1].view())
    ctx.tune_reset()
    b.destroy()


def test_uneven_batch_full_shape(ctx, oracle):
    """The uneven batch at its full shape: 64 pairs of 640x480 .. 3840x2160 with 500-20,000 points each, +-24 px, 3-5
    levels, all pyramids built together and all points in one LK launch. Every pair's records equal its own
    single-pair call, and one pair of every size class matches the oracle: its pyramids bit for bit and every LK record
    equal to the lane tier's."""
    import numpy as np
    from oracle_ffi import lk_cfg
    from test_gpu_lk import _assert_same
    from workloads import Batch, batch_spec
    spec = batch_spec(12, npairs=64, nmin=500, nmax=20000)
    assert {w for (w, _, _, _, _) in spec} == {640, 1280, 1920, 3840}
    b = Batch(vtgpu, ctx, spec)
    cfg = vtgpu.lk_cfg(21, 21, 10, 0.01, 1e-4, 40.0)
    say("uneven batch full shape", len(spec), "pairs", b.npoints, "points")
    b.enqueue(cfg)
    ctx.sync()
    got = b.out.get(np.zeros(b.npoints, vtgpu.LK_OUT_DTYPE))
    for k, (w, h, L, n, _) in enumerate(spec):
        xy = b.xy_host[2 * b.offsets[k]:2 * b.offsets[k + 1]]
        one = ctx.lk_track_host(b.pyrs[k][0], b.pyrs[k][1], L - 1, cfg, xy[0::2], xy[1::2])
        assert one.tobytes() == got[b.offsets[k]:b.offsets[k + 1]].tobytes(), k
    firsts = {}
    for k, (w, h, L, n, _) in enumerate(spec):
        firsts.setdefault(w, k)
    assert len(firsts) == 4
    for w, k in sorted(firsts.items()):
        w, h, L, n, _ = spec[k]
        say("uneven batch full shape: oracle pair", k, w, h, L, n)
        f0, f1 = b.host[k]
        k0, I = oracle.pyramid(f0, w, h, L, threads=16)
        k1, J = oracle.pyramid(f1, w, h, L, threads=16)
        for lv in range(L):
            assert np.array_equal(b.pyrs[k][0].read_level(lv).view(), k0[lv].view()), (k, lv)
            assert np.array_equal(b.pyrs[k][1].read_level(lv).view(), k1[lv].view()), (k, lv)
        xy = b.xy_host[2 * b.offsets[k]:2 * b.offsets[k + 1]]
        ref = lane_ref(oracle, I, J, L - 1, lk_cfg(21, 21, 10, 0.01, 1e-4, 40.0), xy[0::2], xy[1::2], threads=16,
                       label=f"full batch pair {k}")
        _assert_same(got[b.offsets[k]:b.offsets[k + 1]], ref, f"full batch pair {k} ({w}x{h}, {n} points)")
    b.destroy()


def test_download_async_into_pinned(ctx):
    """vtg_download_async into vtg_host_alloc memory, ordered after work on another context, gives the bytes a waiting
    download gives."""
    other = vtgpu.Context(0)
    rng = np.random.default_rng(5)
    src = rng.integers(0, 256, 3 << 20, dtype=np.uint8)
    d = ctx.to_device(src)
    pin = vtgpu.PinnedArray(ctx, (src.size,), np.uint8)
    say("download_async")
    other.wait(ctx)
    d.get_async(other, pin.arr)
    other.sync()
    assert np.array_equal(pin.arr, src)
    pin.free()
    d.free()
    other.close()


def test_marks(ctx):
    """vtg_ctx_mark / vtg_ctx_mark_wait: marks number 1, 2, ... in order; after waiting for a mark the work before it
    is complete (a pyramid built and LK's records written equal a fully synchronised run) while later work may still
    be queued; marks never recorded are refused; a mark older than the ring of 64 still waits correctly."""
    f0, f1, _ = vtgpu.frame_pair(1280, 720, 41)
    p0, p1 = ctx.pyramid(1280, 720, 4), ctx.pyramid(1280, 720, 4)
    xs, ys = vtgpu.grid(60, 30, 1280, 720)
    cfg = vtgpu.LkCfg(21, 21, 10, 0.01, 1e-4, 40.0, 0)
    say("marks")
    p0.build(f0)
    p1.build(f1)
    ref = ctx.lk_track_host(p0, p1, 3, cfg, xs, ys)
    dxy = ctx.to_device(np.stack([xs, ys], 1).ravel().astype(np.float32))
    out = ctx.alloc(32 * len(xs))
    m0 = ctx.mark()
    ctx.lk_track_many([p0], [p1], [3], [0, len(xs)], cfg, dxy, out=out)
    m1 = ctx.mark()
    assert m1 == m0 + 1
    later = ctx.alloc(32 * len(xs))
    ctx.lk_track_many([p0], [p1], [3], [0, len(xs)], cfg, dxy, out=later)  # later work, maybe still queued
    ctx.mark_wait(m1)
    pin = vtgpu.PinnedArray(ctx, (len(xs),), vtgpu.LK_OUT_DTYPE)
    other = vtgpu.Context(0)
    out.get_async(other, pin.arr)
    other.sync()
    assert pin.arr.tobytes() == ref.tobytes()
    for bad in (0, ctx.mark() + 1):
        with pytest.raises(vtgpu.VtgError, match=r"\(-22\)"):
            ctx.mark_wait(bad)
    first = ctx.mark()
    for _ in range(70):
        ctx.mark()
    ctx.mark_wait(first)
    ctx.sync()
    pin.free()
    for b in (out, later, dxy):
        b.free()
    p0.destroy()
    p1.destroy()
    other.close()


def test_two_contexts_unordered(ctx):
    """The batched step as bench.py issues it: context B cuts the templates and runs the searches on one-level
    pyramids over the device frames, context A builds the pyramids and runs LK, with nothing ordering A and B (they
    share no buffer), steps issued back to back over two groups of frames, and the host waiting for a mark on each
    context before reading a step's results. Every step's LK records and picks equal the same calls made one after
    another on a single context."""
    W_, H_, L, B, T = 640, 480, 3, 2, 32
    groups = []
    for g in range(2):
        fr = []
        for k in range(B):
            f0, f1, _ = vtgpu.frame_pair(W_, H_, 300 + 10 * g + k)
            fr.append((ctx.to_device(f0), ctx.to_device(f1)))
        groups.append(fr)
    xs, ys = vtgpu.grid(40, 30, W_, H_)
    n = len(xs)
    xy = ctx.to_device(np.tile(np.stack([xs, ys], 1).ravel(), B).astype(np.float32))
    cfg = vtgpu.LkCfg(21, 21, 10, 0.01, 1e-4, 40.0, 0)
    corners = [(100 + 37 * k, 80 + 23 * k) for k in range(B)]
    fs = C.sizeof(vtgpu.Found)

    def run(A, S, steps):
        wraps = [[(A.pyramid_wrap(f0, W_, H_), A.pyramid_wrap(f1, W_, H_)) for f0, f1 in fr] for fr in groups]
        pyrs = [(A.pyramid(W_, H_, L, vtgpu.STORE_F32), A.pyramid(W_, H_, L, vtgpu.STORE_F32)) for _ in range(B)]
        tm = [S.ncc_template(wraps[0][k][0], 0, *corners[k], T, T) for k in range(B)]
        S.sync()
        outs = [A.alloc(32 * n * B) for _ in range(steps)]
        founds = [A.alloc(fs * B) for _ in range(steps)]
        marks = []
        for i in range(steps):
            g = i % 2
            S.ncc_templates_update(tm, [w[0] for w in wraps[g]], [0] * B, corners)
            S.ncc_search_many([w[1] for w in wraps[g]], tm, [(1, W_ - T, 1, H_ - T)] * B, 1, founds[i])
            A.build_many([p for pair in pyrs for p in pair], [f for pair in groups[g] for f in pair])
            A.lk_track_many([p[0] for p in pyrs], [p[1] for p in pyrs], [L - 1] * B, [0, n, 2 * n], cfg, xy,
                            out=outs[i])
            marks.append((A.mark(), S.mark()))
        res = []
        for i in range(steps):
            A.mark_wait(marks[i][0])
            S.mark_wait(marks[i][1])
            rec = np.zeros(n * B, vtgpu.LK_OUT_DTYPE)
            outs[i].get(rec)
            fb = np.zeros(fs * B, np.uint8)
            founds[i].get(fb)
            res.append((rec, fb))
        A.sync()
        S.sync()
        for b in outs + founds:
            b.free()
        for t in tm:
            t.destroy()
        for p in [p for pair in pyrs for p in pair] + [p for ws in wraps for pair in ws for p in pair]:
            p.destroy()
        return res

    say("two contexts, unordered, 4 steps over 2 frame groups")
    B2 = vtgpu.Context(0)
    got = run(ctx, B2, 4)
    ref = run(ctx, ctx, 4)
    for i, ((r1, f1), (r2, f2)) in enumerate(zip(got, ref)):
        assert r1.tobytes() == r2.tobytes(), f"step {i}: LK records differ"
        assert f1.tobytes() == f2.tobytes(), f"step {i}: picks differ"
        assert (r1["status"] == 0).mean() > 0.5
    # the two groups' steps differ (the check is not vacuous)
    assert got[0][0].tobytes() != got[1][0].tobytes()
    B2.close()
    xy.free()
    for fr in groups:
        for f0, f1 in fr:
            f0.free()
            f1.free()
