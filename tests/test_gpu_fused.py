"""The fused pyramid + LK launch (vtg_pyr_lk_fused: both pyramids of every pair and all pairs' points in one
persistent launch, items taken in dependency order) against the separate launches (vtg_pyr_build_many, then
vtg_lk_track_many), record for record and bit for bit: uneven sizes and point counts, a pair without points, top levels
below the pyramid's last, the flow-back check, graph replays; and the error word stays clear."""
import numpy as np
import pytest

import vtgpu
from gpu_util import say
from workloads import Batch, batch_spec, enqueue_fused

pytestmark = pytest.mark.gpu


def _records(ctx, b):
    ctx.sync()
    return b.out.get(np.zeros(b.npoints, vtgpu.LK_OUT_DTYPE)).copy()


def _check_pyramids(a, b):
    for (p0, p1), (q0, q1) in zip(a.pyrs, b.pyrs):
        for P, Q in ((p0, q0), (p1, q1)):
            for lv in range(P.levels):
                assert np.array_equal(P.read_level(lv).view(), Q.read_level(lv).view()), lv


@pytest.mark.parametrize("fb", [0.0, 0.5])
def test_fused_equals_separate(ctx, fb):
    spec = batch_spec(77, npairs=12, nmin=300, nmax=3000, sizes=[(640, 480), (1280, 720), (1920, 1080)])
    spec[3] = (spec[3][0], spec[3][1], spec[3][2], 0, spec[3][4])   # a pair without points
    a, b = Batch(vtgpu, ctx, spec), Batch(vtgpu, ctx, spec)
    a.tops = b.tops = [max(0, t - (k % 3)) for k, t in enumerate(a.tops)]   # tops 0 .. levels - 1
    cfg = vtgpu.LkCfg(21, 21, 10, 0.01, 1e-4, 40.0, 0)
    say("fused vs separate", fb)
    a.ctx.build_many([p for pr in a.pyrs for p in pr], [f for fr in a.frames for f in fr])
    a.ctx.lk_track_many([p[0] for p in a.pyrs], [p[1] for p in a.pyrs], a.tops, a.offsets, cfg, a.xy, out=a.out,
                        fb_max=fb)
    enqueue_fused(b, cfg, fb_max=fb)
    ra, rb = _records(ctx, a), _records(ctx, b)
    assert ctx.device_error() == 0
    assert ra.tobytes() == rb.tobytes()
    _check_pyramids(a, b)
    print(f"[fused] {b.npoints} points, {len(spec)} pairs, status 0: {(rb['status'] == 0).mean():.3f}")
    a.destroy()
    b.destroy()


def test_fused_graph_replays(ctx):
    """Captured once, replayed 3 times with eager work and syncs between: every replay gives the eager records."""
    spec = batch_spec(78, npairs=6, nmin=500, nmax=2000, sizes=[(640, 480), (1280, 720)])
    b = Batch(vtgpu, ctx, spec)
    cfg = vtgpu.LkCfg(15, 15, 10, 0.01, 1e-4, 40.0, 0)
    say("fused graph")
    enqueue_fused(b, cfg)
    ref = _records(ctx, b)
    ctx.graph_begin()
    enqueue_fused(b, cfg)
    g = ctx.graph_end()
    for i in range(3):
        b.out.put(np.zeros(b.npoints, vtgpu.LK_OUT_DTYPE))
        ctx.graph_launch(g)
        got = _records(ctx, b)
        assert got.tobytes() == ref.tobytes(), i
        enqueue_fused(b, cfg)   # eager work between replays
        ctx.sync()
    ctx.graph_destroy(g)
    assert ctx.device_error() == 0
    b.destroy()
