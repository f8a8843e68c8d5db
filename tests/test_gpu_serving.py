"""The serving loop: results of a stream served alone equal those served in a mix with other streams (bit for bit),
equal the direct calls on the same pyramids, and a device budget too small for everything at once still serves every
request (in smaller batches) without exceeding it."""
import numpy as np
import pytest

import vtgpu
from gpu_util import say
from workloads import StreamFrames

pytestmark = pytest.mark.gpu

CFG = vtgpu.lk_cfg(15, 15, 10, 0.01, 1e-4, 40.0)


def _serve(srv, streams, frames, nframes, ncc_at=(), burst=True):
    """Submit nframes of every stream (interleaved, all at once when burst) and collect results per stream."""
    ids = [srv.open(sf.w, sf.h, L, len(sf.xy) // 2) for sf, L in zip(frames, streams)]
    res = [[None] * nframes for _ in ids]
    tickets = []
    for i in range(nframes):
        for s, sf in enumerate(frames):
            w, h = sf.w, sf.h
            ncc = (w // 2 - 32, h // 2 - 32) if (s, i) in ncc_at else None
            t = srv.submit(ids[s], sf[i], sf.xy, ncc=ncc)
            if burst:
                tickets.append((s, i, t))
            else:
                res[s][i] = srv.wait(t)
    for s, i, t in tickets:
        res[s][i] = srv.wait(t)
    for sid in ids:
        srv.close(sid)
    return res


def test_alone_equals_mix_equals_direct(ctx):
    a = StreamFrames(vtgpu, 640, 480, 5, k=6, nx=40, ny=30)
    others = [StreamFrames(vtgpu, 1280, 720, 6, k=6, nx=60, ny=30), StreamFrames(vtgpu, 1920, 1080, 7, k=6,
                                                                                 nx=80, ny=45),
              StreamFrames(vtgpu, 640, 480, 8, k=6, nx=40, ny=30)]
    nf = 6
    say("serving alone")
    srv = vtgpu.Server(1 << 30, lanes=2, max_batch=8, cfg=CFG)
    alone = _serve(srv, [3], [a], nf, ncc_at={(0, 3)}, burst=False)[0]
    srv.destroy()
    say("serving mix")
    srv = vtgpu.Server(1 << 30, lanes=2, max_batch=8, cfg=CFG)
    mix = _serve(srv, [3, 4, 4, 3], [a] + others, nf, ncc_at={(0, 3), (2, 4)}, burst=True)
    st = srv.stats()
    srv.destroy()
    assert st["completed"] == 4 * nf and st["max_batch"] > 1, st
    for i in range(nf):
        assert alone[i][0].tobytes() == mix[0][i][0].tobytes(), i
        assert bytes(alone[i][1]) == bytes(mix[0][i][1]), i
    assert len(alone[0][0]) == 0 and alone[0][1].ncc == -2.0  # the first frame only builds its pyramid
    assert alone[3][1].ncc > 0.9 and mix[2][4][1].ncc > 0.9     # the searches ran
    # the direct calls on the same frames
    L = 3
    for i in range(1, nf):
        p0, p1 = ctx.pyramid(a.w, a.h, L), ctx.pyramid(a.w, a.h, L)
        p0.build(a[i - 1])
        p1.build(a[i])
        d = ctx.lk_track_host(p0, p1, L - 1, CFG, a.xy[0::2], a.xy[1::2])
        assert d.tobytes() == alone[i][0].tobytes(), i
        if i == 3:
            t = ctx.ncc_template(p0, 0, a.w // 2 - 32, a.h // 2 - 32, 64, 64)
            f = ctx.ncc_search(p1, t, 1, a.w, 1, a.h, 1)
            assert bytes(f) == bytes(alone[3][1])
            t.destroy()
        p0.destroy()
        p1.destroy()


def _r(n, a):
    return (n + a - 1) // a * a


def _resident(w, h, levels):
    """Device bytes a stream keeps: two Q6 pyramids (u8 level 0, u16 above; rows of 256 B, levels of 4 KiB, a
    256 B descriptor) and a 64 x 64 template."""
    b, lw, lh = 0, w, h
    for lv in range(levels):
        b += _r(_r(lw * (1 if lv == 0 else 2), 256) * lh, 4096)
        lw, lh = (lw + 1) // 2, (lh + 1) // 2
    return 2 * _r(b + 256, 256) + _r(256 + 2 * 64 * 64 * 4 + 4 * 64 * 17 * 4, 256)


def test_tight_budget_admission():
    a = StreamFrames(vtgpu, 640, 480, 9, k=4, nx=40, ny=30)
    b = StreamFrames(vtgpu, 1280, 720, 10, k=4, nx=40, ny=30)
    # room for both streams' resident pyramids and one 1280 x 720 frame in flight, not both frames of a round
    budget = _resident(640, 480, 3) + _resident(1280, 720, 4) + 1280 * 720 + (192 << 10)
    srv = vtgpu.Server(budget, lanes=2, max_batch=8, cfg=CFG)
    say("serving tight budget", budget)
    res = _serve(srv, [3, 4], [a, b], 4, burst=True)
    st = srv.stats()
    assert st["completed"] == 8 and st["hbm_peak"] <= budget and st["max_batch"] == 1, st
    # the same frames with room to spare give the same records
    srv2 = vtgpu.Server(1 << 30, lanes=2, max_batch=8, cfg=CFG)
    ref = _serve(srv2, [3, 4], [a, b], 4, burst=True)
    srv2.destroy()
    for s in range(2):
        for i in range(4):
            assert res[s][i][0].tobytes() == ref[s][i][0].tobytes()
    # a stream whose resident pyramids do not fit is refused
    with pytest.raises(vtgpu.VtgError):
        srv.open(3840, 2160, 5, 100)
    srv.destroy()


def test_budget_too_small_for_any_frame():
    """A stream whose resident pyramids fit the budget but whose frames never could (what is left is smaller than one
    frame and its points) is refused at open with VTG_ENOMEM instead of queueing frames that never run; a stream that
    fits is served as usual on the same server."""
    srv = vtgpu.Server(8 << 20, lanes=1, max_batch=4, cfg=CFG)
    say("serving budget too small")
    with pytest.raises(vtgpu.VtgError) as e:
        srv.open(1920, 1080, 4, 16200)   # 2 x 3.4 MB resident, then 2.9 MB per frame: 8 MB cannot hold both
    assert "(-12)" in str(e.value)
    sf = StreamFrames(vtgpu, 320, 240, 9, k=3, nx=10, ny=8)
    sid = srv.open(320, 240, 3, len(sf.xy) // 2)
    outs = [srv.wait(srv.submit(sid, sf[i], sf.xy)) for i in range(3)]
    assert len(outs[0][0]) == 0 and len(outs[2][0]) == len(sf.xy) // 2
    srv.close(sid)
    srv.destroy()


def test_config5_full_shape(ctx, oracle):
    """The serving workload at its full shape over a bounded run: 8 streams (3 x 1920x1080 with 16,200 points, 3 x
    1280x720 with 8,160, 2 x 640x480 with 4,080; 21 x 21 LK; a random walk of 1.5 px per frame), 10 frames each, an
    NCC search on 2 of the 80 frames (2.5%), a 2 GiB budget, continuous batching (lanes 2, batches up to 16). Every
    stream's records served in the mix equal the same stream served alone, bit for bit; one frame of every resolution
    class matches the oracle (every LK record equal to the lane tier's), and both searches match the oracle's (same
    pick, score and second peak within 1e-5)."""
    from gpu_util import lane_ref
    from oracle_ffi import lk_cfg
    from test_gpu_lk import _assert_same
    from workloads import STREAMS
    cfg = vtgpu.lk_cfg(21, 21, 10, 0.01, 1e-4, 40.0)
    nf = 10
    frames = [StreamFrames(vtgpu, w, h, 200 + i, k=6, nx=nx, ny=ny) for i, (w, h, _, _, nx, ny) in enumerate(STREAMS)]
    levels = [L for (_, _, _, L, _, _) in STREAMS]
    ncc_at = {(0, 5), (6, 7)}
    say("config 5: mix of 8 streams")
    srv = vtgpu.Server(2 << 30, lanes=2, max_batch=16, cfg=cfg)
    mix = _serve(srv, levels, frames, nf, ncc_at=ncc_at, burst=True)
    st = srv.stats()
    srv.destroy()
    assert st["completed"] == 8 * nf and st["max_batch"] > 1 and st["hbm_peak"] <= 2 << 30, st
    for s in range(8):
        say("config 5: stream alone", s)
        srv = vtgpu.Server(2 << 30, lanes=2, max_batch=16, cfg=cfg)
        alone = _serve(srv, [levels[s]], [frames[s]], nf, ncc_at={(0, i) for (t, i) in ncc_at if t == s},
                       burst=False)[0]
        srv.destroy()
        for i in range(nf):
            assert alone[i][0].tobytes() == mix[s][i][0].tobytes(), (s, i)
            assert bytes(alone[i][1]) == bytes(mix[s][i][1]), (s, i)
    cfgo = lk_cfg(21, 21, 10, 0.01, 1e-4, 40.0)
    for s, i in ((0, 4), (3, 6), (6, 3)):
        sf, L = frames[s], levels[s]
        say("config 5: oracle frame", s, i)
        k0, I = oracle.pyramid(sf[i - 1], sf.w, sf.h, L, threads=16)
        k1, J = oracle.pyramid(sf[i], sf.w, sf.h, L, threads=16)
        ref = lane_ref(oracle, I, J, L - 1, cfgo, sf.xy[0::2], sf.xy[1::2], threads=16, label=f"stream {s} frame {i}")
        _assert_same(mix[s][i][0], ref, f"config 5 stream {s} frame {i}")
    for s, i in sorted(ncc_at):
        sf, L = frames[s], levels[s]
        say("config 5: oracle search", s, i)
        k0, I = oracle.pyramid(sf[i - 1], sf.w, sf.h, 1, threads=16)
        k1, J = oracle.pyramid(sf[i], sf.w, sf.h, 1, threads=16)
        app, norm = oracle.ncc_template(k0[0], sf.w // 2 - 32, sf.h // 2 - 32, 64, 64)
        of, _ = oracle.ncc_search(k1[0], app, 64, 64, norm, 1, sf.w, 1, sf.h, 1, threads=16)
        f = mix[s][i][1]
        assert (f.hx, f.hy) == (of.hx, of.hy), (s, i)
        assert abs(f.ncc - float(of.ncc)) <= 1e-5 and abs(f.second - float(of.second)) <= 1e-5, (s, i)
        assert f.ncc > 0.8   # the thing was found (noise and a sub-pixel walk keep it under 1)
