"""Still-cell vectors on the GPU (vtg_still_*, src/kern/still.c) against the oracle's restatement (itself pinned to the
reference's still.c in test_oracle_still.py), phase by phase over runs of frames with noise, a walking square and a
jump in brightness:
  vectors  every cell bit for bit: u, v, state (LK's status), the mean difference and sigma, against the oracle with
           its lane64 LK (the reference's fp32 lane arithmetic under the fp64 driver the GPU runs); the x87 lane tier
           gives the same cells but for at most 1 in 1,000 where the two drivers round apart;
  decide   given the same vectors (the GPU's, fed to the oracle too): every cell's state, count and the motion summary
           bit for bit;
  update   the new mean bit for bit; and the still_hash over cells and mean equal.
Frames are a level of a device pyramid (level 0 and 1, ragged sizes) or a padded host plane; the one-call form equals
the four phases."""
import numpy as np
import pytest

import vtgpu
from gpu_util import say
from oracle_ffi import OracleStill, StillCfg
from test_oracle_still import q6_plane, still_frames

pytestmark = pytest.mark.gpu


def _level_plane(oracle, img, w, h, lv):
    planes, _ = oracle.pyramid(img, w, h, lv + 1)
    return planes[lv]


def _compare_vectors(gc, oc, k):
    """Every cell equal bit for bit: the per-cell LK is the reference's lane arithmetic under an fp64 driver, and the
    differences and sigmas follow the reference's x87 rounding exactly."""
    for f in ("u", "v", "diff", "sigma"):
        bad = np.nonzero(gc[f].view(np.uint32) != oc[f].view(np.uint32))[0]
        assert len(bad) == 0, f"frame {k}: {f} differs on {len(bad)} cells, first {bad[:8]}"
    bad = np.nonzero(gc["state"] != oc["state"])[0]
    assert len(bad) == 0, f"frame {k}: state differs on {len(bad)} cells, first {bad[:8]}"
    return 0, 0.0


def run_phases(ctx, oracle, W, H, lv, cfg_t, nframes, seed, jump, storage=vtgpu.STORE_Q6, use_plane=False):
    w, h = W, H
    for _ in range(lv):
        w, h = (w + 1) // 2, (h + 1) // 2
    imgs = still_frames(oracle, W, H, nframes, seed, jump=jump)
    g = ctx.still(w, h, *cfg_t)
    o = OracleStill(oracle, StillCfg(w, h, *cfg_t, 1))
    assert (g.cols, g.rows) == (o.s.cols, o.s.rows)
    pyr = ctx.pyramid(W, H, lv + 1, storage)
    worst, flips, moved, lane_apart = 0.0, 0, 0, 0
    for k, img in enumerate(imgs):
        cur = _level_plane(oracle, img, W, H, lv)
        pyr.build(np.ascontiguousarray(img))
        if k:
            if use_plane:
                g.vectors(plane=cur)
            else:
                g.vectors(pyr, lv)
            o.s.lk_tier = 0
            o.vectors(cur)
            c87 = o.cells.copy()
            o.s.lk_tier = 2
            o.vectors(cur)
            nd = int((c87.view(np.uint32).reshape(len(c87), -1) !=
                      o.cells.view(np.uint32).reshape(len(c87), -1)).any(1).sum())
            assert nd <= max(1, len(c87) // 1000), f"frame {k}: lane64 vs x87 lane tier differ on {nd} cells"
            lane_apart += nd
            f, d = _compare_vectors(g.cells(), o.cells.copy(), k)
            flips += f
            worst = max(worst, d)
            o.cells[:] = g.cells()  # the decision from the same vectors
        g.decide()
        o.decide()
        gc = g.cells()
        assert gc.tobytes() == o.cells.tobytes(), f"frame {k}: decided cells"
        gm, om = g.motion(), o.motion()
        assert gm == om, f"frame {k}: motion {gm} vs {om}"
        if use_plane:
            g.update(plane=cur)
        else:
            g.update(pyr, lv)
        o.update(cur)
        g.finish()
        o.finish()
        assert np.array_equal(g.mean().view().view(np.uint32), o.mean().view(np.uint32)), f"frame {k}: mean"
        assert g.hash() == o.hash(), f"frame {k}: hash"
        moved += gm["moving"]
    say(f"still {w}x{h} lv{lv}: {g.ncells} cells, {nframes} frames, status flips {flips}, max |d| {worst:.2e} px, "
        f"moving cells summed {moved}, cells where the x87 and fp64 drivers round apart {lane_apart}")
    assert moved > 0
    g.destroy()
    o.free()
    pyr.destroy()


@pytest.mark.parametrize("W,H,lv,cfg_t,jump", [
    (320, 240, 0, (16, 0.5, 4.0), 5),
    (301, 203, 0, (4, 0.25, 2.0), 12),      # ragged cells at both edges, everything moves at frame 6
    (642, 481, 1, (8, 0.5, 3.0), 12),       # level 1 (u16 Q6) of a ragged frame
])
def test_still_phases_match_oracle(ctx, oracle, W, H, lv, cfg_t, jump):
    run_phases(ctx, oracle, W, H, lv, cfg_t, 9, 21, jump)


def test_still_1080p_level1(ctx, oracle):
    """The reference's live load: 1080p frames read at level 1 (960 x 540: 60 x 68 = 4,080 cells), 5 frames."""
    run_phases(ctx, oracle, 1920, 1080, 1, (16, 0.5, 4.0), 5, 33, 12)


def test_still_host_planes(ctx, oracle):
    """The _plane forms (padded host planes in) give the same phases."""
    run_phases(ctx, oracle, 320, 240, 0, (16, 0.5, 4.0), 6, 44, 12, use_plane=True)


def test_still_frame_equals_phases(ctx, oracle):
    """vtg_still_frame (one call) equals vectors / decide / update / finish called one by one, frame after frame; the
    same camera fed from an F32 pyramid gives the same hashes as from the Q6 one."""
    W, H = 256, 192
    imgs = still_frames(oracle, W, H, 7, 55, jump=12)
    a, b, c = ctx.still(W, H), ctx.still(W, H), ctx.still(W, H)
    pq, pf = ctx.pyramid(W, H, 1, vtgpu.STORE_Q6), ctx.pyramid(W, H, 1, vtgpu.STORE_F32)
    for k, img in enumerate(imgs):
        pq.build(np.ascontiguousarray(img))
        pf.build(np.ascontiguousarray(img))
        a.frame(pq, 0)
        if not b.first():
            b.vectors(pq, 0)
        b.decide()
        b.update(pq, 0)
        b.finish()
        c.frame(pf, 0)
        assert a.hash() == b.hash() == c.hash(), f"frame {k}"
        assert a.motion() == b.motion()
    for s in (a, b, c):
        s.destroy()
    pq.destroy()
    pf.destroy()


def test_still_closed_forms(ctx, oracle):
    """Identical frames: every vector and difference is exactly 0 and the counts reach avg; the mean is the frame; a
    frame one grey level brighter everywhere: every difference is exactly 1.0."""
    W, H = 160, 96
    img = still_frames(oracle, W, H, 1, 9, square=False)[0]
    brighter = np.clip(img.astype(np.int32) + 1, 0, 255).astype(np.uint8)
    assert (img < 255).all()
    s = ctx.still(W, H, 3, 0.5, 4.0)
    p = ctx.pyramid(W, H, 1)
    p.build(img)
    for k in range(4):
        if k:
            s.vectors(p, 0)
            c = s.cells()
            assert (c["u"] == 0).all() and (c["v"] == 0).all() and (c["diff"] == 0).all()
        s.decide()
        s.update(p, 0)
        s.finish()
        assert np.array_equal(s.mean().view(), q6_plane(oracle, img).view())
    assert ((s.cells()["state"] >> 16) == 3).all() and s.motion()["moving"] == 0
    p.build(brighter)
    s.vectors(p, 0)
    assert (s.cells()["diff"] == 1.0).all()
    s.destroy()
    p.destroy()


def test_still_rejects_bad_sizes(ctx):
    """A camera smaller than a patch and its margin, or a level of another size, is VTG_EINVAL."""
    with pytest.raises(vtgpu.VtgError):
        ctx.still(16, 100)
    s = ctx.still(64, 64)
    p = ctx.pyramid(128, 128, 2)
    with pytest.raises(vtgpu.VtgError):
        s.vectors(p, 0)
    s.vectors(p, 1)   # level 1 is 64 x 64
    s.destroy()
    p.destroy()


def test_still_frame_many_equals_single(ctx, oracle):
    """Five cameras of different sizes and levels through vtg_still_frame_many (one LK launch for all cells) end every
    frame exactly as five cameras through vtg_still_frame one by one: the same still_hash and motion."""
    sizes = [(320, 240, 0), (301, 203, 0), (642, 481, 1), (256, 192, 0), (400, 300, 1)]
    frames = [still_frames(oracle, W, H, 6, 90 + i, jump=12) for i, (W, H, _) in enumerate(sizes)]
    pyrs = [ctx.pyramid(W, H, lv + 1) for (W, H, lv) in sizes]

    def lvl(W, H, lv):
        for _ in range(lv):
            W, H = (W + 1) // 2, (H + 1) // 2
        return W, H

    many = [ctx.still(*lvl(W, H, lv), 8, 0.5, 3.0) for (W, H, lv) in sizes]
    one = [ctx.still(*lvl(W, H, lv), 8, 0.5, 3.0) for (W, H, lv) in sizes]
    for k in range(6):
        for i, p in enumerate(pyrs):
            p.build(np.ascontiguousarray(frames[i][k]))
        ctx.still_frame_many(many, pyrs, [lv for (_, _, lv) in sizes])
        for i, p in enumerate(pyrs):
            one[i].frame(p, sizes[i][2])
        for i in range(len(sizes)):
            assert many[i].hash() == one[i].hash(), (k, i)
            assert many[i].motion() == one[i].motion()
    for s in many + one:
        s.destroy()
    for p in pyrs:
        p.destroy()
