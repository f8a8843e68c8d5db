"""The Scharr gradient variant (vtg_lk_set_gradient(VTG_GRAD_SCHARR); no counterpart in the reference, whose LK takes
Sobel gradients, lk_ref.c:30-31): every record equal bit for bit to the oracle's lane64 Scharr tier (the reference's
fp32 lane arithmetic with [3 10 3] / 32 smoothing in place of [1 2 1] / 8, and the fp64 driver) in both lane layouts,
for exact and u8-level storage, with the flow-back check; the fused launch equal to the separate launches; and Sobel
restored exactly. The oracle's Scharr tiers are pinned on the CPU by a closed form and translation truth."""
import numpy as np
import pytest

import vtgpu
from gpu_util import say
from oracle_ffi import lk_cfg
from test_gpu_lk import _assert_same, _gpu_pyr
from workloads import Batch, batch_spec, enqueue_fused

pytestmark = pytest.mark.gpu
LANE64, LANE64_SCHARR = 2, 3


@pytest.mark.parametrize("storage", [vtgpu.STORE_Q6, vtgpu.STORE_F32_U8])
@pytest.mark.parametrize("ppw", [3, 4])
def test_scharr_equals_oracle(ctx, oracle, storage, ppw):
    w, h, L = 640, 480, 3
    f0, f1, _ = vtgpu.frame_pair(w, h, 31)
    xs, ys = vtgpu.points(31, 1500, w, h, 8)
    cfg = lk_cfg(21, 21, 10, 0.01, 1e-4, 40.0)
    kI, I = oracle.pyramid(f0, w, h, L)
    kJ, J = oracle.pyramid(f1, w, h, L)
    ref = oracle.lk_batch(LANE64_SCHARR, I, J, L - 1, cfg, xs, ys, fb_max=0.5)
    sob = oracle.lk_batch(LANE64, I, J, L - 1, cfg, xs, ys, fb_max=0.5)
    assert (ref["x"] != sob["x"]).any()   # the variant changes answers
    Lb = vtgpu.lib()
    PI, PJ = _gpu_pyr(ctx, f0, w, h, L, storage), _gpu_pyr(ctx, f1, w, h, L, storage)
    ctx.check(Lb.vtg_lk_set_ppw(ctx.h, ppw))
    try:
        say("lk scharr storage", storage, "ppw", ppw)
        ctx.check(Lb.vtg_lk_set_gradient(ctx.h, vtgpu.GRAD_SCHARR))
        got = ctx.lk_track_host(PI, PJ, L - 1, cfg, xs, ys, fb_max=0.5)
        ctx.check(Lb.vtg_lk_set_gradient(ctx.h, vtgpu.GRAD_SOBEL))
        back = ctx.lk_track_host(PI, PJ, L - 1, cfg, xs, ys, fb_max=0.5)
    finally:
        ctx.check(Lb.vtg_lk_set_gradient(ctx.h, vtgpu.GRAD_SOBEL))
        ctx.check(Lb.vtg_lk_set_ppw(ctx.h, 0))
    _assert_same(got, ref, f"scharr storage {storage} ppw {ppw}")
    _assert_same(back, sob, f"sobel again, storage {storage} ppw {ppw}")
    assert Lb.vtg_lk_set_gradient(ctx.h, 2) == -22
    PI.destroy()
    PJ.destroy()


def test_scharr_fused_equals_separate(ctx):
    spec = batch_spec(78, npairs=4, nmin=300, nmax=1500, sizes=[(640, 480), (1280, 720)])
    a, b = Batch(vtgpu, ctx, spec), Batch(vtgpu, ctx, spec)
    cfg = vtgpu.LkCfg(21, 21, 10, 0.01, 1e-4, 40.0, 0)
    Lb = vtgpu.lib()
    say("lk scharr fused vs separate")
    ctx.check(Lb.vtg_lk_set_gradient(ctx.h, vtgpu.GRAD_SCHARR))
    try:
        a.ctx.build_many([p for pr in a.pyrs for p in pr], [f for fr in a.frames for f in fr])
        a.ctx.lk_track_many([p[0] for p in a.pyrs], [p[1] for p in a.pyrs], a.tops, a.offsets, cfg, a.xy, out=a.out,
                            fb_max=0.5)
        enqueue_fused(b, cfg, fb_max=0.5)
        ctx.sync()
    finally:
        ctx.check(Lb.vtg_lk_set_gradient(ctx.h, vtgpu.GRAD_SOBEL))
    ra = a.out.get(np.zeros(a.npoints, vtgpu.LK_OUT_DTYPE))
    rb = b.out.get(np.zeros(b.npoints, vtgpu.LK_OUT_DTYPE))
    assert ctx.device_error() == 0
    assert ra.tobytes() == rb.tobytes()
    a.destroy()
    b.destroy()
