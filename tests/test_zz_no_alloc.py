"""vtg_ncc_search_fixed, the NCC search that allocates nothing (its working memory is a fixed device array), against
vtg_ncc_search_async through the C-ABI: two calls in a row, the second over a smaller window, give the same found
record and cell map bit for bit; a search larger than the fixed array is refused with VTG_ERANGE."""
import ctypes as C

import numpy as np
import pytest

import vtgpu
from gpu_util import say

pytestmark = pytest.mark.gpu


def _found(ctx, buf):
    f = vtgpu.Found()
    ctx.check(vtgpu.lib().vtg_download(ctx.h, C.byref(f), buf.ptr, C.sizeof(f)))
    return f


def test_fixed_search_equals_search():
    Lb = vtgpu.lib()
    Lb.vtg_ncc_search_fixed.argtypes = [C.c_void_p] * 3 + [C.c_int] * 5 + [C.c_void_p, C.c_void_p]
    Lb.vtg_ncc_search_fixed.restype = C.c_int
    ctx = vtgpu.Context(0)
    w, h = 1920, 1080
    f0, f1, _ = vtgpu.frame_pair(w, h, 31)
    p0, p1 = ctx.pyramid(w, h, 1), ctx.pyramid(w, h, 1)
    p0.build(f0)
    p1.build(f1)
    t = ctx.ncc_template(p0, 0, 700, 400, 64, 64)
    fs = C.sizeof(vtgpu.Found)
    for k, (x0, x1, y0, y1) in enumerate([(1, w, 1, h), (500, 900, 300, 520)]):
        ncell = ((min(x1, w - 66) - x0) // 16 + 1) * ((min(y1, h - 66) - y0) // 16 + 1)
        ca, cb = ctx.alloc(12 * ncell), ctx.alloc(12 * ncell)
        fa, fb = ctx.alloc(fs), ctx.alloc(fs)
        say("ncc fixed search", k, (x0, x1, y0, y1))
        ctx.ncc_search_async(p1, t, x0, x1, y0, y1, 1, ca, fa)
        ctx.check(Lb.vtg_ncc_search_fixed(ctx.h, p1.h, t.h, x0, x1, y0, y1, 1, cb.ptr, fb.ptr), "ncc_search_fixed")
        ctx.sync()
        a, b = _found(ctx, fa), _found(ctx, fb)
        assert bytes(a) == bytes(b), k
        assert a.cells == ncell and a.ncc > 0.9
        ma = ca.get(np.zeros(3 * ncell, np.int32))
        mb = cb.get(np.zeros(3 * ncell, np.int32))
        assert np.array_equal(ma, mb), k
        # and with the map kept in the fixed array too (no cells buffer)
        ctx.check(Lb.vtg_ncc_search_fixed(ctx.h, p1.h, t.h, x0, x1, y0, y1, 1, None, fb.ptr), "ncc_search_fixed")
        ctx.sync()
        assert bytes(_found(ctx, fb)) == bytes(a), k
        for buf in (ca, cb, fa, fb):
            buf.free()
    # a 3840 x 2160 exhaustive search needs more than the fixed array holds
    g0, g1, _ = vtgpu.frame_pair(3840, 2160, 32)
    q0, q1 = ctx.pyramid(3840, 2160, 1), ctx.pyramid(3840, 2160, 1)
    q0.build(g0)
    q1.build(g1)
    t4 = ctx.ncc_template(q0, 0, 1000, 1000, 64, 64)
    fb = ctx.alloc(fs)
    VTG_ERANGE = -34
    assert Lb.vtg_ncc_search_fixed(ctx.h, q1.h, t4.h, 1, 3840, 1, 2160, 1, None, fb.ptr) == VTG_ERANGE
    fb.free()
    for x in (t, t4):
        x.destroy()
    for p in (p0, p1, q0, q1):
        p.destroy()
    ctx.close()
