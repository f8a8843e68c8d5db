"""The layered scenes the tracker-level tests run (libvtsynth, vts_scene_*), pinned on the CPU by what their motion
fixes: a translating object's interior at two times a whole-pixel offset apart is the same pixels moved by exactly that
offset while the still background is unchanged, the truth map moves a point with the layer on top of it, and frames are
the same under any thread count."""
import numpy as np
import pytest

import vtgpu

W, H = 640, 480


def test_translate_interior_moves_by_whole_pixels():
    # the object moves (48, 17.6) px/s: from t = 0 to t = 0.625 s exactly (30, 11) px
    f0 = vtgpu.scene_frame("translate", 0.0, W, H)
    f1 = vtgpu.scene_frame("translate", 0.625, W, H)
    cx, cy, hx, hy = vtgpu.scene_object("translate", W, H, 0.0)
    cx1, cy1, _, _ = vtgpu.scene_object("translate", W, H, 0.625)
    assert (round(cx1 - cx, 9), round(cy1 - cy, 9)) == (30.0, 11.0)
    x0, x1 = int(np.ceil(cx - hx)) + 1, int(np.floor(cx + hx)) - 1
    y0, y1 = int(np.ceil(cy - hy)) + 1, int(np.floor(cy + hy)) - 1
    assert np.array_equal(f1[y0 + 11:y1 + 11, x0 + 30:x1 + 30], f0[y0:y1, x0:x1])
    # the background far from both boxes does not move
    assert np.array_equal(f0[:, :int(cx - hx) - 2], f1[:, :int(cx - hx) - 2])
    # a planted mistake: one pixel off is not the same picture
    assert not np.array_equal(f1[y0 + 11:y1 + 11, x0 + 31:x1 + 31], f0[y0:y1, x0:x1])


def test_camera_moves_every_layer_together():
    # the camera offset is (128 sin 0.8t, 48 sin 1.3t): every layer maps by its difference
    t0, t1 = 0.4, 0.9
    du = 128 * (np.sin(0.8 * t1) - np.sin(0.8 * t0))
    dv = 48 * (np.sin(1.3 * t1) - np.sin(1.3 * t0))
    for x, y in ((50.0, 60.0), (320.0, 240.0), (600.0, 400.0)):
        k, mx, my = vtgpu.scene_map("camera", W, H, t0, x, y, t1)
        assert k in (0, 1) and abs(mx - x - du) < 1e-9 and abs(my - y - dv) < 1e-9


def test_map_takes_the_layer_on_top():
    # occlusion: the bar (layer 3) covers the object's upper half when over it, and moves at 88 px/s
    t = 2.9
    ox, oy, _, _ = vtgpu.scene_object("occlusion", W, H, t)
    bar_x = W / 2 - 256 + 88 * t
    k, mx, my = vtgpu.scene_map("occlusion", W, H, t, bar_x, H / 2 - 32, t + 0.1)
    assert k == 3 and abs(mx - bar_x - 8.8) < 1e-9 and abs(my - (H / 2 - 32)) < 1e-9
    k, mx, my = vtgpu.scene_map("occlusion", W, H, t, ox, oy + 40, t + 0.1)  # the object's lower half
    assert k == 1 and abs(mx - ox - 1.28) < 1e-9
    k, mx, my = vtgpu.scene_map("occlusion", W, H, t, 5.0, 5.0, t + 0.1)  # background
    assert (k, mx, my) == (0, 5.0, 5.0)
    # shapes: a point on the disc moves with it
    cx, cy, r, _ = vtgpu.scene_object("shapes", W, H, 1.0)
    cx1, cy1, _, _ = vtgpu.scene_object("shapes", W, H, 1.1)
    k, mx, my = vtgpu.scene_map("shapes", W, H, 1.0, cx + 10, cy - 5, 1.1)
    assert k == 1 and abs(mx - (cx1 + 10)) < 1e-9 and abs(my - (cy1 - 5)) < 1e-9


@pytest.mark.parametrize("scene", vtgpu.SCENES)
def test_frames_deterministic_under_threads(scene):
    a = vtgpu.scene_frame(scene, 1.3, W, H, seed=3, noise_seed=9, threads=1)
    b = vtgpu.scene_frame(scene, 1.3, W, H, seed=3, noise_seed=9, threads=7)
    assert np.array_equal(a, b) and a.std() > 10


def test_brightness_is_the_frame_transform():
    # t where the gain is 1 and the offset 0 exist only at t = 0; there the frame equals the translate-like object scene
    f = vtgpu.scene_frame("brightness", 0.0, W, H)
    g = vtgpu.scene_frame("brightness", 1.0, W, H)
    gain, off = 1 + 0.25 * np.sin(1.3), 16 * np.sin(0.7)
    # the background (still) at t = 1 is the t = 0 background under the transform, to the rounding of both frames
    bg0 = f[:40, :].astype(np.float64)
    bg1 = g[:40, :].astype(np.float64)
    ok = (bg0 > 2) & (bg0 < 190)
    assert np.abs(bg1[ok] - (gain * bg0[ok] + off)).max() <= gain * 0.5 + 0.5 + 1e-9
