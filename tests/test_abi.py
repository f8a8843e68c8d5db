"""The C-ABI library without a GPU: it loads, exports every entry point include/*.h declares, and the synthetic
generator (plain C) is deterministic and exact about its motion."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import vtgpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header, prefix):
    text = open(os.path.join(REPO, "include", header)).read()
    return sorted(set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", text)))


def test_vtgpu_exports_every_declared_symbol():
    lib = C.CDLL(vtgpu.LIB_GPU)
    names = _declared("vtgpu.h", "vtg_")
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(vtgpu.EXPORTS) == names


def test_vtsynth_exports_every_declared_symbol():
    lib = C.CDLL(vtgpu.LIB_SYNTH)
    names = _declared("vtsynth.h", "vts_")
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_no_device_is_an_error_not_a_crash():
    """Without a visible GPU, creating a context returns a negative code."""
    if os.path.exists("/dev/kfd") and os.environ.get("HIP_VISIBLE_DEVICES", "x") != "":
        pytest.skip("a GPU may be present")
    L = vtgpu.lib()
    h = C.c_void_p()
    assert L.vtg_ctx_create(0, C.byref(h)) < 0


def test_synth_deterministic():
    a0, a1, m = vtgpu.frame_pair(320, 200, 7)
    b0, b1, n = vtgpu.frame_pair(320, 200, 7, threads=3)
    assert np.array_equal(a0, b0) and np.array_equal(a1, b1)
    assert (m.dx, m.dy, m.theta_deg) == (n.dx, n.dy, n.theta_deg)
    assert -8 <= m.dx <= 8 and -8 <= m.dy <= 8 and -2 <= m.theta_deg <= 2
    c0, _, _ = vtgpu.frame_pair(320, 200, 8)
    assert not np.array_equal(a0, c0)
    assert a0.max() > 60 and a0.std() > 10


def test_synth_motion_is_rigid():
    """A noiseless pure translation by whole pixels moves frame 0 exactly; vts_truth is the rigid map."""
    m = vtgpu.Motion(3.0, -2.0, 0.0)
    f0, f1, _ = vtgpu.frame_pair(200, 160, 9, motion=m, noise=0.0)
    assert np.array_equal(f1[2 + 10:150, 10:190], f0[10:148 - 0, 7:187])
    r = vtgpu.Motion(1.5, -0.5, 2.0)
    u = vtgpu.truth(r, 200, 160, [99.5, 10.0], [79.5, 20.0])
    assert np.allclose(u[0], [1.5, -0.5])
    th = np.deg2rad(2.0)
    px = np.cos(th) * (10 - 99.5) - np.sin(th) * (20 - 79.5) + 99.5 + 1.5
    py = np.sin(th) * (10 - 99.5) + np.cos(th) * (20 - 79.5) + 79.5 - 0.5
    assert np.allclose(u[1], [px - 10, py - 20])


def test_grid_points():
    xs, ys = vtgpu.grid(180, 90, 1920, 1080)
    assert len(xs) == 16200 and xs[0] == np.float32(1920 / 360) and ys[-1] == np.float32(1080 - 6)
    px, py = vtgpu.points(3, 1000, 640, 480, 20)
    assert px.min() >= 20 and px.max() <= 619 and py.min() >= 20 and py.max() <= 459


def test_vtg_track_exports_every_declared_symbol():
    lib = C.CDLL(vtgpu.LIB_GPU)
    names = _declared("vtg_track.h", "vtg_tk_")
    assert len(names) >= 12
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_headers_compile_as_c(tmp_path):
    """A C program includes the public headers and takes every tracker entry point's address (gcc, no link)."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("no gcc")
    names = _declared("vtg_track.h", "vtg_tk_")
    src = tmp_path / "use.c"
    src.write_text('#include "vtg_track.h"\n#include "vtsynth.h"\nvoid *table[] = {' +
                   ", ".join(f"(void *){n}" for n in names) + "};\n")
    r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-c", "-o", str(tmp_path / "use.o"),
                        "-I" + os.path.join(REPO, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
