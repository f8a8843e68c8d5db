"""Helpers shared by the GPU parity tests."""
import sys

import numpy as np


def say(*a):
    """Name the case before each GPU step, flushed, so a failing run says where it was."""
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def flow_diff(a, b):
    """Per-point |delta| of two lk_out arrays (max over x and y)."""
    return np.maximum(np.abs(a["x"].astype(np.float64) - b["x"]), np.abs(a["y"].astype(np.float64) - b["y"]))


NUDGES = [(1e-3, 0.0), (-1e-3, 0.0), (0.0, 1e-3), (0.0, -1e-3)]


def conditioning(oracle, I, J, top, cfg, xs, ys, gx=None, gy=None, fb=0.0, tol=0.005, threads=16):
    """(well, f64): the FP64 oracle's result, and which points are well-conditioned: moving the start point by
    1e-3 px along either axis, either way, keeps the status and moves the displacement by at most `tol` px
    (amplification <= 5). On the other points the answer depends on rounding far below the 0.01 px parity band
    (a window barely above min_eig, an iteration that runs to max_iter): no arithmetic pins them."""
    xs = np.asarray(xs, np.float32)
    ys = np.asarray(ys, np.float32)
    base = oracle.lk_batch(1, I, J, top, cfg, xs, ys, gx, gy, fb, threads)
    well = np.ones(len(xs), bool)
    for dx, dy in NUDGES:
        px, py = xs + np.float32(dx), ys + np.float32(dy)
        p = oracle.lk_batch(1, I, J, top, cfg, px, py, gx, gy, fb, threads)
        move = np.maximum(np.abs((p["x"].astype(np.float64) - px) - (base["x"].astype(np.float64) - xs)),
                          np.abs((p["y"].astype(np.float64) - py) - (base["y"].astype(np.float64) - ys)))
        well &= (p["status"] == base["status"]) & ((base["status"] != 0) | (move <= tol))
    return well, base
