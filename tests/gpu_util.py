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
