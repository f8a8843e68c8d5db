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


def conditioning(oracle, I, J, top, cfg, xs, ys, gx=None, gy=None, fb=0.0, tol=0.005, threads=16, tier=1):
    """(well, f64): the FP64 oracle's result (tier 1, or 4 for the Scharr gradient variant), and which points are
    well-conditioned: moving the start point by 1e-3 px along either axis, either way, keeps the status and moves the
    displacement by at most `tol` px (amplification <= 5). On the other points the answer depends on rounding far
    below the 0.01 px parity band (a window barely above min_eig, an iteration that runs to max_iter): no arithmetic
    pins them."""
    xs = np.asarray(xs, np.float32)
    ys = np.asarray(ys, np.float32)
    base = oracle.lk_batch(tier, I, J, top, cfg, xs, ys, gx, gy, fb, threads)
    well = np.ones(len(xs), bool)
    for dx, dy in NUDGES:
        px, py = xs + np.float32(dx), ys + np.float32(dy)
        p = oracle.lk_batch(tier, I, J, top, cfg, px, py, gx, gy, fb, threads)
        move = np.maximum(np.abs((p["x"].astype(np.float64) - px) - (base["x"].astype(np.float64) - xs)),
                          np.abs((p["y"].astype(np.float64) - py) - (base["y"].astype(np.float64) - ys)))
        well &= (p["status"] == base["status"]) & ((base["status"] != 0) | (move <= tol))
    return well, base


def flat_rects(w, h, seed, frac=0.10, n=8):
    """n seeded rectangles (x0, y0, x1, y1), inclusive, together about `frac` of the frame, clear of the border."""
    rng = np.random.default_rng(seed)
    area = frac * w * h / n
    out = []
    for _ in range(n):
        aspect = rng.uniform(0.7, 1.4)
        rw = int(round(np.sqrt(area * aspect)))
        rh = int(round(area / max(rw, 1)))
        x0 = int(rng.integers(40, w - rw - 40))
        y0 = int(rng.integers(40, h - rh - 40))
        out.append((x0, y0, x0 + rw - 1, y0 + rh - 1))
    return out


def in_rects(xs, ys, rects):
    m = np.zeros(len(xs), bool)
    for x0, y0, x1, y1 in rects:
        m |= (xs >= x0) & (xs <= x1) & (ys >= y0) & (ys <= y1)
    return m


def cov_condition(out):
    """cond(G) at level 0 from the covariance s^2 G^-1 the kernel reports (the same condition number); inf where the
    point did not converge."""
    a, b, c = (out["cov_xx"].astype(np.float64), out["cov_xy"].astype(np.float64), out["cov_yy"].astype(np.float64))
    tr, det = a + c, a * c - b * b
    disc = np.sqrt(np.maximum(tr * tr / 4 - det, 0.0))
    hi, lo = tr / 2 + disc, tr / 2 - disc
    with np.errstate(divide="ignore", invalid="ignore"):
        cond = np.where((out["status"] == 0) & (lo > 0), hi / lo, np.inf)
    return cond


def precision_eval(vtgpu, oracle, ctx, w=1920, h=1080, levels=4, nx=180, ny=90, win=21, seed=77, threads=16,
                   grad=0):
    """Pyramid storage x solve mode against the FP64 oracle on one frame pair with about 10% of the grid in near-flat
    patches (texture amplitude 3 grey). Per combination: statuses equal to the oracle's, and the flow error over the
    converged, well-conditioned points, textured and flat apart, plus the share of converged points whose G has a
    condition number above 1e4 (where the mixed solve goes to fp64). grad: the gradient operator (0 Sobel, 1 the
    Scharr variant, against the oracle's FP64 Scharr tier); the context is left on Sobel."""
    rects = flat_rects(w, h, seed)
    f0, f1, _ = vtgpu.frame_pair(w, h, seed, flat_amp=3.0, flat=np.array(rects).ravel())
    xs, ys = vtgpu.grid(nx, ny, w, h)
    flat = in_rects(xs, ys, rects)
    k0, I = oracle.pyramid(f0, w, h, levels, threads=threads)
    k1, J = oracle.pyramid(f1, w, h, levels, threads=threads)
    cfg = vtgpu.LkCfg(win, win, 10, 0.01, 1e-4, 40.0, 0)
    well, f64 = conditioning(oracle, I, J, levels - 1, cfg, xs, ys, threads=threads, tier=4 if grad else 1)
    res = {"points": int(len(xs)), "flat_points": int(flat.sum()), "well_conditioned": int(well.sum()),
           "oracle_converged": int((f64["status"] == 0).sum()), "gradient": "scharr" if grad else "sobel"}
    ctx.check(vtgpu.lib().vtg_lk_set_gradient(ctx.h, grad))
    names = {vtgpu.STORE_Q6: "q6", vtgpu.STORE_F32: "f32", vtgpu.STORE_F16: "f16", vtgpu.STORE_BF16: "bf16"}
    for st, sname in names.items():
        p0, p1 = ctx.pyramid(w, h, levels, st), ctx.pyramid(w, h, levels, st)
        p0.build(f0)
        p1.build(f1)
        for solve, vname in ((vtgpu.SOLVE_F64, "f64"), (vtgpu.SOLVE_MIXED, "mixed")):
            g = ctx.lk_track_host(p0, p1, levels - 1, cfg, xs, ys, solve=solve)
            both = (g["status"] == 0) & (f64["status"] == 0)
            e = flow_diff(g, f64)
            r = {"status_equal": float((g["status"] == f64["status"]).mean()),
                 "within_0.05px_converged": float((e[both] <= 0.05).mean()),
                 "cond_over_1e4": float((cov_condition(g)[g["status"] == 0] > 1e4).mean())}
            for part, m in (("textured", ~flat), ("flat", flat)):
                ee = e[both & well & m]
                r[part] = {"n": int(len(ee)),
                           "p50_px": float(np.median(ee)) if len(ee) else None,
                           "p99_px": float(np.percentile(ee, 99)) if len(ee) else None,
                           "max_px": float(ee.max()) if len(ee) else None}
            res[f"{sname}/{vname}"] = r
        p0.destroy()
        p1.destroy()
    ctx.check(vtgpu.lib().vtg_lk_set_gradient(ctx.h, 0))
    return res


def lane_ref(oracle, I, J, top, cfg, xs, ys, gx=None, gy=None, fb_max=0.0, threads=8, label=""):
    """The records the GPU must equal: the lane64 tier's, after checking that they equal the x87 lane tier's on all but
    at most 1 in 1,000 records (the fp64 and x87 drivers rounding apart)."""
    r64 = oracle.lk_batch(2, I, J, top, cfg, xs, ys, gx, gy, fb_max=fb_max, threads=threads)
    r87 = oracle.lk_batch(0, I, J, top, cfg, xs, ys, gx, gy, fb_max=fb_max, threads=threads)
    a = np.ascontiguousarray(r64).view(np.uint32).reshape(-1, 8)
    b = np.ascontiguousarray(r87).view(np.uint32).reshape(-1, 8)
    nd = int((a != b).any(1).sum())
    print(f"[lk] {label}: lane64 vs x87 lane tier: {nd} of {len(r64)} records differ")
    assert nd * 1000 <= max(len(r64), 1)
    return r64
