#!/usr/bin/env python3
"""bench.py: 1080p frame-pair throughput of the vision path on one MI355X.

The work per frame pair: a 1920 x 1080 u8 pair already in HBM; both 4-level pyramids; sparse pyramidal LK of 16,200
grid points (180 x 90, 21 x 21 window, 10 iterations, eps 0.01 px, min_eig 1e-4); a 64 x 64 template cut from frame 0
searched exhaustively over frame 1 (cell map, top 8, half-pixel refinement); LK records and the search's pick copied to
the host.

Two forms run, on the same inputs:
  plain    the reference's fp32 Q6 planes, one pair at a time, one launch per pyramid level per frame, the NCC map
           in the reference's own fp32 arithmetic (the straightforward port);
  batched  the headline: B pairs per step, exact fp32 Q6 pyramids built with one launch per level for all 2B
           pyramids, one LK launch for all B pairs, the B searches on a second stream overlapping the pyramids and
           LK, each step's launches issued as they are (the same step replayed as captured HIP graphs is measured
           beside it: `other_launch_form`).
A step of the batched form is one batch of B pairs; `value` is pairs per second. Inputs rotate over more than the
256 MB Infinity Cache. Prints one JSON line; details go to stderr.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--graph] [--cpu-seconds S] [--batch-steps N] [--serve-frames F]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "intel-phi-visiontrack_amd"))

METRIC = "LK pts/s + 1080p frame-pairs/s vs plain; p50/p99 ms @ ≤0.05 px err"
W, H, LEVELS, NX, NY, WIN, TPL = 1920, 1080, 4, 180, 90, 21, 64
NPTS = NX * NY
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector, published (dual-issue packed FP32)
PEAK_INT8_MFMA_TOPS = 5000.0  # MI355X dense int8 matrix peak, published (5.0 POP/s)
UNIQ = 8                   # distinct seeded pairs; device copies rotate through more than the Infinity Cache


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_MD = os.path.join(REPO, "profiles", "traffic_1080p_batch8.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE / WRITE_SIZE summary of this workload (the
    counters cannot be read from inside the process), or None when the summary is absent."""
    try:
        text = open(TRAFFIC_MD).read()
        blob = json.loads(text[text.index("```json") + 7:text.rindex("```")])
        return int(blob["kernels"][kernel]["traffic_bytes"])
    except (OSError, ValueError, KeyError):
        return None


def lk_flops(out, xs, ys, top, w, h, win):
    """Algorithmic fp32 flops of one LK batch: per level processed 6(w+2)(h+2) + 26wh (build), per iteration
    3w(h+1) + 8wh, one residual 3w(h+1) + 11wh at level 0 for a converged point."""
    build = 6 * (win + 2) ** 2 + 26 * win * win
    it = 3 * win * (win + 1) + 8 * win * win
    res = 3 * win * (win + 1) + 11 * win * win
    m = win // 2 + 1
    levels = np.zeros(len(xs))
    for L in range(top + 1):
        lw, lh = w, h
        for _ in range(L):
            lw, lh = (lw + 1) // 2, (lh + 1) // 2
        px, py = xs / 2 ** L, ys / 2 ** L
        levels += (px >= m) & (py >= m) & (px < lw - m - 1) & (py < lh - m - 1)
    return float((levels * build).sum() + out["iters"].astype(np.float64).sum() * it
                 + (out["status"] == 0).sum() * res)


def pyr_bytes(w, h, levels, elem):
    """Algorithmic HBM bytes of one pyramid: the u8 frame read once, each upper level written once."""
    b, lw, lh = w * h, w, h
    for _ in range(1, levels):
        lw, lh = (lw + 1) // 2, (lh + 1) // 2
        b += lw * lh * elem
    return b


class Batched:
    """B pairs per step on two contexts (A: pyramids and LK; S: the template cuts and searches, reading the device
    frames in place, beside them), the results of each step downloaded on one of two copy contexts while the next step
    runs (double-buffered pinned host records). Issued launches need no order between A and S: the two share no
    buffer, each runs its steps in its own order, and the host waits for a mark on each before a step's download.
    Captured graphs fork S from A at each step's start and join it at the end."""

    def __init__(self, vtgpu, d_frames, corners, batch, xs, ys, graph, one_stream=False, pyr_sets=0,
                 storage=None, join=False):
        self.v, self.B, self.graph, self.one_stream = vtgpu, batch, graph, one_stream
        # a captured graph's second branch must fork from A and join it again; issued launches need neither
        self.join = graph or join
        self.A, self.S = vtgpu.Context(0), vtgpu.Context(0)
        # the points are a raster grid, already in spatial order: no reordering launch
        self.A.check(self.A.L.vtg_lk_set_order(self.A.h, 0))
        self.copy = [vtgpu.Context(0), vtgpu.Context(0)]
        self.frames, self.corners = d_frames, corners
        self.groups = len(d_frames) // batch
        A = self.A
        # one-level pyramids over the device frames themselves: the template cuts and searches read the frames in
        # place, so they need not wait for the pyramids
        self.wraps = [(A.pyramid_wrap(f0, W, H), A.pyramid_wrap(f1, W, H)) for f0, f1 in d_frames]
        st = vtgpu.STORE_F32 if storage is None else storage
        self.pyrs = [[(A.pyramid(W, H, LEVELS, st), A.pyramid(W, H, LEVELS, st))
                      for _ in range(batch)]
                     for _ in range(pyr_sets or (self.groups if graph else 1))]
        self.xy = A.to_device(np.tile(np.stack([xs, ys], 1).ravel(), batch))
        # each group's own records, so a step's download never races the next step's writes
        self.outs = [A.alloc(32 * NPTS * batch) for _ in range(self.groups)]
        self.cfg = vtgpu.LkCfg(WIN, WIN, 10, 0.01, 1e-4, 40.0, 0)
        self.offs = [k * NPTS for k in range(batch + 1)]
        ncell = ((W - 1) // 16 + 1) * ((H - 1) // 16 + 1)   # the header's bound for corners [1, W] x [1, H]
        self.cells = [self.A.alloc(12 * ncell) for _ in range(batch)]
        self.fs = C.sizeof(vtgpu.Found)
        self.founds = [A.alloc(self.fs * batch) for _ in range(self.groups)]
        self.tmpl = [None] * batch
        self.h_out = [vtgpu.PinnedArray(A, (NPTS * batch,), vtgpu.LK_OUT_DTYPE) for _ in range(2)]
        self.h_found = [vtgpu.PinnedArray(A, (self.fs * batch,), np.uint8) for _ in range(2)]
        self.graphs = {}
        self.marks = {}

    def enqueue(self, g, serial=False):
        """All device work of group g's batch. Context S: the B template cuts from the frames 0 and the B exhaustive
        searches of the frames 1 (cell maps on the matrix cores, top 8, refinement), all reading the device frames in
        place; context A beside it: the 2B pyramids, then LK of all pairs (with join: S forked from A at the start and
        joined to it at the end). serial: everything on A, one kernel after another (for timing each kernel alone)."""
        A, S, B = self.A, self.S, self.B
        serial = serial or self.one_stream
        pyrs = self.pyrs[g % len(self.pyrs)]
        fr = self.frames[g * B:(g + 1) * B]
        wr = self.wraps[g * B:(g + 1) * B]
        N = A if serial else S
        if not serial and self.join:
            S.wait(A)
        corners = [self.corners[(g * B + k) % UNIQ] for k in range(B)]
        if self.tmpl[0] is None:
            self.tmpl = [N.ncc_template(wr[k][0], 0, *corners[k], TPL, TPL) for k in range(B)]
        else:  # the B templates re-cut in one launch
            N.ncc_templates_update(self.tmpl, [w[0] for w in wr], [0] * B, corners)
        # the B exhaustive searches: one launch per stage for all of them
        regions = [(1, W, 1, H)] * B
        N.ncc_search_many([w[1] for w in wr], self.tmpl, regions, 1, self.founds[g], cells=self.cells)
        A.build_many([p for pair in pyrs for p in pair], [f for pair in fr for f in pair])
        A.lk_track_many([p[0] for p in pyrs], [p[1] for p in pyrs], [LEVELS - 1] * B, self.offs, self.cfg, self.xy,
                        out=self.outs[g])
        if not serial and self.join:
            A.wait(S)

    def tune(self):
        """The autotuner on pair 0 of the workload: its picks (reduce tile height, LK schedule) for the 1080p class
        apply to every later build and LK launch of that class on context A."""
        p0, p1 = self.pyrs[0][0]
        f0, f1 = self.frames[0]
        t = self.A.tune_pair(p0, p1, f0, f1, LEVELS - 1, self.cfg, self.xy)
        log(f"autotuner picks for the step's class: {t}")
        return t

    def prepare(self):
        """Grow scratch and templates eagerly once per group, then capture each group's step."""
        for g in range(self.groups):
            self.enqueue(g)
        self.A.sync()
        self.S.sync()
        if self.graph:
            for g in range(self.groups):
                self.A.graph_begin()
                self.enqueue(g)
                self.graphs[g] = self.A.graph_end()

    def launch(self, i, after=None):
        """Step i: its group's work and a mark after it on each context it ran on, no host wait. after: a context
        whose work so far the step's work waits for (the frames' upload)."""
        g = i % self.groups
        if after is not None:
            self.A.wait(after)
            if not self.join:
                self.S.wait(after)
        if self.graph:
            self.A.graph_launch(self.graphs[g])
        else:
            self.enqueue(g)
        self.marks[i] = (self.A.mark(), None if self.join or self.one_stream else self.S.mark())

    def finish(self, i):
        """Wait for step i's work (its mark), then download its results (every LK record, every pick) on copy
        context i % 2 into pinned buffer i % 2 and return them. The host waits for the mark instead of the copy
        context waiting for A on the device: a wait queued on the device holds up everything behind it on its
        hardware queue, which the graphs' own nodes may share (measured 10 % slower)."""
        g = i % self.groups
        ma, ms = self.marks.pop(i)
        self.A.mark_wait(ma)
        if ms is not None:
            self.S.mark_wait(ms)
        cp = self.copy[i % 2]
        self.outs[g].get_async(cp, self.h_out[i % 2].arr)
        self.founds[g].get_async(cp, self.h_found[i % 2].arr)
        cp.sync()
        found = (self.v.Found * self.B).from_buffer_copy(self.h_found[i % 2].arr.tobytes())
        return self.h_out[i % 2].arr, found

    def close(self):
        """Everything the steps own, then the contexts (the device frames are the caller's)."""
        self.A.sync()
        self.S.sync()
        for g in self.graphs.values():
            self.A.graph_destroy(g)
        for p in [p for ps in self.pyrs for pair in ps for p in pair] + [p for pair in self.wraps for p in pair]:
            p.destroy()
        for t in self.tmpl:
            if t is not None:
                t.destroy()
        for b in [self.xy] + self.outs + self.cells + self.founds + self.h_out + self.h_found:
            b.free()
        for c in [self.A, self.S] + self.copy:
            c.close()


def plain_rate(vtgpu, d_frames, corners, xs, ys, steps):
    """The straightforward port: fp32 Q6 planes, one pair at a time, the NCC map in the reference's arithmetic."""
    ctx = vtgpu.Context(0)
    p0, p1 = ctx.pyramid(W, H, LEVELS, vtgpu.STORE_F32), ctx.pyramid(W, H, LEVELS, vtgpu.STORE_F32)
    xy = ctx.to_device(np.stack([xs, ys], 1).ravel())
    out = ctx.alloc(32 * NPTS)
    cfg = vtgpu.LkCfg(WIN, WIN, 10, 0.01, 1e-4, 40.0, 0)
    h_out = np.zeros(NPTS, vtgpu.LK_OUT_DTYPE)
    tmpl = None

    def one(i):
        nonlocal tmpl
        fa, fb = d_frames[i % len(d_frames)]
        p0.build(fa, on_device=True)
        p1.build(fb, on_device=True)
        ctx.lk_track(p0, p1, LEVELS - 1, cfg, xy, None, 0.0, out)
        cx, cy = corners[i % UNIQ]
        if tmpl is None:
            tmpl = ctx.ncc_template(p0, 0, cx, cy, TPL, TPL)
        else:
            tmpl.update(p0, 0, cx, cy)
        ctx.ncc_search(p1, tmpl, 1, W, 1, H, 1)
        out.get(h_out)

    one(0)
    ctx.sync()
    t = time.perf_counter()
    for i in range(steps):
        one(i)
    ctx.sync()
    rate = steps / (time.perf_counter() - t)
    # the same pairs once more with HIP events around every launch: milliseconds per pair by stage
    ctx.timing(True)
    ctx.timing_reset()
    for i in range(steps):
        one(i)
    ctx.sync()
    names = {"pyramids": ("pyr_level0", "pyr_reduce"), "lk": ("lk_track",),
             "ncc": ("ncc_template", "ncc_cells_generic", "ncc_cells", "ncc_top8", "ncc_refine_final")}
    stage = {s: sum(ctx.timing_read(k)[0] for k in ks) / steps for s, ks in names.items()}
    cells_ms = ctx.timing_read("ncc_cells_generic")[0] / steps
    ctx.timing(False)
    npos = (W - TPL - 2) * (H - TPL - 2)
    ncc_tflops = 2.0 * TPL * TPL * npos / (cells_ms * 1e-3) / 1e12 if cells_ms > 0 else None
    for b in (p0, p1):
        b.destroy()
    return rate, {"stage_ms_per_pair": {k: round(v, 4) for k, v in stage.items()},
                  "ncc_map_kernel": {"name": "ncc_cells_generic (the reference's fp32 lane arithmetic)",
                                     "ms_per_search": round(cells_ms, 3),
                                     "achieved_tflops": round(ncc_tflops, 2) if ncc_tflops else None}}


def batching(vtgpu, steps, seed=2024):
    """The uneven batch: 64 pairs (640x480 .. 3840x2160, 500-20,000 points, +-24 px, 3-5 levels), captured as one HIP
    graph and replayed, in three forms: one launch per pyramid level for all 128 pyramids then one LK launch, under the
    default schedules; the same after the autotuner has picked per size class; and the fused form, every pyramid and
    every point in ONE persistent launch (work items in dependency order, a pair's LK overlapping later pyramids).
    Per form: the median and p99 of a whole batch (graph launch to LK records on the host) and the rates."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from workloads import Batch, batch_spec, enqueue_fused
    ctx = vtgpu.Context(0)
    spec = batch_spec(seed)
    b = Batch(vtgpu, ctx, spec, storage=vtgpu.STORE_F32_U8)
    cfg = vtgpu.LkCfg(WIN, WIN, 10, 0.01, 1e-4, 40.0, 0)
    h_out = np.zeros(b.npoints, vtgpu.LK_OUT_DTYPE)

    def timed(label, fused=False):
        enq = (lambda: enqueue_fused(b, cfg)) if fused else (lambda: b.enqueue(cfg))
        enq()  # eager once: scratch and schedules settle
        ctx.sync()
        ctx.graph_begin()
        enq()
        g = ctx.graph_end()
        lat = []
        for i in range(steps + 2):
            t = time.perf_counter()
            ctx.graph_launch(g)
            b.out.get(h_out)
            if i >= 2:
                lat.append(time.perf_counter() - t)
        ctx.graph_destroy(g)
        ctx.timing(True)
        ctx.timing_reset()
        enq()
        ctx.sync()
        kern = {k: ctx.timing_read(k) for k in (("pyr_lk_fused",) if fused else ("pyr_reduce", "lk_track"))}
        ctx.timing(False)
        ms = np.array(lat) * 1e3
        r = {"batch_ms_p50": round(float(np.percentile(ms, 50)), 3),
             "batch_ms_p99": round(float(np.percentile(ms, 99)), 3),
             "pairs_per_s": round(len(spec) / (np.median(ms) * 1e-3), 1),
             "lk_pts_per_s": round(b.npoints / (kern["pyr_lk_fused" if fused else "lk_track"][0] * 1e-3), 1),
             "kernel_ms": {k: round(v[0], 4) for k, v in kern.items()},
             "launches": {k: int(v[1]) for k, v in kern.items()}}
        log(f"batching {label}: {r}")
        return r

    res = {"pairs": len(spec), "points": b.npoints, "megapixels_per_frame_set": round(b.pixels() / 1e6, 2),
           "levels": {str(L): sum(1 for x in spec if x[2] == L) for L in (3, 4, 5)},
           "sizes": {f"{w}x{h}": sum(1 for x in spec if x[0] == w) for (w, h) in
                     ((640, 480), (1280, 720), (1920, 1080), (3840, 2160))},
           "graph": True, "storage": "f32_u8 (fused form: q6, the storage it builds)"}
    res["default"] = timed("default")
    picks = {}
    for (w, h, L), k in sorted(b.classes().items()):
        n = b.offsets[k + 1] - b.offsets[k]
        picks[f"{w}x{h}/L{L}"] = ctx.tune_pair(b.pyrs[k][0], b.pyrs[k][1], b.frames[k][0], b.frames[k][1], L - 1, cfg,
                                               b.xy.view(8 * b.offsets[k], 8 * n))
    res["autotune"] = picks
    res["tuned"] = timed("tuned")
    ctx.tune_reset()
    b.destroy()
    # one persistent launch for all 128 pyramids and all points (items in dependency order); it builds Q6 pyramids
    b = Batch(vtgpu, ctx, spec)
    res["fused"] = timed("fused", fused=True)
    res["fused"]["device_error"] = ctx.device_error()
    res["fused"]["storage"] = "q6"
    b.destroy()
    ctx.close()
    return res


def serving(vtgpu, frames_per_stream, seed=31, budget=2 << 30):
    """The serving loop under 8 streams (3 x 1080p60, 3 x 720p30, 2 x 480p60; grids of 16,200 / 8,160 / 4,080
    points; a random walk of 1.5 px per frame; an NCC search on 2% of frames), first with Poisson arrivals at the
    streams' rates (latency from submit to results on the host, p50 / p99 per class and overall), then saturated: every
    frame queued at once, which gives the server's capacity in frames per second."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from workloads import STREAMS, StreamFrames
    rng = np.random.default_rng(seed)
    frames = [StreamFrames(vtgpu, w, h, 100 + i, k=24, nx=nx, ny=ny) for i, (w, h, _, _, nx, ny) in enumerate(STREAMS)]
    cfg = vtgpu.LkCfg(WIN, WIN, 10, 0.01, 1e-4, 40.0, 0)

    def run(poisson):
        srv = vtgpu.Server(budget, lanes=2, max_batch=16, cfg=cfg)
        ids = [srv.open(w, h, L, nx * ny) for (w, h, _, L, nx, ny) in STREAMS]
        events = []
        for s, (w, h, fps, _, _, _) in enumerate(STREAMS):
            times = np.cumsum(rng.exponential(1.0 / fps, frames_per_stream)) if poisson else np.zeros(frames_per_stream)
            ncc = rng.random(frames_per_stream) < 0.02
            events += [(float(times[i]), s, i, bool(ncc[i])) for i in range(frames_per_stream)]
        events.sort()
        pending, lat = {}, []

        def collect(block=False):
            for tk in list(pending):
                r = srv.wait(tk) if block else srv.poll(tk)
                if r is not None:
                    lat.append((pending.pop(tk), r[2]))

        t0 = time.perf_counter()
        for (t, s, i, ncc) in events:
            while poisson and time.perf_counter() - t0 < t:
                collect()
                time.sleep(0.0002)
            w, h = STREAMS[s][0], STREAMS[s][1]
            tk = srv.submit(ids[s], frames[s][i], frames[s].xy, ncc=(w // 2 - 32, h // 2 - 32) if ncc else None)
            pending[tk] = s
            if poisson:
                collect()
        collect(block=True)
        elapsed = time.perf_counter() - t0
        st = srv.stats()
        srv.destroy()
        ms = np.array([x[1] for x in lat])
        cls = np.array([STREAMS[x[0]][1] for x in lat])
        r = {"frames": len(lat), "seconds": round(elapsed, 3), "frames_per_s": round(len(lat) / elapsed, 1),
             "p50_ms": round(float(np.percentile(ms, 50)), 3), "p99_ms": round(float(np.percentile(ms, 99)), 3),
             "per_class": {f"{h}p": {"p50_ms": round(float(np.percentile(ms[cls == h], 50)), 3),
                                      "p99_ms": round(float(np.percentile(ms[cls == h], 99)), 3)}
                           for h in (1080, 720, 480)},
             "lk_points_per_s": round(st["lk_points"] / elapsed, 1), "batches": st["batches"],
             "mean_batch": round(st["completed"] / max(1, st["batches"]), 2), "max_batch": st["max_batch"],
             "max_batches_in_flight": st["max_inflight"], "hbm_budget": st["hbm_budget"],
             "hbm_resident": st["hbm_resident"], "hbm_peak": st["hbm_peak"]}
        log(f"serving {'poisson' if poisson else 'saturated'}: {r}")
        return r

    return {"streams": "3 x 1920x1080@60, 3 x 1280x720@30, 2 x 640x480@60", "offered_frames_per_s": 390,
            "frames_per_stream": frames_per_stream, "lanes": 2, "max_batch": 16, "ncc_share": 0.02,
            "poisson": run(True), "saturated": run(False)}


def still_bench(vtgpu, nframes, cpu_frames=4):
    """Still-cell vectors (still.c) at the reference's live loads: 1080p frames read at level 1 (960 x 540, 4,080 cells)
    and at level 0 (1920 x 1080, 16,200 cells); the camera jitters by up to 1 px (a pair's two frames alternate). A frame
    is vtg_still_frame (per-cell LK and differences on the GPU, the decision on this thread, the mean update) with the
    stream drained after it: p50 / p99 ms per frame and cells per second. Beside it, the oracle's still phases
    (lane-tier LK on 16 threads) on the same frames, a few frames each."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from oracle_ffi import Oracle, OracleStill, StillCfg
    f0, f1, _ = vtgpu.frame_pair(W, H, 77, max_shift=1.0, max_rot=0.0)
    ctx = vtgpu.Context(0)
    oracle = Oracle()
    out = {}
    for lv in (1, 0):
        w, h = W, H
        for _ in range(lv):
            w, h = (w + 1) // 2, (h + 1) // 2
        pyrs = []
        for f in (f0, f1):
            pyr = ctx.pyramid(W, H, lv + 1)
            pyr.build(f)
            pyrs.append(pyr)
        st = ctx.still(w, h, 16, 0.5, 4.0)
        for i in range(3):
            st.frame(pyrs[i % 2], lv)
        ctx.sync()
        ts = []
        for i in range(nframes):
            t = time.perf_counter()
            st.frame(pyrs[i % 2], lv)
            ctx.sync()
            ts.append(time.perf_counter() - t)
        ms = np.array(ts) * 1e3
        m = st.motion()
        planes = [oracle.pyramid(f, W, H, lv + 1)[0][lv] for f in (f0, f1)]
        o = OracleStill(oracle, StillCfg(w, h, 16, 0.5, 4.0, 1))
        o.decide()
        o.update(planes[0])
        o.finish()
        t = time.perf_counter()
        for i in range(cpu_frames):
            o.vectors(planes[(i + 1) % 2], threads=16)
            o.decide()
            o.update(planes[(i + 1) % 2])
            o.finish()
        cpu_ms = (time.perf_counter() - t) * 1e3 / cpu_frames
        o.free()
        # eight such cameras a frame: one vtg_still_frame_many against eight vtg_still_frame calls
        cams = [ctx.still(w, h, 16, 0.5, 4.0) for _ in range(8)]
        for i in range(3):
            ctx.still_frame_many(cams, [pyrs[i % 2]] * 8, [lv] * 8)
        ctx.sync()
        tm = []
        for i in range(nframes):
            t = time.perf_counter()
            ctx.still_frame_many(cams, [pyrs[i % 2]] * 8, [lv] * 8)
            ctx.sync()
            tm.append(time.perf_counter() - t)
        many_ms = float(np.percentile(np.array(tm) * 1e3, 50))
        for cam in cams:
            cam.destroy()
        key = f"{w}x{h}"
        out[key] = {"camera": f"{W}x{H} read at level {lv}", "cells": st.ncells, "frames": nframes,
                    "eight_cameras_ms_p50": round(many_ms, 4),
                    "eight_cameras_cells_per_s": round(8 * st.ncells / (many_ms * 1e-3), 1),
                    "gpu_ms_p50": round(float(np.percentile(ms, 50)), 4),
                    "gpu_ms_p99": round(float(np.percentile(ms, 99)), 4),
                    "cells_per_s": round(st.ncells / (float(np.percentile(ms, 50)) * 1e-3), 1),
                    "tracked": m["tracked"], "moving": m["moving"], "diff_median": m["diff_median"],
                    "cpu_ms_per_frame": round(cpu_ms, 3), "cpu_threads": 16, "cpu_kind": "port (oracle lane tier)"}
        log(f"still {key}: {out[key]}")
        st.destroy()
        for pyr in pyrs:
            pyr.destroy()
    return out


def stereo_bench(vtgpu, calls, nfeat=512, d0=40):
    """The tracker's stereo step at 1080p (stereo.c, track.c:1859-1914) with its dataset settings (dataset.c:154-157:
    disparities 2 .. w/8, min_ncc 0.6, ambiguity 0.85; the scan at level 2, the lowest with w/8 / 2^level <= 64) on a
    rectified pair (the right frame is the left moved d0 px, fresh noise): VT_MAX_FEATURES = 512 features, first all
    without a prior (scan + LK), then all with one (LK only). ms per call p50 / p99 (features in, results on the host),
    and the oracle's same step on one thread."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from oracle_ffi import Oracle, Plane, stereo_cfg
    f0, _, _ = vtgpu.frame_pair(W, H, 91, max_shift=0.0, max_rot=0.0)
    rng = np.random.default_rng(5)
    right = np.clip(np.roll(f0, -d0, axis=1).astype(np.int32) + rng.normal(0, 1.5, f0.shape).round().astype(np.int32),
                    0, 255).astype(np.uint8)
    ctx = vtgpu.Context(0)
    levels = 3
    pl, pr = ctx.pyramid(W, H, levels), ctx.pyramid(W, H, levels)
    pl.build(f0)
    pr.build(right)
    cfg = stereo_cfg(W, H, level=2, min_disp=2.0, max_disp=W / 8.0, min_ncc=0.6, ambiguity=0.85, fb_max=0.5,
                     max_reproj=1.0)
    base = (vtgpu.StereoFeature * nfeat)()
    for i in range(nfeat):
        base[i].x = float(rng.uniform(W / 8.0 + 20, W - 20))
        base[i].y = float(rng.uniform(20, H - 20))
    out = {"pair": f"{W}x{H}, right = left moved {d0} px", "features": nfeat, "stereo_level": 2,
           "disparities": [2, W // 8]}
    for label, prior in (("scan_and_lk", 0.0), ("lk_with_prior", float(d0) + 0.3)):
        for i in range(nfeat):
            base[i].disparity = prior
        ts = []
        for k in range(calls + 3):
            f = type(base).from_buffer_copy(base)
            t = time.perf_counter()
            ctx.stereo_match(pl, pr, cfg, f)
            if k >= 3:
                ts.append(time.perf_counter() - t)
        ms = np.array(ts) * 1e3
        fl = np.array([q.flags for q in f])
        out[label] = {"gpu_ms_p50": round(float(np.percentile(ms, 50)), 4),
                      "gpu_ms_p99": round(float(np.percentile(ms, 99)), 4),
                      "features_per_s": round(nfeat / (float(np.percentile(ms, 50)) * 1e-3), 1),
                      "matched": int(((fl & 1) > 0).sum()), "3d": int(((fl & 2) > 0).sum()),
                      "ambiguous": int(((fl & 4) > 0).sum())}
    oracle = Oracle()
    _, L = oracle.pyramid(f0, W, H, levels)
    _, R = oracle.pyramid(right, W, H, levels)
    oracle.lib.vo_stereo_match.argtypes = [C.POINTER(Plane), C.POINTER(Plane), C.POINTER(type(cfg)), C.c_int,
                                           C.c_void_p]
    for i in range(nfeat):
        base[i].disparity = 0.0
    f = type(base).from_buffer_copy(base)
    t = time.perf_counter()
    oracle.lib.vo_stereo_match(L, R, C.byref(cfg), nfeat, C.addressof(f))
    out["cpu_ms_scan_and_lk"] = round((time.perf_counter() - t) * 1e3, 3)
    out["cpu_threads"] = 1
    log(f"stereo: {out}")
    pl.destroy()
    pr.destroy()
    return out


def host_cpu():
    """The host's processor model and logical CPU count, and the threads this run uses (a one-GPU box's share)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "threads_used": min(os.cpu_count() or 1, 16),
            "note": "a one-GPU box's CPU share is 16 threads; the baseline uses that many"}


def config1(o, cores, reps=7):
    """The reference's plain CPU case (640 x 480 pair, 3 levels, 1,200 points on a 40 x 30 grid, 15 x 15, 10 iterations,
    eps 0.01, min_eig 1e-4) on the oracle's lane tier: both pyramids and the LK of all points, median of `reps` pairs,
    one thread and all threads; the GPU's plain form on the same pair beside it."""
    import vtgpu
    from oracle_ffi import lk_cfg
    w, h, L = 640, 480, 3
    f0, f1, _ = vtgpu.frame_pair(w, h, 1)
    xs1, ys1 = vtgpu.grid(40, 30, w, h)
    cfg = lk_cfg(15, 15, 10, 0.01, 1e-4, 40.0)
    res = {}
    for threads in (1, cores):
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            k0, I = o.pyramid(f0, w, h, L, threads=threads)
            k1, J = o.pyramid(f1, w, h, L, threads=threads)
            o.lk_batch(0, I, J, L - 1, cfg, xs1, ys1, threads=threads)
            ts.append(time.perf_counter() - t)
        ms = float(np.median(ts)) * 1e3
        res[f"cpu_{threads}_threads"] = {"ms_per_pair": round(ms, 3), "lk_pts_per_s": round(len(xs1) / (ms * 1e-3), 1)}
    ctx = vtgpu.Context(0)
    p0, p1 = ctx.pyramid(w, h, L), ctx.pyramid(w, h, L)
    d0, d1 = ctx.to_device(f0), ctx.to_device(f1)
    xy = ctx.to_device(np.stack([xs1, ys1], 1).ravel())
    gcfg = vtgpu.LkCfg(15, 15, 10, 0.01, 1e-4, 40.0, 0)
    out = ctx.alloc(32 * len(xs1))
    h_out = np.zeros(len(xs1), vtgpu.LK_OUT_DTYPE)
    ts = []
    for i in range(reps + 3):
        t = time.perf_counter()
        p0.build(d0, on_device=True)
        p1.build(d1, on_device=True)
        ctx.lk_track(p0, p1, L - 1, gcfg, xy, None, 0.0, out)
        out.get(h_out)
        if i >= 3:
            ts.append(time.perf_counter() - t)
    ms = float(np.median(ts)) * 1e3
    res["gpu_plain_one_pair"] = {"ms_per_pair": round(ms, 3), "lk_pts_per_s": round(len(xs1) / (ms * 1e-3), 1),
                                 "note": "host to results on the host, one pair per call (launch-bound)"}
    res["workload"] = "640x480 pair, 3 levels, 1,200 points (40x30 grid), 15x15, 10 iterations, eps 0.01"
    for b in (out, d0, d1, xy):
        b.free()
    for b in (p0, p1):
        b.destroy()
    ctx.close()
    log(f"config 1: {res}")
    return res


def store_of(vtgpu, name):
    return {"f32": vtgpu.STORE_F32, "f32_u8": vtgpu.STORE_F32_U8}[name]


def timed_steps(bt, steps, warmup):
    """warmup untimed steps, then steps timed: step i+1 is launched before step i's results are waited for, so
    downloads overlap the next step. Returns the seconds and each step's latency in ms (the launch of a step to its
    results on the host)."""
    for i in range(warmup):
        bt.launch(i)
        bt.finish(i)
    lat, ts = [], []
    t0 = time.perf_counter()
    for i in range(steps):
        ts.append(time.perf_counter())
        bt.launch(i)
        if i >= 1:
            bt.finish(i - 1)
            lat.append(time.perf_counter() - ts[i - 1])
    bt.finish(steps - 1)
    lat.append(time.perf_counter() - ts[steps - 1])
    return time.perf_counter() - t0, np.array(lat) * 1e3


def other_form(vtgpu, d_frames, corners, a, xs, ys):
    """The headline step in its other launch form (captured graphs replayed, or with --graph the launches issued as
    they are), same inputs, a.graph_steps steps after 3 untimed ones."""
    bt = Batched(vtgpu, d_frames, corners, a.batch, xs, ys, graph=not a.graph, one_stream=a.one_stream,
                 pyr_sets=a.pyr_sets, storage=store_of(vtgpu, a.storage), join=a.join)
    bt.prepare()
    total, lat_ms = timed_steps(bt, a.graph_steps, 3)
    r = {"graph": not a.graph, "steps": a.graph_steps, "pairs_per_s": round(a.graph_steps * a.batch / total, 2),
         "p50_ms": round(float(np.percentile(lat_ms, 50)), 4), "p99_ms": round(float(np.percentile(lat_ms, 99)), 4),
         "late_steps": int((lat_ms > np.median(lat_ms) + 0.5).sum())}
    bt.close()
    log(f"other launch form: {r}")
    return r


def pcie_rate(vtgpu, bt, f0s, f1s, steps):
    """The batched step with its frames arriving in pinned host memory: each step's 2B frames are uploaded over PCIe on
    an upload context while the previous step computes, and the step waits for its upload. Pairs per second
    including the uploads."""
    B = bt.B
    up = [vtgpu.Context(0), vtgpu.Context(0)]
    host = [vtgpu.PinnedArray(up[0], f0s[k].shape, np.uint8) for k in range(UNIQ)] + \
           [vtgpu.PinnedArray(up[0], f1s[k].shape, np.uint8) for k in range(UNIQ)]
    for k in range(UNIQ):
        host[k].arr[...] = f0s[k]
        host[UNIQ + k].arr[...] = f1s[k]

    def upload(i):
        g = i % bt.groups
        c = up[i % 2]
        for k in range(B):
            j = g * B + k
            a0, a1 = bt.frames[j]
            a0.put_async(c, host[j % UNIQ].arr)
            a1.put_async(c, host[UNIQ + j % UNIQ].arr)

    nbytes = 2 * B * W * H
    upload(0)
    up[0].sync()
    for i in range(3):  # warm
        bt.launch(i, after=up[i % 2])
        upload(i + 1)
        bt.finish(i)
    t0 = time.perf_counter()
    for i in range(steps):
        bt.launch(i, after=up[i % 2])
        upload(i + 1)
        if i >= 1:
            bt.finish(i - 1)
    bt.finish(steps - 1)
    total = time.perf_counter() - t0
    for c in up:
        c.sync()
    for hp in host:
        hp.free()
    r = {"pairs_per_s": round(steps * B / total, 2), "ms_per_step": round(1e3 * total / steps, 4),
         "upload_bytes_per_step": nbytes, "upload_gb_per_s": round(nbytes * steps / total / 1e9, 2),
         "note": "frames uploaded from pinned host memory each step, overlapped with the previous step's compute"}
    log(f"pcie-inclusive: {r}")
    return r


def cpu_baseline(f0s, f1s, corners, xs, ys, budget_s, gpu_out):
    """The oracle's reference-faithful tier (C, the reference's lane arithmetic) on this host, on the same workload:
    whole 1080p pairs (both pyramids, LK of all 16,200 points, the 64 x 64 template's exhaustive NCC map). All cores
    (pthreads over rows and points): the median over at least 7 pairs, for 60 % of the budget; one thread: whole pairs
    for the rest. Also the GPU's flow error against the FP64 oracle on pair 0, over its well-conditioned points."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from gpu_util import conditioning, precision_eval
    from oracle_ffi import Oracle, lk_cfg
    o = Oracle()
    cores = min(os.cpu_count() or 1, 16)
    cfg = lk_cfg(WIN, WIN, 10, 0.01, 1e-4, 40.0)
    x1, y1 = W - TPL - 2, H - TPL - 2
    rows = (y1 - 1) // 16 + 1

    def pair(k, threads, lk_stride=1, band=rows):
        """Seconds of one pair: pyramids, LK of every lk_stride-th point (scaled), NCC over `band` cell rows
        (scaled)."""
        f0, f1 = f0s[k % UNIQ], f1s[k % UNIQ]
        t = time.perf_counter()
        k0, I = o.pyramid(f0, W, H, LEVELS, threads=threads)
        k1, J = o.pyramid(f1, W, H, LEVELS, threads=threads)
        t_pyr = time.perf_counter() - t
        t = time.perf_counter()
        ref = o.lk_batch(0, I, J, LEVELS - 1, cfg, xs[::lk_stride], ys[::lk_stride], threads=threads)
        t_lk = (time.perf_counter() - t) * lk_stride
        cx, cy = corners[k % UNIQ]
        app, norm = o.ncc_template(k0[0], cx, cy, TPL, TPL)
        t = time.perf_counter()
        o.ncc_search(k1[0], app, TPL, TPL, norm, 1, x1, 1, min(y1, 16 * band), 1, threads=threads)
        t_ncc = (time.perf_counter() - t) * rows / band
        return t_pyr, t_lk, t_ncc, (I, J, ref, k0, k1)  # k0, k1 own the planes I and J point into

    t_start = time.perf_counter()
    rounds = []
    # all cores: whole pairs over the rotating seeded inputs, at least 7, until 60 % of the budget is spent
    while len(rounds) < 7 or (time.perf_counter() - t_start < 0.6 * budget_s and len(rounds) < 1000):
        rounds.append(pair(len(rounds), cores))
        if len(rounds) == 1:
            I0, J0, ref0, keep0, keep1 = rounds[0][3]
        else:
            rounds[-1] = rounds[-1][:3]  # only pair 0's planes are kept
    tot = sorted(r[0] + r[1] + r[2] for r in rounds)
    med = tot[len(tot) // 2]
    stage = [float(np.median([r[q] for r in rounds])) for q in range(3)]
    # one thread: whole pairs for the rest of the budget (at least one)
    ones = []
    while not ones or time.perf_counter() - t_start < budget_s:
        p1, l1, n1, _ = pair(len(ones), 1)
        ones.append(p1 + l1 + n1)
    one = float(np.median(ones))
    sample = (f"{len(rounds)} whole 1080p pairs on {cores} threads (median; {UNIQ} seeded pairs rotating), each: both "
              f"pyramids, LK of all {NPTS} points, the full NCC map ({rows} cell rows); {len(ones)} whole pairs on one "
              f"thread; lane-model tier")
    log(f"cpu: {cores} threads: pyramids {stage[0] * 1e3:.1f} ms, LK {stage[1] * 1e3:.0f} ms, NCC "
        f"{stage[2] * 1e3:.0f} ms -> {1 / med:.3f} pairs/s ({len(rounds)} pairs); 1 thread: {1 / one:.4f} pairs/s "
        f"({len(ones)} pairs); {time.perf_counter() - t_start:.1f} s")
    # the GPU's records of pair 0 against the reference tier (the reference's lane arithmetic and x87 driver) and the
    # FP64 oracle, over every point
    well, f64 = conditioning(o, I0, J0, LEVELS - 1, cfg, xs, ys, threads=cores)
    ref, g = ref0, gpu_out
    both = (g["status"] == 0) & (f64["status"] == 0)
    e_all = np.maximum(np.abs(g["x"] - f64["x"]), np.abs(g["y"] - f64["y"]))
    err = e_all[both]
    same_status = float((g["status"] == ref["status"]).mean())
    same_record = float((np.ascontiguousarray(g).view(np.uint32).reshape(-1, 8) ==
                         np.ascontiguousarray(ref).view(np.uint32).reshape(-1, 8)).all(1).mean())
    ok_ref = (g["status"] == 0) & (ref["status"] == 0)
    dref = np.maximum(np.abs(g["x"] - ref["x"]), np.abs(g["y"] - ref["y"]))[ok_ref]
    # storage precision against the FP64 oracle (pyramids fp16 / bf16 / exact, fp64 or mixed solve)
    import vtgpu
    pctx = vtgpu.Context(0)
    prec = precision_eval(vtgpu, o, pctx, threads=cores)
    # the Scharr gradient variant (no reference counterpart) against the oracle's FP64 Scharr tier, per storage
    prec["scharr_variant"] = precision_eval(vtgpu, o, pctx, threads=cores, grad=1)
    for tag, pr in (("sobel", prec), ("scharr", prec["scharr_variant"])):
        log(f"precision ({tag}): " + ", ".join(f"{k} p99 {v['textured']['p99_px']:.2e} px" for k, v in pr.items()
                                               if isinstance(v, dict) and "textured" in v))
    return ({"value": round(1.0 / med, 4), "unit": "frame-pairs/s", "cores": cores, "kind": "port",
             "sample": sample, "host": host_cpu(), "one_thread": {"value": round(1.0 / one, 5), "unit": "frame-pairs/s",
                                                                  "cores": 1},
             "config1_640x480": config1(o, cores),
             "lk_pts_per_s": round(NPTS / stage[1], 1), "stage_ms": {
                 "pyramids": round(stage[0] * 1e3, 2), "lk": round(stage[1] * 1e3, 1), "ncc": round(stage[2] * 1e3, 1)},
             "cpu_seconds": round(time.perf_counter() - t_start, 2)},
            {"points": int(len(xs)), "converged_in_both_vs_f64": int(both.sum()),
             "well_conditioned": int(well.sum()),
             "vs_f64_median_px": float(np.median(err)), "vs_f64_p99_px": float(np.percentile(err, 99)),
             "vs_f64_max_px": float(err.max()),
             "share_within_0.05px_all_converged": float((err <= 0.05).mean()),
             "vs_f64_max_px_well_conditioned": float(e_all[both & well].max()),
             "status_equal_vs_reference_tier": same_status,
             "records_bit_equal_vs_reference_tier": same_record,
             "vs_reference_tier_max_px": float(dref.max()) if len(dref) else 0.0}, prec)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=8, help="frame pairs per step of the batched form")
    ap.add_argument("--pairs", type=int, default=72, help="device-resident frame pairs rotated through")
    ap.add_argument("--graph", action="store_true", help="the headline steps replay captured HIP graphs instead of "
                    "issuing their launches")
    ap.add_argument("--graph-steps", type=int, default=30,
                    help="steps of the other form (graph replay, or issued launches with --graph) measured beside the "
                         "headline (0: skip)")
    ap.add_argument("--join", action="store_true", help="issued launches: fork the searches' context from the LK "
                    "context at each step's start and join it at the end, as a captured graph must")
    ap.add_argument("--tune", action="store_true", help="run the autotuner on pair 0 before the timed steps and use "
                    "its picks")
    ap.add_argument("--one-stream", action="store_true",
                    help="the searches on the LK stream too (no fork and join inside the step)")
    ap.add_argument("--storage", choices=["f32", "f32_u8"], default="f32",
                    help="level storage of the batched step's pyramids (f32_u8 also keeps the u8 frame beside level 0)")
    ap.add_argument("--pyr-sets", type=int, default=0, help="pyramid sets the steps rotate through (0: one per "
                    "frame group with graphs, one without)")
    ap.add_argument("--plain-steps", type=int, default=8)
    ap.add_argument("--batch-steps", type=int, default=20, help="graph replays of the uneven 64-pair batch (0: skip)")
    ap.add_argument("--serve-frames", type=int, default=300, help="frames per stream of the serving runs (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU baseline sample (0: skip)")
    ap.add_argument("--still-frames", type=int, default=60, help="timed frames of each still-cell camera (0: skip)")
    ap.add_argument("--stereo-calls", type=int, default=30, help="timed calls of the 1080p stereo step (0: skip)")
    ap.add_argument("--pcie-steps", type=int, default=1, help="1: also time the step with its frames uploaded from "
                    "pinned host memory each step (0: skip)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="one device-to-device copy of this many MiB first: a plain copy of known size that calibrates "
                         "the HBM byte counters of a profiled run (0: none)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported (streams shard with no "
              f"exchange, so N GPUs run N independent copies of this step)", file=sys.stderr)
        sys.exit(2)
    import vtgpu
    pairs = max(a.batch, (a.pairs // a.batch) * a.batch)
    f0s, f1s = [], []
    for s in range(UNIQ):
        f0, f1, _ = vtgpu.frame_pair(W, H, 100 + s)
        f0s.append(f0)
        f1s.append(f1)
    xs, ys = vtgpu.grid(NX, NY, W, H)
    rng = np.random.default_rng(2)
    corners = [(int(rng.integers(100, W - 200)), int(rng.integers(100, H - 200))) for _ in range(UNIQ)]
    up = vtgpu.Context(0)
    if a.copy_mib > 0:
        src, dst = up.alloc(a.copy_mib << 20), up.alloc(a.copy_mib << 20)
        up.check(vtgpu.lib().vtg_copy_d2d(up.h, dst.ptr, src.ptr, a.copy_mib << 20), "copy_d2d")
        up.sync()
        log(f"calibration copy: {a.copy_mib} MiB device to device")
        src.free()
        dst.free()
    d_frames = [(up.to_device(f0s[i % UNIQ]), up.to_device(f1s[i % UNIQ])) for i in range(pairs)]
    up.sync()

    bt = Batched(vtgpu, d_frames, corners, a.batch, xs, ys, graph=a.graph, one_stream=a.one_stream,
                 pyr_sets=a.pyr_sets, storage=store_of(vtgpu, a.storage), join=a.join)
    tuned = bt.tune() if a.tune else None
    bt.prepare()
    total, lat_ms = timed_steps(bt, a.steps, a.warmup)
    pairs_per_s = a.steps * a.batch / total
    # per-pair algorithmic counts from each pair's own LK records (group 0 holds pairs 0..B-1)
    bt.launch(0)
    h_out, h_found = bt.finish(0)
    flops = [lk_flops(h_out[k * NPTS:(k + 1) * NPTS], xs, ys, LEVELS - 1, W, H, WIN) for k in range(a.batch)]
    gpu_out0 = h_out[:NPTS].copy()
    found0 = h_found[0]
    log(f"batched: {pairs_per_s:.1f} pairs/s, batch p50 {np.percentile(lat_ms, 50):.3f} ms p99 "
        f"{np.percentile(lat_ms, 99):.3f} ms; pair 0 NCC pick ({found0.hx / 2},{found0.hy / 2}) {found0.ncc:.4f}")

    # per-kernel times: the same batched steps launched eagerly with the two streams serialised (each kernel alone on
    # the device), HIP events around every launch on its own stream
    bt.A.timing(True)
    bt.S.timing(True)
    ksteps = min(a.steps, 10)
    for i in range(ksteps):
        bt.enqueue(i % bt.groups, serial=True)
    bt.A.sync()
    kernels = {}
    for ctx in (bt.A, bt.S):
        for k in ("pyr_reduce", "lk_track", "ncc_template", "ncc_xcorr_cells", "ncc_top8", "ncc_refine_final"):
            ms, n = ctx.timing_read(k)
            if n:
                kernels[k] = {"avg_us": round(1e3 * ms / n, 2), "launches": n, "ms_per_step": round(ms / ksteps, 4)}
    bt.A.timing(False)
    bt.S.timing(False)
    npos = (W - TPL - 2) * (H - TPL - 2)
    lk_flop_step = float(np.sum(flops))
    work = {  # per step (one batch): fp32 flops, int8 MAC ops, HBM bytes
        "lk_track": ("valu", lk_flop_step / 1e12, PEAK_FP32_TFLOPS, "TFLOP/s"),
        # the cell map: the cross term's 2 tw th ops per corner on the matrix cores (the band's zero products are not
        # counted), then each corner's box sums and score in the same workgroup
        "ncc_xcorr_cells": ("mfma", a.batch * 2.0 * TPL * TPL * npos / 1e12, PEAK_INT8_MFMA_TOPS, "TOP/s"),
        "pyr_reduce": ("hbm", a.batch * 2 * pyr_bytes(W, H, LEVELS, 4) / 1e9, PEAK_HBM_GBS, "GB/s"),
    }
    for k, (b, wk, pk, u) in work.items():
        if k in kernels:
            ach = wk / (kernels[k]["ms_per_step"] * 1e-3)
            kernels[k].update({"achieved": round(ach, 3), "unit": u, "frac": round(ach / pk, 4)})
    log("kernels:", json.dumps(kernels))
    dom = max((k for k in kernels if k in work), key=lambda k: kernels[k]["ms_per_step"])
    bound, per_step, peak, unit = work[dom]
    achieved = per_step / (kernels[dom]["ms_per_step"] * 1e-3)
    lk_us = kernels["lk_track"]["avg_us"]

    plain, plain_detail = plain_rate(vtgpu, d_frames, corners, xs, ys, max(1, a.plain_steps))
    log(f"plain: {plain:.2f} pairs/s {plain_detail}")
    # the batched form's stages per pair (serial kernel pass) beside the plain form's
    bstage = {"pyramids": kernels.get("pyr_reduce", {}).get("ms_per_step", 0.0) / a.batch,
              "lk": kernels["lk_track"]["ms_per_step"] / a.batch,
              "ncc": sum(kernels[k]["ms_per_step"] for k in kernels if k.startswith("ncc_")) / a.batch}
    by_stage = {s: {"plain_ms_per_pair": plain_detail["stage_ms_per_pair"][s],
                    "batched_ms_per_pair": round(v, 4),
                    "ratio": round(plain_detail["stage_ms_per_pair"][s] / v, 2) if v > 0 else None}
                for s, v in bstage.items()}
    pcie = pcie_rate(vtgpu, bt, f0s, f1s, a.steps) if a.pcie_steps else {}
    other = other_form(vtgpu, d_frames, corners, a, xs, ys) if a.graph_steps > 0 else {}
    batch = batching(vtgpu, a.batch_steps) if a.batch_steps > 0 else {}
    serve = serving(vtgpu, a.serve_frames) if a.serve_frames > 0 else {}
    still = still_bench(vtgpu, a.still_frames) if a.still_frames > 0 else {}
    stereo = stereo_bench(vtgpu, a.stereo_calls) if a.stereo_calls > 0 else {}
    cpu, err, prec = ({}, {}, {})
    if a.cpu_seconds > 0:
        cpu, err, prec = cpu_baseline(f0s, f1s, corners, xs, ys, a.cpu_seconds, gpu_out0)
    line = {
        "metric": METRIC, "value": round(pairs_per_s, 2), "unit": "frame-pairs/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(1e3 * total / a.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "u8 frames, fp32 Q6 levels (exact integers), fp32 patch sums, fp64 per-point driver",
        "data": "synthetic (seeded 200-blob frames, rigid motion + noise)",
        "config": {"workload": "1080p frame pairs: 2 x 4-level pyramid + LK 16,200 pts 21x21 + NCC 64x64 exhaustive "
                               "per pair", "w": W, "h": H, "levels": LEVELS, "points_per_pair": NPTS,
                   "window": f"{WIN}x{WIN}", "ncc_template": f"{TPL}x{TPL}", "ncc_corners": npos,
                   "pairs_per_step": a.batch,
                   "storage": {"f32": "f32 (fp32 Q6 levels: exact integers; the searches read the u8 frames in "
                                       "place; the level-1 reduce also writes the fp32 level 0: 4 B per level-0 pixel "
                                       "beyond the pyramid's algorithmic bytes, which LK's fp32 reads repay)",
                               "f32_u8": "f32_u8 (as f32, and the u8 frame kept beside level 0: 5 B per level-0 "
                                         "pixel beyond the pyramid's algorithmic bytes)"}[a.storage],
                   "lk_points_per_wavefront": 3, "graph": a.graph, "fork_join": a.graph or a.join,
                   "autotuned": tuned,
                   "parallelism": "1 GPU, 1 stream" if a.one_stream else "1 GPU, 2 streams",
                   "input_rotation": f"{pairs} device frame pairs ({pairs * 2 * W * H / 2 ** 20:.0f} MiB)"},
        "plain_pairs_per_s": round(plain, 2), "vs_plain": round(pairs_per_s / plain, 2),
        "vs_plain_by_stage": by_stage, "plain_ncc_map_kernel": plain_detail["ncc_map_kernel"],
        "pcie_inclusive": pcie,
        "other_launch_form": other,
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 4), "p99_ms": round(float(np.percentile(lat_ms, 99)), 4),
        "latency_unit": f"one step = {a.batch} pairs, submit to results on the host",
        "lk_pts_per_s": round(a.batch * NPTS / (lk_us * 1e-6), 1),
        "kernels": kernels,
        "roofline": {"bound": bound, "kernel": dom, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
                     "frac": round(achieved / peak, 4), "traffic": measured_traffic(dom),
                     "traffic_source": os.path.relpath(TRAFFIC_MD, REPO)},
        "flow_error": err,
        "batching": batch,
        "serving": serve,
        "precision": prec,
        "still_cells": still,
        "stereo": stereo,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
