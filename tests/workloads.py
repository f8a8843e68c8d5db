"""Synthetic workloads shared by the GPU tests and bench.py.

Uneven batch: frame pairs whose sizes are drawn from 640x480, 1280x720, 1920x1080 and 3840x2160, with 500-20,000
random points each and a translation of up to 24 px per axis; each pair gets the levels its size class wants (3 for
640x480 and 1280x720, 4 for 1920x1080, 5 for 3840x2160) or more, so that the top level's displacement is at most a
quarter of the window.
"""
import math

import numpy as np

SIZES = [(640, 480), (1280, 720), (1920, 1080), (3840, 2160)]
SIZE_LEVELS = {640: 3, 1280: 3, 1920: 4, 3840: 5}


def levels_for(w, shift, win):
    need = 1 + max(0, math.ceil(math.log2(max(shift, 1e-9) / (win / 4.0))))
    return int(min(5, max(SIZE_LEVELS[w], need)))


def batch_spec(seed, npairs=64, nmin=500, nmax=20000, max_shift=24.0, win=21, sizes=SIZES):
    """[(w, h, levels, n, pair_seed)]: the batch's pairs, seeded."""
    import vtgpu
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(npairs):
        w, h = sizes[int(rng.integers(len(sizes)))]
        n = int(rng.integers(nmin, nmax + 1))
        ps = int(rng.integers(1, 2 ** 31 - 1))
        m = vtgpu.Motion()
        vtgpu.synth().vts_draw_motion(ps, max_shift, 0.0, vtgpu.C.byref(m))
        out.append((w, h, levels_for(w, max(abs(m.dx), abs(m.dy)), win), n, ps))
    return out


class Batch:
    """A batch resident on the device: both frames of every pair, one pyramid per frame, the points of all pairs in one
    array (pair k's slice offsets[k] .. offsets[k+1]), an output array, and the host copies for checking."""

    def __init__(self, vtgpu, ctx, spec, max_shift=24.0, storage=0, margin=24):
        self.spec, self.ctx = spec, ctx
        self.frames, self.host, self.pyrs, self.motion = [], [], [], []
        xy, self.offsets = [], [0]
        for (w, h, L, n, ps) in spec:
            f0, f1, mo = vtgpu.frame_pair(w, h, ps, max_shift=max_shift, max_rot=0.0, threads=16)
            self.host.append((f0, f1))
            self.motion.append(mo)
            self.frames.append((ctx.to_device(f0), ctx.to_device(f1)))
            self.pyrs.append((ctx.pyramid(w, h, L, storage), ctx.pyramid(w, h, L, storage)))
            xs, ys = vtgpu.points(ps, n, w, h, margin)
            xy.append(np.stack([xs, ys], 1).ravel())
            self.offsets.append(self.offsets[-1] + n)
        self.xy_host = np.concatenate(xy).astype(np.float32)
        self.xy = ctx.to_device(self.xy_host)
        self.out = ctx.alloc(32 * self.offsets[-1])
        self.tops = [L - 1 for (_, _, L, _, _) in spec]

    @property
    def npoints(self):
        return self.offsets[-1]

    def pixels(self):
        return sum(w * h for (w, h, _, _, _) in self.spec)

    def enqueue(self, cfg, solve=0):
        """Both pyramids of every pair (one launch per level and tile height), then one LK launch per schedule."""
        self.ctx.build_many([p for pr in self.pyrs for p in pr], [f for fr in self.frames for f in fr])
        self.ctx.lk_track_many([p[0] for p in self.pyrs], [p[1] for p in self.pyrs], self.tops, self.offsets, cfg,
                               self.xy, out=self.out, solve=solve)

    def classes(self):
        """Index of the first pair of each (size, levels) class."""
        seen = {}
        for k, (w, h, L, _, _) in enumerate(self.spec):
            seen.setdefault((w, h, L), k)
        return seen

    def destroy(self):
        for p0, p1 in self.pyrs:
            p0.destroy()
            p1.destroy()


# Serving: (w, h, fps, levels, grid nx x ny) per stream class; 3 x 1080p60, 3 x 720p30, 2 x 480p60.
STREAMS = [(1920, 1080, 60.0, 4, 180, 90)] * 3 + [(1280, 720, 30.0, 4, 120, 68)] * 3 + [(640, 480, 60.0, 3, 68, 60)] * 2


class StreamFrames:
    """A stream's frames: K distinct frames of one scene whose position walks randomly (sigma px per frame and axis),
    served in ping-pong order 0, 1, .., K-1, K-2, .., 1, 0, 1, .. so consecutive frames are always neighbours of the
    walk; its tracking points are a grid."""

    def __init__(self, vtgpu, w, h, seed, k=24, sigma=1.5, nx=None, ny=None, noise=2.0):
        rng = np.random.default_rng(seed)
        steps = rng.normal(0.0, sigma, (k, 2))
        steps[0] = 0.0
        walk = np.cumsum(steps, 0)
        self.w, self.h, self.k = w, h, k
        self.frames = [vtgpu.frame(w, h, seed, float(walk[i, 0]), float(walk[i, 1]), 0.0, noise, seed * 1000 + i,
                                   threads=16) for i in range(k)]
        xs, ys = vtgpu.grid(nx or 16, ny or 12, w, h)
        self.xy = np.stack([xs, ys], 1).ravel().astype(np.float32)

    def index(self, i):
        period = 2 * (self.k - 1)
        j = i % period
        return j if j < self.k else period - j

    def __getitem__(self, i):
        return self.frames[self.index(i)]


def enqueue_fused(batch, cfg, solve=0, fb_max=0.0):
    """The same batch as Batch.enqueue in ONE persistent launch (pyramids and LK fused)."""
    b = batch
    b.ctx.pyr_lk_fused([p[0] for p in b.pyrs], [p[1] for p in b.pyrs], [f[0] for f in b.frames],
                       [f[1] for f in b.frames], b.tops, b.offsets, cfg, b.xy, b.out, fb_max=fb_max, solve=solve)
