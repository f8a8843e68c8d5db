// lk.hip: sparse pyramidal Lucas-Kanade on the GPU, one wavefront per point (lk_driver.inc, src/kern/lk_driver.inc:7-121;
// patch operations lk_vx.c:36-175; the OP_LK batch loop, src/kern/track.c:275-292).
//
// Lane layout. A w x h window is cut into G = 64 / w row groups of w lanes each (21 x 21: 3 groups of 7 rows, 63 lanes;
// 15 x 15: 4 groups of 4 rows, 60 lanes). Lane (k, g) owns window column k and rows [g * chunk, (g + 1) * chunk): its
// template values T, Tx, Ty live in registers for the whole level. A lane walks MAXR consecutive window rows (a
// compile-time count no smaller than the chunk); a group whose chunk would run past the window's last row walks the
// last MAXR rows instead, and rows that are not the lane's own carry zero gradients, so they add nothing to any sum.
// Direct schedule (the default). The window build reads the level in memory straight into registers: per image row,
// four taps and three x-interpolations give the lane's three columns of the bilinear window S, and consecutive rows
// give S rows and the Sobel gradients, with no staging and no index division. Each iteration reads two taps per row of
// J, x-interpolated once per row and y-interpolated between rows, as the reference's vector code carries a row
// (lk_vx.c:100-126). The level bounds test (a window m = w / 2 + 1 pixels clear of the left/top edge and m + 1 of the
// right/bottom edge) keeps every pixel read inside the level, so no coordinate is clamped.
// Staged schedule (margin > 0). At the start of a level the wave loads, in one batch, the template's (w + 3) x (h + 3)
// patch of I and J's neighbourhood of the level's starting estimate (the window plus a margin of m pixels) into its
// LDS slice; S, the gradients and every iteration's samples then come from LDS. The arithmetic of both schedules is
// the same operation for operation, so their results agree bit for bit.
// Element types. Level values are read in their own units (u8: grey, u16: Q6, f16/bf16: grey) and carried at that
// scale through the interpolation; the Q6 factor is applied where it is exact (a power of two), so the results equal
// those of interpolating Q6 values.
// Sums over the window (G, b, the residual) reduce across the wave with DPP (quad permutes, row mirrors, row
// broadcasts) and one readlane, so every lane holds the same total. The per-point driver (level walk, eigenvalue test,
// 2x2 solve, convergence, bounds) runs in fp64 on those uniform values, the reference's being x87 long double; the
// patch arithmetic is fp32 with fused multiply-adds where the reference's vector code fuses them.
// Batches. A launch covers points of up to 64 frame pairs of any sizes: a table in the kernel arguments maps each pair
// to its pyramids' device descriptors, its top level and its slice of the point array. Workgroups are mapped to points
// so that each of the 8 XCDs (which take workgroups in turn) works on one contiguous run of points, keeping
// neighbouring windows in one L2.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <climits>

#include "reduce_tile.hpp"
#include "vtg_internal.hpp"

namespace vtg {

namespace {

struct Cfg {
    int w, h, max_iter;
    float eps, min_eig, max_res;
};

constexpr int BMAX = 64; // pairs per launch (inline in the kernel arguments)
constexpr int NXCD = 8;  // XCDs of the device; workgroup i runs on XCD i mod 8

// A launch's pairs: pair q's points are xy[start[q] ..] (global indices), numbered pre[q] .. pre[q + 1] - 1 in the
// launch.
struct LkArgs {
    const PyrView *I[BMAX], *J[BMAX];
    int top[BMAX];
    int start[BMAX];
    int pre[BMAX + 1];
    int npairs, n, per_xcd, waves;
    Cfg cfg;
    int margin, lds_floats;
    const float *xy, *guess;
    float fb_max;
    int fb_top;                 // the flowback's top level (-1: the pair's own)
    lk_out *out;
};

// A level element in its own units, and how many Q6 units one of them is.
template <typename T> struct Ld;
template <> struct Ld<uint8_t> {
    static constexpr float scale = 64.0f;
    static __device__ __forceinline__ float get(uint8_t v) { return (float)v; }
};
template <> struct Ld<uint16_t> {
    static constexpr float scale = 1.0f;
    static __device__ __forceinline__ float get(uint16_t v) { return (float)v; }
};
template <> struct Ld<float> {
    static constexpr float scale = 1.0f;
    static __device__ __forceinline__ float get(float v) { return v; }
};
template <> struct Ld<__half> {
    static constexpr float scale = 64.0f;
    static __device__ __forceinline__ float get(__half v) { return __half2float(v); }
};
template <> struct Ld<__hip_bfloat16> {
    static constexpr float scale = 64.0f;
    static __device__ __forceinline__ float get(__hip_bfloat16 v) { return __bfloat162float(v); }
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// A level read through a buffer resource: the base and size are wave-uniform (scalar registers), each load is the
// resource plus a lane's 32-bit byte offset o plus a uniform offset so (a row step, kept in a scalar register), and a
// read past the level's bytes returns 0 instead of touching memory. Two or four neighbouring elements of a row come
// in one load.
template <typename T> struct Row;
template <> struct Row<uint8_t> {
    static __device__ __forceinline__ void pair(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so, float &a0, float &a1)
    {
        const unsigned v = __builtin_amdgcn_raw_buffer_load_b16(r, o, so, 0);
        a0 = (float)(v & 0xffu);
        a1 = (float)(v >> 8);
    }
    static __device__ __forceinline__ void quad(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so, float *a)
    {
        const unsigned v = __builtin_amdgcn_raw_buffer_load_b32(r, o, so, 0);
        a[0] = (float)(v & 0xffu);
        a[1] = (float)((v >> 8) & 0xffu);
        a[2] = (float)((v >> 16) & 0xffu);
        a[3] = (float)(v >> 24);
    }
};
template <> struct Row<uint16_t> {
    static __device__ __forceinline__ void pair(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so, float &a0, float &a1)
    {
        const unsigned v = __builtin_amdgcn_raw_buffer_load_b32(r, o, so, 0);
        a0 = (float)(v & 0xffffu);
        a1 = (float)(v >> 16);
    }
    static __device__ __forceinline__ void quad(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so, float *a)
    {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, o, so, 0);
        a[0] = (float)(v[0] & 0xffffu);
        a[1] = (float)(v[0] >> 16);
        a[2] = (float)(v[1] & 0xffffu);
        a[3] = (float)(v[1] >> 16);
    }
};
template <> struct Row<float> {
    static __device__ __forceinline__ void pair(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so, float &a0, float &a1)
    {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, o, so, 0);
        a0 = __uint_as_float(v[0]);
        a1 = __uint_as_float(v[1]);
    }
    static __device__ __forceinline__ void quad(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so, float *a)
    {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, o, so, 0);
        for (int i = 0; i < 4; i++)
            a[i] = __uint_as_float(v[i]);
    }
};
template <typename H> struct Row16 {
    static __device__ __forceinline__ float cvt(unsigned short b)
    {
        H h;
        __builtin_memcpy(&h, &b, 2);
        return Ld<H>::get(h);
    }
    static __device__ __forceinline__ void pair(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so, float &a0, float &a1)
    {
        const unsigned v = __builtin_amdgcn_raw_buffer_load_b32(r, o, so, 0);
        a0 = cvt((unsigned short)(v & 0xffffu));
        a1 = cvt((unsigned short)(v >> 16));
    }
    static __device__ __forceinline__ void quad(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so, float *a)
    {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, o, so, 0);
        a[0] = cvt((unsigned short)(v[0] & 0xffffu));
        a[1] = cvt((unsigned short)(v[0] >> 16));
        a[2] = cvt((unsigned short)(v[1] & 0xffffu));
        a[3] = cvt((unsigned short)(v[1] >> 16));
    }
};
template <> struct Row<__half> : Row16<__half> {};
template <> struct Row<__hip_bfloat16> : Row16<__hip_bfloat16> {};

template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t level_rsrc(const Level &lv, unsigned &pitch_bytes)
{
    const uint64_t v = (uint64_t)lv.p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    const int pitch = __builtin_amdgcn_readfirstlane(lv.pitch), h = __builtin_amdgcn_readfirstlane(lv.h);
    pitch_bytes = (unsigned)pitch * (unsigned)sizeof(T);
    return __builtin_amdgcn_make_buffer_rsrc((void *)(((uint64_t)hi << 32) | lo), 0,
                                             (int)((unsigned)pitch * (unsigned)h * (unsigned)sizeof(T)), 0x00020000);
}

// LDS written by one lane and read by another lane of the same wave: order the accesses, no workgroup barrier.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_add(float v)
{
    const int moved = __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false);
    return v + __int_as_float(moved);
}

// Sum over the 64 lanes, the same value in every lane. Every lane must be active.
__device__ __forceinline__ float wave_sum(float v)
{
    v = dpp_add<0xb1, 0xf>(v);  // quad_perm [1,0,3,2]
    v = dpp_add<0x4e, 0xf>(v);  // quad_perm [2,3,0,1]
    v = dpp_add<0x141, 0xf>(v); // row_half_mirror
    v = dpp_add<0x140, 0xf>(v); // row_mirror: each row of 16 holds its sum
    v = dpp_add<0x142, 0xa>(v); // row_bcast:15 into rows 1 and 3
    v = dpp_add<0x143, 0xc>(v); // row_bcast:31 into rows 2 and 3: lane 63 holds the total
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// This lane's place in the window: column k, its own rows [r0, r0 + nr), the MAXR rows it walks from re (inside the
// window when h >= MAXR), and which of those are its own (bit i: row re + i).
struct Geo {
    int W, H, k, r0, nr, re;
    unsigned on;
};

template <int MAXR>
__device__ __forceinline__ Geo geometry(const Cfg &c, int lane)
{
    Geo g;
    g.W = c.w;
    g.H = c.h;
    const int groups = 64 / c.w, chunk = (c.h + groups - 1) / groups;
    const int grp = lane / c.w;
    g.k = lane - grp * c.w;
    g.r0 = grp * chunk;
    int nr = c.h - g.r0;
    nr = nr < chunk ? nr : chunk;
    g.nr = grp < groups && nr > 0 ? nr : 0;
    const int last = c.h - MAXR > 0 ? c.h - MAXR : 0;
    g.re = g.nr == 0 ? 0 : (g.r0 < last ? g.r0 : last);
    g.on = 0;
#pragma unroll
    for (int i = 0; i < MAXR; i++)
        if (g.re + i >= g.r0 && g.re + i < g.r0 + g.nr)
            g.on |= 1u << i;
    return g;
}

template <int MAXR> struct Tile {
    float t[MAXR], tx[MAXR], ty[MAXR];
};

// Sobel gradients of window rows re .. re + MAXR - 1 at the lane's column from the lane's three columns of S rows
// re .. re + MAXR + 1 (values in units of `ks` Q6), with T = S interior; returns this lane's part of G. Rows that are not
// the lane's own get zero gradients (build_vx, lk_vx.c:36-97).
template <int MAXR>
__device__ __forceinline__ void sobel_rows(const float (&S)[MAXR + 2][3], float ks, const Geo &g, Tile<MAXR> &tl,
                                           float &sxx, float &sxy, float &syy)
{
    const float kg = ks * (1.0f / 512.0f);
#pragma unroll
    for (int i = 0; i < MAXR; i++) {
        const float *u = S[i], *m = S[i + 1], *d = S[i + 2];
        const float du = u[2] - u[0], dm = m[2] - m[0], dd = d[2] - d[0];
        float tx = fmaf(dm, 2.0f, du + dd) * kg;
        const float vl = d[0] - u[0], vc = d[1] - u[1], vr = d[2] - u[2];
        float ty = fmaf(vc, 2.0f, vl + vr) * kg;
        const bool on = (g.on >> i) & 1u;
        tx = on ? tx : 0.0f;
        ty = on ? ty : 0.0f;
        tl.t[i] = m[1] * ks;
        tl.tx[i] = tx;
        tl.ty[i] = ty;
        sxx = fmaf(tx, tx, sxx);
        sxy = fmaf(tx, ty, sxy);
        syy = fmaf(ty, ty, syy);
    }
}

// One iteration's (RES = false: b = totals of It Tx, It Ty; iterate_vx, lk_vx.c:100-140) or the residual's (RES = true:
// totals of |It| / 64 and (It / 64)^2 over the lane's own rows; residual_vx, lk_vx.c:142-175) sums from the lane's
// x-interpolated J rows A[0 .. MAXR] (units of `ks` Q6): J = y-interpolation of consecutive rows, It = T - J.
template <int MAXR, bool RES>
__device__ __forceinline__ void mismatch(const float (&A)[MAXR + 1], float fy, float ks, const Geo &g,
                                         const Tile<MAXR> &tl, float *b)
{
    float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
    for (int i = 0; i < MAXR; i++) {
        const float jv = fmaf(A[i + 1] - A[i], fy, A[i]);
        const float it = fmaf(jv, -ks, tl.t[i]); // T - ks J, the product exact
        if (RES) {
            const bool on = (g.on >> i) & 1u;
            const float q = it * (1.0f / 64.0f);
            s0 = on ? s0 + fabsf(q) : s0;
            s1 = on ? fmaf(q, q, s1) : s1;
        } else {
            s0 = fmaf(it, tl.tx[i], s0);
            s1 = fmaf(it, tl.ty[i], s1);
        }
    }
    b[0] = wave_sum(s0);
    b[1] = wave_sum(s1);
}

// ---- the direct schedule

// The template at (px, py), straight from the level into registers: returns the wave totals of Tx^2, Tx Ty, Ty^2.
template <int MAXR, typename T>
__device__ __forceinline__ void build_direct(const Level &lv, double px, double py, const Geo &g, Tile<MAXR> &tl,
                                             float *G)
{
    const double flx = floor(px), fly = floor(py);
    const float fx = (float)(px - flx), fy = (float)(py - fly);
    // S row 0, column 0 samples level pixel (flx - w/2 - 1 + fx, fly - h/2 - 1 + fy); this lane needs S columns
    // k .. k + 2 of S rows re .. re + MAXR + 1, i.e. image rows re .. re + MAXR + 2 (at most h + 2) and 4 columns
    unsigned pitch;
    const __amdgpu_buffer_rsrc_t rs = level_rsrc<T>(lv, pitch);
    const unsigned x0 = (unsigned)((int)flx - g.W / 2 - 1 + g.k) * (unsigned)sizeof(T);
    const int y0 = (int)fly - g.H / 2 - 1;
    float X[MAXR + 3][3];
    if (g.H >= MAXR) {
        // rows re .. re + MAXR + 2 lie inside the window's h + 3 rows: one lane offset, the row steps uniform
        const unsigned o = (unsigned)(y0 + g.re) * pitch + x0;
#pragma unroll
        for (int i = 0; i < MAXR + 3; i++) {
            float a[4];
            Row<T>::quad(rs, o, (unsigned)i * pitch, a);
            X[i][0] = fmaf(a[1] - a[0], fx, a[0]);
            X[i][1] = fmaf(a[2] - a[1], fx, a[1]);
            X[i][2] = fmaf(a[3] - a[2], fx, a[2]);
        }
    } else {
#pragma unroll
        for (int i = 0; i < MAXR + 3; i++) {
            const int j = g.re + i < g.H + 2 ? g.re + i : g.H + 2;
            float a[4];
            Row<T>::quad(rs, (unsigned)(y0 + j) * pitch + x0, 0u, a);
            X[i][0] = fmaf(a[1] - a[0], fx, a[0]);
            X[i][1] = fmaf(a[2] - a[1], fx, a[1]);
            X[i][2] = fmaf(a[3] - a[2], fx, a[2]);
        }
    }
    float S[MAXR + 2][3];
#pragma unroll
    for (int i = 0; i < MAXR + 2; i++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            S[i][c] = fmaf(X[i + 1][c] - X[i][c], fy, X[i][c]);
    float sxx = 0.0f, sxy = 0.0f, syy = 0.0f;
    sobel_rows<MAXR>(S, Ld<T>::scale, g, tl, sxx, sxy, syy);
    G[0] = wave_sum(sxx);
    G[1] = wave_sum(sxy);
    G[2] = wave_sum(syy);
}

// J at (qx, qy) straight from the level: rows re .. re + MAXR (at most h) of the window at floor(q), 2 taps each.
template <int MAXR, typename T, bool RES>
__device__ __forceinline__ void sample_direct(const Level &lv, double qx, double qy, const Geo &g,
                                              const Tile<MAXR> &tl, float *b)
{
    const double flx = floor(qx), fly = floor(qy);
    const float fx = (float)(qx - flx), fy = (float)(qy - fly);
    unsigned pitch;
    const __amdgpu_buffer_rsrc_t rs = level_rsrc<T>(lv, pitch);
    const unsigned x = (unsigned)((int)flx - g.W / 2 + g.k) * (unsigned)sizeof(T);
    const int y0 = (int)fly - g.H / 2;
    float A[MAXR + 1];
    if (g.H >= MAXR) {
        // rows re .. re + MAXR lie inside the window's h + 1 rows: one lane offset, the row steps uniform
        const unsigned o = (unsigned)(y0 + g.re) * pitch + x;
#pragma unroll
        for (int i = 0; i <= MAXR; i++) {
            float a0, a1;
            Row<T>::pair(rs, o, (unsigned)i * pitch, a0, a1);
            A[i] = fmaf(a1 - a0, fx, a0);
        }
    } else {
#pragma unroll
        for (int i = 0; i <= MAXR; i++) {
            const int j = g.re + i < g.H ? g.re + i : g.H;
            float a0, a1;
            Row<T>::pair(rs, (unsigned)(y0 + j) * pitch + x, 0u, a0, a1);
            A[i] = fmaf(a1 - a0, fx, a0);
        }
    }
    mismatch<MAXR, RES>(A, fy, Ld<T>::scale, g, tl, b);
}

// ---- the staged schedule

// A rectangle of a level copied into LDS: rows ry .. ry + rh - 1, columns rx .. rx + rw - 1, coordinates clamped to the
// level (the reference's replicated border).
struct Region {
    int rx, ry, rw, rh;
};

// The wave's LDS slice: P = the template's patch of I ((h + 3) rows), directly followed by R = J's neighbourhood (both
// rows of the region width w + 1 + 2 m), and S = the bilinear window ((w + 2) x (h + 2)).
struct Slice {
    float *P, *S, *R;
};

// The wave's loads of one level, issued together, as Q6 values: rows 0 .. np - 1 of the combined region come from
// level li at (rp.rx, rp.ry) (the template patch, np = 0 to skip it), the next rg.rh rows from level lj at
// (rg.rx, rg.ry) (J's neighbourhood); every row is rg.rw wide. Element i = lane + 64 k sits at row i / rw, column
// i % rw, stepped without a division per element.
template <typename T, int BATCH>
__device__ __forceinline__ void load_regions(const Level &li, const Region &rp, int np, const Level &lj,
                                             const Region &rg, float *dst, int lane)
{
    const int rw = rg.rw, n = (np + rg.rh) * rw;
    const int dq = 64 / rw, dr = 64 - dq * rw;
    int r = lane / rw, cc = lane - (lane / rw) * rw;
    const T *pi = (const T *)li.p, *pj = (const T *)lj.p;
    wave_sync(); // earlier readers of the slice are done
    for (int base = 0; base < n; base += 64 * BATCH) {
        float v[BATCH];
        int rr = r, c2 = cc;
#pragma unroll
        for (int k = 0; k < BATCH; k++) {
            const bool fromI = rr < np;
            const int y = fromI ? clampi(rp.ry + rr, 0, li.h - 1) : clampi(rg.ry + rr - np, 0, lj.h - 1);
            const int x = fromI ? clampi(rp.rx + c2, 0, li.w - 1) : clampi(rg.rx + c2, 0, lj.w - 1);
            v[k] = Ld<T>::get(fromI ? pi[(size_t)y * li.pitch + x] : pj[(size_t)y * lj.pitch + x]) * Ld<T>::scale;
            c2 += dr;
            rr += dq;
            if (c2 >= rw)
                c2 -= rw, rr++;
        }
#pragma unroll
        for (int k = 0; k < BATCH; k++) {
            const int i = base + lane + 64 * k;
            if (i < n)
                dst[i] = v[k];
        }
        r = rr;
        cc = c2;
    }
    wave_sync();
}

// The template at (px, py) from the staged patch P (origin = the window's corner minus one row and column): S into
// LDS, then the lane's rows as in the direct schedule.
template <int MAXR>
__device__ __forceinline__ void build(const Slice &sl, int PW, float fx, float fy, const Geo &g, int lane,
                                      Tile<MAXR> &tl, float *G)
{
    const int SW = g.W + 2, SH = g.H + 2;
    for (int i = lane; i < SW * SH; i += 64) {
        const int r = i / SW, cc = i - r * SW;
        const float *p = sl.P + r * PW + cc;
        const float top = fmaf(p[1] - p[0], fx, p[0]), bot = fmaf(p[PW + 1] - p[PW], fx, p[PW]);
        sl.S[i] = fmaf(bot - top, fy, top);
    }
    wave_sync();
    float S[MAXR + 2][3];
#pragma unroll
    for (int i = 0; i < MAXR + 2; i++) {
        const int j = g.re + i < SH - 1 ? g.re + i : SH - 1;
#pragma unroll
        for (int c = 0; c < 3; c++)
            S[i][c] = sl.S[j * SW + g.k + c];
    }
    float sxx = 0.0f, sxy = 0.0f, syy = 0.0f;
    sobel_rows<MAXR>(S, 1.0f, g, tl, sxx, sxy, syy);
    G[0] = wave_sum(sxx);
    G[1] = wave_sum(sxy);
    G[2] = wave_sum(syy);
}

// Whether the window at integer corner (x0, y0), with its extra column and row, lies inside the region.
__device__ __forceinline__ bool covers(const Region &rg, int x0, int y0, int W, int H)
{
    return x0 >= rg.rx && y0 >= rg.ry && x0 + W <= rg.rx + rg.rw - 1 && y0 + H <= rg.ry + rg.rh - 1;
}

// J sampled at (qx, qy) from the staged neighbourhood R (Q6 values).
template <int MAXR, bool RES>
__device__ __forceinline__ void sample(const float *R, const Region &rg, double qx, double qy, const Geo &g,
                                       const Tile<MAXR> &tl, float *b)
{
    const double flx = floor(qx), fly = floor(qy);
    const float fx = (float)(qx - flx), fy = (float)(qy - fly);
    const int x = clampi((int)flx - g.W / 2 + g.k - rg.rx, 0, rg.rw - 2);
    const int y0 = (int)fly - g.H / 2 - rg.ry;
    float A[MAXR + 1];
#pragma unroll
    for (int i = 0; i <= MAXR; i++) {
        const int j = g.re + i < g.H ? g.re + i : g.H;
        const float *p = R + clampi(y0 + j, 0, rg.rh - 1) * rg.rw + x;
        A[i] = fmaf(p[1] - p[0], fx, p[0]);
    }
    mismatch<MAXR, RES>(A, fy, 1.0f, g, tl, b);
}

struct Track {
    float x, y, min_eig, residual, cxx, cxy, cyy;
    int iters, status;
};

// The 2x2 system of one level: eigenvalue test and inverse. MIXED: fp32 unless G's condition number passes 1e4.
struct Solve {
    double ia, ib, ic, det, eig;
};

template <int SOLVE>
__device__ __forceinline__ Solve solve2(double gxx, double gxy, double gyy, double area)
{
    Solve s;
    if (SOLVE == VTG_SOLVE_MIXED) {
        const float fxx = (float)gxx, fxy = (float)gxy, fyy = (float)gyy;
        const float tr = fxx + fyy, df = fxx - fyy, rt = sqrtf(df * df + 4.0f * fxy * fxy);
        const float lo = 0.5f * (tr - rt), hi = 0.5f * (tr + rt);
        if (lo > 0.0f && hi < 1e4f * lo) {
            const float det = fxx * fyy - fxy * fxy;
            const float inv = 1.0f / det;
            s.det = det;
            s.eig = (double)(lo / (float)area);
            s.ia = fyy * inv;
            s.ib = -fxy * inv;
            s.ic = fxx * inv;
            return s;
        }
    }
    s.det = gxx * gyy - gxy * gxy;
    const double tr = gxx + gyy, df = gxx - gyy;
    s.eig = (tr - sqrt(df * df + 4.0 * gxy * gxy)) * 0.5 / area;
    const double inv = 1.0 / s.det;
    s.ia = gyy * inv;
    s.ib = -gxy * inv;
    s.ic = gxx * inv;
    return s;
}

// Load the level's template patch (with_patch) and J's neighbourhood together (T: the level's element type).
template <typename T, int BATCH>
__device__ __forceinline__ void level_loads(const Level &li, const Region &rp, const Level &lj, const Region &rg,
                                            const Slice &sl, bool with_patch, int lane)
{
    if (with_patch)
        load_regions<T, BATCH>(li, rp, rp.rh, lj, rg, sl.P, lane);
    else
        load_regions<T, BATCH>(li, rp, 0, lj, rg, sl.R, lane);
}

// One point from pyramid I to J: the reference's level walk (lk_driver.inc:7-109).
template <int MAXR, typename T0, typename T1, int SOLVE, bool STAGED>
__device__ void track(const PyrView &I, const PyrView &J, int top, const Cfg &c, const Geo &g, int lane,
                      const Slice &sl, int margin, float x, float y, float gx0, float gy0, Track &o)
{
    constexpr int BATCH = MAXR <= 8 ? 26 : 16;
    Region rg, rp;
    rg.rw = c.w + 1 + 2 * margin;
    rg.rh = c.h + 1 + 2 * margin;
    rg.rx = rg.ry = INT_MIN; // nothing staged
    rp.rw = rg.rw; // the patch needs w + 3 columns; it shares the region's row width (margin >= 1)
    rp.rh = c.h + 3;
    Tile<MAXR> tl;
    double gx = ldexp((double)gx0, -top), gy = ldexp((double)gy0, -top);
    const double area = (double)(c.w * c.h), eps2 = (double)c.eps * (double)c.eps;
    const int mx = c.w / 2 + 1, my = c.h / 2 + 1;
    float G0[3] = {0.0f, 0.0f, 0.0f}; // level 0's structure tensor, for the covariance
    int status = 0, iters = 0;
    bool valid0 = false;
    float eig0 = 0.0f;
    for (int L = top; L >= 0; L--) {
        const double px = ldexp((double)x, -L), py = ldexp((double)y, -L);
        const Level li = I.l[L], lj = J.l[L]; // copies: the descriptors stay in registers through the iterations
        if (px < mx || py < my || px >= (double)(li.w - mx - 1) || py >= (double)(li.h - my - 1)) {
            if (L == 0)
                status |= LK_OUT;
            else
                gx *= 2, gy *= 2;
            continue;
        }
        float G[3];
        if (STAGED) {
            // the template patch and J around the level's first estimate, one batch
            const double flx = floor(px), fly = floor(py);
            rp.rx = (int)flx - c.w / 2 - 1;
            rp.ry = (int)fly - c.h / 2 - 1;
            rg.rx = (int)floor(px + gx) - c.w / 2 - margin;
            rg.ry = (int)floor(py + gy) - c.h / 2 - margin;
            if (L > 0)
                level_loads<T1, BATCH>(li, rp, lj, rg, sl, true, lane);
            else
                level_loads<T0, BATCH>(li, rp, lj, rg, sl, true, lane);
            build<MAXR>(sl, rg.rw, (float)(px - flx), (float)(py - fly), g, lane, tl, G);
        } else if (L > 0) {
            build_direct<MAXR, T1>(li, px, py, g, tl, G);
        } else {
            build_direct<MAXR, T0>(li, px, py, g, tl, G);
        }
        const double gxx = G[0], gxy = G[1], gyy = G[2];
        const Solve s = solve2<SOLVE>(gxx, gxy, gyy, area);
        if (L == 0)
            eig0 = (float)s.eig;
        if (s.eig < (double)c.min_eig || s.det <= 0) {
            if (L == 0)
                status |= LK_LOWEIG;
            else
                gx *= 2, gy *= 2;
            continue;
        }
        if (L == 0) {
            G0[0] = G[0], G0[1] = G[1], G0[2] = G[2];
            valid0 = true;
        }
        const double lx = (double)(lj.w - mx - 1), ly = (double)(lj.h - my - 1);
        // the inverse with the sums' 1/64 folded in (a power of two: the products round exactly as b / 64 would)
        const double ka = s.ia * (1.0 / 64.0), kb = s.ib * (1.0 / 64.0), kc = s.ic * (1.0 / 64.0);
        double vx = 0, vy = 0;
        for (int it = 0; it < c.max_iter; it++) {
            const double qx = px + gx + vx, qy = py + gy + vy;
            if (qx < mx || qy < my || qx >= lx || qy >= ly) {
                if (L == 0)
                    status |= LK_OUT, valid0 = false;
                break;
            }
            float b[2];
            if (STAGED) {
                const int wx = (int)floor(qx) - c.w / 2, wy = (int)floor(qy) - c.h / 2;
                if (!covers(rg, wx, wy, c.w, c.h)) {
                    rg.rx = wx - margin;
                    rg.ry = wy - margin;
                    if (L > 0)
                        level_loads<T1, BATCH>(li, rp, lj, rg, sl, false, lane);
                    else
                        level_loads<T0, BATCH>(li, rp, lj, rg, sl, false, lane);
                }
                sample<MAXR, false>(sl.R, rg, qx, qy, g, tl, b);
            } else if (L > 0) {
                sample_direct<MAXR, T1, false>(lj, qx, qy, g, tl, b);
            } else {
                sample_direct<MAXR, T0, false>(lj, qx, qy, g, tl, b);
            }
            const double bx = (double)b[0], by = (double)b[1];
            const double ex = ka * bx + kb * by, ey = kb * bx + kc * by;
            vx += ex;
            vy += ey;
            iters++;
            if (ex * ex + ey * ey < eps2)
                break;
        }
        gx += vx;
        gy += vy;
        if (L > 0)
            gx *= 2, gy *= 2;
    }
    const double qx = (double)x + gx, qy = (double)y + gy;
    o.x = (float)qx;
    o.y = (float)qy;
    o.min_eig = eig0;
    o.iters = iters;
    o.residual = 0.0f;
    o.cxx = o.cxy = o.cyy = 0.0f;
    if (valid0 && !status) {
        const Level &l0 = J.l[0];
        if (qx < mx || qy < my || qx >= (double)(l0.w - mx - 1) || qy >= (double)(l0.h - my - 1)) {
            status |= LK_OUT;
        } else {
            float r[2];
            if (STAGED) {
                // level 0 was the last level walked, so rg describes level 0's neighbourhood in R
                const int wx = (int)floor(qx) - c.w / 2, wy = (int)floor(qy) - c.h / 2;
                if (!covers(rg, wx, wy, c.w, c.h)) {
                    rg.rx = wx - margin;
                    rg.ry = wy - margin;
                    level_loads<T0, BATCH>(I.l[0], rp, l0, rg, sl, false, lane);
                }
                sample<MAXR, true>(sl.R, rg, qx, qy, g, tl, r);
            } else {
                sample_direct<MAXR, T0, true>(l0, qx, qy, g, tl, r);
            }
            const double n = area;
            o.residual = (float)((double)r[0] / n);
            if (o.residual > c.max_res)
                status |= LK_RESIDUAL;
            const double s2 = ((double)r[1] / n) * area / (area - 2);
            const double gxx0 = G0[0], gxy0 = G0[1], gyy0 = G0[2], det0 = gxx0 * gyy0 - gxy0 * gxy0;
            o.cxx = (float)(s2 * gyy0 / det0);
            o.cxy = (float)(-s2 * gxy0 / det0);
            o.cyy = (float)(s2 * gxx0 / det0);
        }
    }
    o.status = status;
}

// The direct schedule of windows up to 7 rows per lane fits 128 VGPRs without spilling: bound it to 4 waves per SIMD
// (latency hiding is what the iterations need); the other forms keep the compiler's allocation.
template <int MAXR, typename T0, typename T1, int SOLVE, bool STAGED>
__global__ __launch_bounds__(512, (!STAGED && MAXR <= 7) ? 4 : 1) void lk_kernel(LkArgs a)
{
    extern __shared__ float smem[];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    // workgroup b runs on XCD b mod 8; XCD x takes the contiguous run of workgroups [x per_xcd, (x + 1) per_xcd)
    const int wg = (int)(blockIdx.x % NXCD) * a.per_xcd + (int)(blockIdx.x / NXCD);
    const int lp = wg * a.waves + wave;
    if (lp >= a.n)
        return;
    // this point's pair: the last whose numbering starts at or before it (uniform per wave)
    int lo = 0, hi = a.npairs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (lp >= a.pre[mid])
            lo = mid;
        else
            hi = mid - 1;
    }
    const int pt = a.start[lo] + (lp - a.pre[lo]);
    const PyrView &I = *a.I[lo], &J = *a.J[lo];
    const int top = a.top[lo];
    const int rw = a.cfg.w + 1 + 2 * a.margin, ph = a.cfg.h + 3;
    float *base = smem + wave * a.lds_floats;
    // layout: S first, then P and R contiguous (one combined region)
    const int sS = (a.cfg.w + 2) * (a.cfg.h + 2);
    const Slice sl{base + sS, base, base + sS + rw * ph};
    const Geo g = geometry<MAXR>(a.cfg, lane);
    const float x = a.xy[2 * pt], y = a.xy[2 * pt + 1];
    const float gx = a.guess ? a.guess[2 * pt] : 0.0f, gy = a.guess ? a.guess[2 * pt + 1] : 0.0f;
    Track o;
    track<MAXR, T0, T1, SOLVE, STAGED>(I, J, top, a.cfg, g, lane, sl, a.margin, x, y, gx, gy, o);
    if (a.fb_max > 0.0f && o.status == 0) {
        // back from J to I, the template rebuilt from J (lk_driver.inc:111-121)
        Track b;
        const int fb_top = a.fb_top >= 0 ? a.fb_top : top;
        track<MAXR, T0, T1, SOLVE, STAGED>(J, I, fb_top, a.cfg, g, lane, sl, a.margin, o.x, o.y,
                                           (float)((double)x - (double)o.x), (float)((double)y - (double)o.y), b);
        const double ex = (double)b.x - (double)x, ey = (double)b.y - (double)y;
        if (b.status || ex * ex + ey * ey > (double)a.fb_max * (double)a.fb_max)
            o.status |= LK_FLOWBACK;
    }
    if (lane == 0) {
        lk_out r;
        r.x = o.x;
        r.y = o.y;
        r.min_eig = o.min_eig;
        r.residual = o.residual;
        r.cov_xx = o.cxx;
        r.cov_xy = o.cxy;
        r.cov_yy = o.cyy;
        r.iters = (uint16_t)o.iters;
        r.status = (uint16_t)o.status;
        a.out[pt] = r;
    }
}

template <int MAXR, int SOLVE, bool STAGED>
int dispatch_storage(vtg_ctx *c, const LkArgs &a, int storage, dim3 grid, dim3 block, size_t lds)
{
    switch (storage) {
    case VTG_STORE_Q6:
        hipLaunchKernelGGL((lk_kernel<MAXR, uint8_t, uint16_t, SOLVE, STAGED>), grid, block, lds, c->stream, a);
        break;
    case VTG_STORE_F32:
        hipLaunchKernelGGL((lk_kernel<MAXR, float, float, SOLVE, STAGED>), grid, block, lds, c->stream, a);
        break;
    case VTG_STORE_Q6W:
        hipLaunchKernelGGL((lk_kernel<MAXR, uint16_t, uint16_t, SOLVE, STAGED>), grid, block, lds, c->stream, a);
        break;
    case VTG_STORE_Q6F:
        hipLaunchKernelGGL((lk_kernel<MAXR, uint8_t, float, SOLVE, STAGED>), grid, block, lds, c->stream, a);
        break;
    case VTG_STORE_F32_U8:
        hipLaunchKernelGGL((lk_kernel<MAXR, float, float, SOLVE, STAGED>), grid, block, lds, c->stream, a);
        break;
    case VTG_STORE_F16:
        hipLaunchKernelGGL((lk_kernel<MAXR, uint8_t, __half, SOLVE, STAGED>), grid, block, lds, c->stream, a);
        break;
    case VTG_STORE_BF16:
        hipLaunchKernelGGL((lk_kernel<MAXR, uint8_t, __hip_bfloat16, SOLVE, STAGED>), grid, block, lds, c->stream,
                           a);
        break;
    default:
        return fail(c, VTG_EINVAL, "storage %d", storage);
    }
    VTG_HIP(c, hipGetLastError());
    return VTG_OK;
}

// ---------------------------------------------------------------- fused pyramid + LK, one persistent launch

// The batched tk_build + OP_LK (track.c:584-629, 275-292) as one launch: work items in the order of their
// dependencies: bands of 16 output rows of every pair's level 1 (which also write the frames into level 0), then of
// every pair's level 2, ..., then the points. Workgroups take items from a counter in that order (a workgroup's first
// point turns its four waves into independent takers of one point each, as no band follows). A level-l band waits until
// its pyramid's level l - 1 is complete; a wave's first point waits until every pyramid is complete, then the wave
// acquires once for all its points (an agent-scope acquire invalidates the XCD's L2, which LK's overlapping windows
// live in, so one per pair change costs far more than the wait). Every item waited for was taken before the waiter's
// own, by a workgroup that is running and waits only for earlier items, so the earliest unfinished item always
// progresses. There is one queue per XCD (workgroup w serves queue w mod 8, the XCD it runs on), each in that same order
// of classes (level 1, level 2, ..., then the points): a level's bands are dealt round-robin over the queues and the
// queue's points are the q-th of 8 contiguous runs of all points, so an XCD's LK windows stay in a few pairs. The
// earliest unfinished class has all its dependencies finished and its items in every queue either finished or held by
// running workgroups (the grid is at most what is resident at once), so the queues cannot deadlock either.
// A band publishes with an agent-scope release (every thread) before its counts, a waiter reads a count relaxed and then
// acquires, so levels written on any XCD are seen fresh. A wait gives up after a bounded number of sleeps, setting the
// context's error word (the results are then not valid), so every wave reaches the exit.
constexpr int FB = 64;          // pairs per fused launch
constexpr int FBAND = 16;       // output rows per pyramid item (a band across the whole level)
constexpr int FQ = NXCD;        // work queues: one per XCD (workgroup w serves queue w mod 8, the XCD it runs on)
constexpr int FERR = FQ;        // the error word
constexpr int FTOT = FQ + 1;    // bands finished, all pyramids
constexpr int FCTR = 16;        // counter words before the done counts: [0, FQ) each queue's next item, [FERR] error,
                                // [FTOT] bands finished

struct FusedArgs {
    LkArgs lk;                  // pairs' pyramids, top levels, point slices, window configuration, output
    const uint8_t *f0[FB], *f1[FB];
    int p0[FB], p1[FB];         // frame pitches (bytes)
    int total_bands;            // bands of every pyramid's levels 1 ..
    int *ctr;                   // FCTR words, then done[pair][side][level]
};

__device__ __forceinline__ int level_bands(const Level &L) { return (L.h + FBAND - 1) / FBAND; }

// Wait until *p >= need, called by a whole wave (every lane reads the same word; the value each test uses is lane 0's,
// so the loop and its exit are uniform): sleeps between reads; gives up (setting the error word) after 2^24 sleeps, or
// at once once another wait has given up, so a launch with a broken wait still ends.
__device__ __forceinline__ void wait_count(int *p, int need, int *err)
{
    for (int spins = 0;
         __builtin_amdgcn_readfirstlane(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < need;
         spins++) {
        __builtin_amdgcn_s_sleep(2);
        if ((spins & 1023) == 1023 &&
            __builtin_amdgcn_readfirstlane(__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
            break;
        if (spins == (1 << 24)) {
            __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
    }
}

// Add `add` to *p once for the calling wave (lane 0's operand; the others add 0) and return the old value, uniform.
// Called by whole waves only, so no branch on the lane index sits in the persistent loops: a loop whose back edge
// passes a lane-0-only block can be compiled into one that lane 0 leaves alone while the other lanes go round.
__device__ __forceinline__ int wave_fetch_add(int *p, int add, int lane)
{
    return __builtin_amdgcn_readfirstlane(
        __hip_atomic_fetch_add(p, lane == 0 ? add : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// The counters of a fused launch to zero; the error word only when clear_err (the context's first launch): it stays
// set until the host reads it.
__global__ __launch_bounds__(256) void fused_reset(int *ctr, int n, int clear_err)
{
    for (int i = threadIdx.x; i < n; i += 256)
        if (i != FERR || clear_err)
            ctr[i] = 0;
}

// Bands of a level of one side in queue q: b = q, q + 8, ...
__device__ __forceinline__ int queue_bands(int nb, int q) { return nb > q ? (nb - q + FQ - 1) / FQ : 0; }


template <int MAXR, int SOLVE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void fused_kernel(FusedArgs a)
{
    __shared__ int s_item;
    // this workgroup's queue: first the bands of every pair's level 1, then of level 2, ... (qa[(l - 1) np + k]: the
    // first item of level l of pair k, both sides), then the q-th of 8 contiguous runs of all points (pair by pair)
    __shared__ int qa[(kMaxLevels - 1) * FB + 1];
    __shared__ int lds[rt::tile_lds_elems<FBAND>()];
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int np = a.lk.npairs, q = blockIdx.x % FQ;
    int *err = a.ctr + FERR, *done = a.ctr + FCTR;
    int nl = 1; // the most levels of any pair
    for (int k = 0; k < np; k++)
        nl = a.lk.I[k]->levels > nl ? a.lk.I[k]->levels : nl;
    if (tid == 0) {
        int acc = 0;
        for (int l = 1; l < nl; l++)
            for (int k = 0; k < np; k++) {
                qa[(l - 1) * np + k] = acc;
                const PyrView &I = *a.lk.I[k], &J = *a.lk.J[k];
                if (l < I.levels)
                    acc += queue_bands(level_bands(I.l[l]), q) + queue_bands(level_bands(J.l[l]), q);
            }
        qa[(nl - 1) * np] = acc;
    }
    __syncthreads();
    const int npa = (nl - 1) * np, nbands = __builtin_amdgcn_readfirstlane(qa[npa]);
    const int run = (a.lk.n + FQ - 1) / FQ, pfirst = q * run;
    const int plast = pfirst + run < a.lk.n ? pfirst + run : a.lk.n;
    const int total = nbands + (plast > pfirst ? plast - pfirst : 0);
    for (;;) {
        if (wave == 0) // uniform: wave 0 takes the next item and every lane of it stores the same value
            s_item = wave_fetch_add(a.ctr + q, 1, lane);
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads(); // s_item is rewritten next round
        if (item >= total)
            break;
        if (item < nbands) {
            // a band: its level and pair from the last entry of qa at or before it, then its side
            int lo = 0, hi = npa - 1;
            while (lo < hi) { // values read from LDS are per-lane to the compiler: keep the search in scalars
                const int mid = (lo + hi + 1) >> 1;
                if (item >= __builtin_amdgcn_readfirstlane(qa[mid]))
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const int k = lo % np, l = lo / np + 1;
            int loc = item - __builtin_amdgcn_readfirstlane(qa[lo]);
            const int ni = queue_bands(level_bands(a.lk.I[k]->l[l]), q), sd = loc >= ni ? 1 : 0;
            const PyrView &P = sd ? *a.lk.J[k] : *a.lk.I[k];
            loc = q + FQ * (loc - sd * ni); // the band of the level
            int *cnt = done + (k * 2 + sd) * kMaxLevels;
            if (l >= 2) {
                if (wave == 0)
                    wait_count(cnt + l - 1, level_bands(P.l[l - 1]), err);
                __syncthreads();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            rt::RJob job;
            job.copy8 = nullptr;
            job.cp8 = 0;
            job.dst = (void *)P.l[l].p;
            job.dw = P.l[l].w;
            job.dh = P.l[l].h;
            job.dp = P.l[l].pitch;
            job.tiles_x = (P.l[l].w + 63) / 64;
            job.t0 = 0;
            // the band's tiles, left to right (tile rows are FBAND = 16 output rows)
            if (l == 1) {
                job.src = sd ? a.f1[k] : a.f0[k];
                job.sp = sd ? a.p1[k] : a.p0[k];
                job.sw = P.l[0].w;
                job.sh = P.l[0].h;
                job.copy = (void *)P.l[0].p;
                job.cp = P.l[0].pitch;
                for (int tx = 0; tx < job.tiles_x; tx++) {
                    rt::reduce_tile<uint8_t, uint16_t, uint8_t, true, true, FBAND>(job, loc * job.tiles_x + tx, lds);
                    __syncthreads(); // the next tile reuses the LDS
                }
            } else {
                job.src = P.l[l - 1].p;
                job.sp = P.l[l - 1].pitch;
                job.sw = P.l[l - 1].w;
                job.sh = P.l[l - 1].h;
                job.copy = nullptr;
                job.cp = 0;
                for (int tx = 0; tx < job.tiles_x; tx++) {
                    rt::reduce_tile<uint16_t, uint16_t, uint16_t, false, true, FBAND>(job, loc * job.tiles_x + tx,
                                                                                      lds);
                    __syncthreads();
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __syncthreads();
            if (wave == 0) {
                wave_fetch_add(cnt + l, 1, lane);
                wave_fetch_add(a.ctr + FTOT, 1, lane);
            }
            continue;
        }
        // Every later item of this queue is an LK item (one point): the waves now take them one at a time, each on its
        // own (wave 0 starts with the item just taken), and no longer meet at barriers.
        int it = wave == 0 ? item : -1;
        bool ready = false;
        for (;;) {
            if (it < 0)
                it = wave_fetch_add(a.ctr + q, 1, lane);
            if (it >= total)
                return;
            const int gp = pfirst + (it - nbands); // the point in the numbering of all pairs' points
            int lo = 0, hi = np - 1; // its pair: the last whose numbering starts at or before it
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (gp >= a.lk.pre[mid])
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const int k = lo;
            const PyrView &I = *a.lk.I[k], &J = *a.lk.J[k];
            const int top = a.lk.top[k];
            if (!ready) {
                // every pyramid complete, then one acquire for all the wave's points (an agent-scope acquire
                // invalidates the XCD's L2, which LK's overlapping windows live in)
                wait_count(a.ctr + FTOT, a.total_bands, err);
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                ready = true;
            }
            const int pt = a.lk.start[k] + (gp - a.lk.pre[k]);
            it = -1;
            const Geo g = geometry<MAXR>(a.lk.cfg, lane);
            const float x = a.lk.xy[2 * pt], y = a.lk.xy[2 * pt + 1];
            const float gx = a.lk.guess ? a.lk.guess[2 * pt] : 0.0f, gy = a.lk.guess ? a.lk.guess[2 * pt + 1] : 0.0f;
            const Slice sl{nullptr, nullptr, nullptr};
            Track o;
            track<MAXR, uint8_t, uint16_t, SOLVE, false>(I, J, top, a.lk.cfg, g, lane, sl, 1, x, y, gx, gy, o);
            if (a.lk.fb_max > 0.0f && o.status == 0) {
                Track b;
                track<MAXR, uint8_t, uint16_t, SOLVE, false>(J, I, top, a.lk.cfg, g, lane, sl, 1, o.x, o.y,
                                                             (float)((double)x - (double)o.x),
                                                             (float)((double)y - (double)o.y), b);
                const double ex = (double)b.x - (double)x, ey = (double)b.y - (double)y;
                if (b.status || ex * ex + ey * ey > (double)a.lk.fb_max * (double)a.lk.fb_max)
                    o.status |= LK_FLOWBACK;
            }
            if (lane == 0) {
                lk_out r;
                r.x = o.x;
                r.y = o.y;
                r.min_eig = o.min_eig;
                r.residual = o.residual;
                r.cov_xx = o.cxx;
                r.cov_xy = o.cxy;
                r.cov_yy = o.cyy;
                r.iters = (uint16_t)o.iters;
                r.status = (uint16_t)o.status;
                a.lk.out[pt] = r;
            }
        }
    }
}

template <int MAXR, int SOLVE>
int launch_fused_t(vtg_ctx *c, const FusedArgs &a, int total)
{
    static int per_cu = 0, cus = 0; // resident workgroups per CU of this kernel, CUs of device 0 class
    if (!per_cu) {
        VTG_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fused_kernel<MAXR, SOLVE>, 256, 0));
        VTG_HIP(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
        if (per_cu < 1)
            per_cu = 1;
    }
    // every queue needs workgroups: a multiple of FQ, at least FQ
    const int grid = std::max(FQ, std::min((total + FQ - 1) / FQ * FQ, per_cu * cus / FQ * FQ));
    hipEvent_t ev;
    timer_begin(c, "pyr_lk_fused", &ev);
    hipLaunchKernelGGL((fused_kernel<MAXR, SOLVE>), dim3(grid), dim3(256), 0, c->stream, a);
    VTG_HIP(c, hipGetLastError());
    timer_end(c, "pyr_lk_fused", ev);
    return VTG_OK;
}

template <int SOLVE>
int launch_fused_rows(vtg_ctx *c, const FusedArgs &a, int total, int chunk)
{
    if (chunk <= 4)
        return launch_fused_t<4, SOLVE>(c, a, total);
    if (chunk <= 7)
        return launch_fused_t<7, SOLVE>(c, a, total);
    if (chunk <= 8)
        return launch_fused_t<8, SOLVE>(c, a, total);
    if (chunk <= 16)
        return launch_fused_t<16, SOLVE>(c, a, total);
    return launch_fused_t<32, SOLVE>(c, a, total);
}

// Rows per lane: the smallest compiled count no smaller than the window's chunk.
template <int SOLVE, bool STAGED>
int dispatch_rows(vtg_ctx *c, const LkArgs &a, int storage, int chunk, dim3 grid, dim3 block, size_t lds)
{
    if (chunk <= 4)
        return dispatch_storage<4, SOLVE, STAGED>(c, a, storage, grid, block, lds);
    if (chunk <= 7)
        return dispatch_storage<7, SOLVE, STAGED>(c, a, storage, grid, block, lds);
    if (chunk <= 8)
        return dispatch_storage<8, SOLVE, STAGED>(c, a, storage, grid, block, lds);
    if (chunk <= 16)
        return dispatch_storage<16, SOLVE, STAGED>(c, a, storage, grid, block, lds);
    return dispatch_storage<32, SOLVE, STAGED>(c, a, storage, grid, block, lds);
}

int check_cfg(vtg_ctx *c, const lk_cfg *cfg, int solve_mode)
{
    if (cfg->w < 5 || cfg->w > LK_MAX_W || cfg->h < 5 || cfg->h > LK_MAX_H || cfg->max_iter < 0)
        return fail(c, VTG_EINVAL, "vtg_lk_track: window %dx%d (5..%d x 5..%d), max_iter %d", cfg->w, cfg->h,
                    LK_MAX_W, LK_MAX_H, cfg->max_iter);
    if (solve_mode != VTG_SOLVE_F64 && solve_mode != VTG_SOLVE_MIXED)
        return fail(c, VTG_EINVAL, "vtg_lk_track: solve mode %d", solve_mode);
    return VTG_OK;
}

int check_pair(vtg_ctx *c, const vtg_pyr *I, const vtg_pyr *J, int top)
{
    if (!I || !J)
        return fail(c, VTG_EINVAL, "vtg_lk_track: null pyramid");
    if (top < 0 || top >= I->levels || top >= J->levels)
        return fail(c, VTG_EINVAL, "vtg_lk_track: top %d with %d / %d levels", top, I->levels, J->levels);
    if (I->storage != J->storage)
        return fail(c, VTG_EINVAL, "vtg_lk_track: pyramids of different storage");
    for (int l = 0; l <= top; l++)
        if (I->lv[l].w != J->lv[l].w || I->lv[l].h != J->lv[l].h)
            return fail(c, VTG_EINVAL, "vtg_lk_track: level %d sizes differ", l);
    return VTG_OK;
}

}  // namespace

// One launch for pairs sel[0 .. m) (m <= BMAX; indices into the caller's arrays) under schedule s.
static int launch_group(vtg_ctx *c, const std::vector<int> &sel, const vtg_pyr *const *I, const vtg_pyr *const *J,
                        const int *top, const int *off, const lk_cfg *cfg, const float *xy, const float *guess,
                        float fb_max, int solve_mode, lk_out *out, LkSched s, int fb_top)
{
    LkArgs a;
    a.npairs = (int)sel.size();
    int acc = 0;
    for (int q = 0; q < a.npairs; q++) {
        const int k = sel[q];
        a.I[q] = I[k]->dview;
        a.J[q] = J[k]->dview;
        a.top[q] = top[k];
        a.start[q] = off[k];
        a.pre[q] = acc;
        acc += off[k + 1] - off[k];
    }
    a.pre[a.npairs] = acc;
    a.n = acc;
    if (a.n <= 0)
        return VTG_OK;
    a.cfg = Cfg{cfg->w, cfg->h, cfg->max_iter, cfg->eps, cfg->min_eig, cfg->max_residual};
    a.margin = s.margin > 0 ? s.margin : 1;
    // the staged schedule's slice per wave: S, then the patch and J's neighbourhood; the direct one uses no LDS
    a.lds_floats = s.margin > 0 ? (cfg->w + 2) * (cfg->h + 2) +
                                      (cfg->w + 1 + 2 * a.margin) * (cfg->h + 3 + cfg->h + 1 + 2 * a.margin)
                                : 0;
    a.xy = xy;
    a.guess = guess;
    a.fb_max = fb_max;
    a.fb_top = fb_top;
    a.out = out;
    a.waves = s.waves;
    const int groups = 64 / cfg->w, chunk = (cfg->h + groups - 1) / groups;
    const int wgs = (a.n + a.waves - 1) / a.waves;
    a.per_xcd = (wgs + NXCD - 1) / NXCD;
    const dim3 grid(a.per_xcd * NXCD), block(64 * a.waves);
    const size_t lds = (size_t)a.waves * a.lds_floats * sizeof(float);
    if (s.waves < 1 || s.waves > 8 || lds > 65536)
        return fail(c, VTG_EINVAL, "vtg_lk_track: schedule margin %d x %d waves needs %zu B of LDS (at most 65536)",
                    s.margin, s.waves, lds);
    hipEvent_t ev;
    timer_begin(c, "lk_track", &ev);
    const bool staged = s.margin > 0;
    int rc;
    if (solve_mode == VTG_SOLVE_MIXED)
        rc = staged ? dispatch_rows<VTG_SOLVE_MIXED, true>(c, a, I[sel[0]]->storage, chunk, grid, block, lds)
                    : dispatch_rows<VTG_SOLVE_MIXED, false>(c, a, I[sel[0]]->storage, chunk, grid, block, lds);
    else
        rc = staged ? dispatch_rows<VTG_SOLVE_F64, true>(c, a, I[sel[0]]->storage, chunk, grid, block, lds)
                    : dispatch_rows<VTG_SOLVE_F64, false>(c, a, I[sel[0]]->storage, chunk, grid, block, lds);
    timer_end(c, "lk_track", ev);
    return rc;
}

uint64_t lk_class(const vtg_pyr *I, int top, const lk_cfg *cfg)
{
    return ((uint64_t)I->w << 48) ^ ((uint64_t)I->h << 32) ^ ((uint64_t)top << 28) ^ ((uint64_t)I->storage << 24) ^
           ((uint64_t)cfg->w << 12) ^ (uint64_t)cfg->h;
}

LkSched tuned_lk(const vtg_ctx *c, const vtg_pyr *I, int top, const lk_cfg *cfg)
{
    const auto it = c->lk_sched.find(lk_class(I, top, cfg));
    if (it != c->lk_sched.end())
        return it->second;
    return LkSched{c->lk_margin, 2};
}

// Pairs [0, np) with points off[k] .. off[k+1] - 1 of xy: grouped by schedule, one launch per group of up to BMAX.
int launch_lk_many(vtg_ctx *c, int np, const vtg_pyr *const *I, const vtg_pyr *const *J, const int *top,
                   const int *off, const lk_cfg *cfg, const float *xy, const float *guess, float fb_max,
                   int solve_mode, lk_out *out, const LkSched *sched, int fb_top)
{
    std::vector<LkSched> sc(np);
    for (int k = 0; k < np; k++)
        sc[k] = sched ? *sched : tuned_lk(c, I[k], top[k], cfg);
    std::vector<bool> done(np, false);
    for (int k = 0; k < np; k++) {
        if (done[k])
            continue;
        std::vector<int> sel;
        for (int q = k; q < np; q++)
            if (!done[q] && sc[q].margin == sc[k].margin && sc[q].waves == sc[k].waves) {
                sel.push_back(q);
                done[q] = true;
            }
        for (size_t b = 0; b < sel.size(); b += BMAX) {
            const std::vector<int> part(sel.begin() + b, sel.begin() + std::min(sel.size(), b + BMAX));
            int rc = launch_group(c, part, I, J, top, off, cfg, xy, guess, fb_max, solve_mode, out, sc[k], fb_top);
            if (rc)
                return rc;
        }
    }
    return VTG_OK;
}

}  // namespace vtg

using namespace vtg;

extern "C" {

int vtg_lk_track_ex(vtg_ctx *c, const vtg_pyr *I, const vtg_pyr *J, int top, const struct lk_cfg *cfg, int n,
                    const float *xy, const float *guess, float fb_max, int solve_mode, struct lk_out *out)
{
    if (!c || !cfg || (n > 0 && (!xy || !out)) || n < 0)
        return fail(c, VTG_EINVAL, "vtg_lk_track: bad argument");
    int rc = check_cfg(c, cfg, solve_mode);
    if (!rc)
        rc = check_pair(c, I, J, top);
    if (rc || n == 0)
        return rc;
    const int off[2] = {0, n};
    return launch_lk_many(c, 1, &I, &J, &top, off, cfg, xy, guess, fb_max, solve_mode, out, nullptr);
}

int vtg_lk_track(vtg_ctx *c, const vtg_pyr *I, const vtg_pyr *J, int top, const struct lk_cfg *cfg, int n,
                 const float *xy, const float *guess, float fb_max, struct lk_out *out)
{
    return vtg_lk_track_ex(c, I, J, top, cfg, n, xy, guess, fb_max, VTG_SOLVE_F64, out);
}

int vtg_lk_track_many(vtg_ctx *c, int npairs, const vtg_pyr *const *I, const vtg_pyr *const *J, const int *top,
                      const int *offsets, const struct lk_cfg *cfg, const float *xy, const float *guess, float fb_max,
                      int solve_mode, struct lk_out *out)
{
    if (!c || npairs < 0 || (npairs > 0 && (!I || !J || !top || !offsets)) || !cfg)
        return fail(c, VTG_EINVAL, "vtg_lk_track_many: bad argument");
    int rc = check_cfg(c, cfg, solve_mode);
    if (rc)
        return rc;
    for (int k = 0; k < npairs; k++) {
        if ((rc = check_pair(c, I[k], J[k], top[k])))
            return rc;
        if (offsets[k + 1] < offsets[k] || I[k]->storage != I[0]->storage)
            return fail(c, VTG_EINVAL, "vtg_lk_track_many: pair %d: offsets or storage", k);
    }
    if (npairs == 0 || offsets[npairs] == offsets[0])
        return VTG_OK;
    if (!xy || !out)
        return fail(c, VTG_EINVAL, "vtg_lk_track_many: null points or output");
    return launch_lk_many(c, npairs, I, J, top, offsets, cfg, xy, guess, fb_max, solve_mode, out, nullptr);
}

int vtg_pyr_lk_fused(vtg_ctx *c, int npairs, vtg_pyr *const *I, vtg_pyr *const *J, const uint8_t *const *frame0,
                     const uint8_t *const *frame1, const int *pitch0, const int *pitch1, const int *top,
                     const int *offsets, const struct lk_cfg *cfg, const float *xy, const float *guess, float fb_max,
                     int solve_mode, struct lk_out *out)
{
    if (!c || !cfg || npairs < 0 || npairs > FB ||
        (npairs > 0 && (!I || !J || !frame0 || !frame1 || !pitch0 || !pitch1 || !top || !offsets)))
        return fail(c, VTG_EINVAL, "vtg_pyr_lk_fused: bad argument (at most %d pairs)", FB);
    int rc = check_cfg(c, cfg, solve_mode);
    if (rc || npairs == 0)
        return rc;
    FusedArgs a;
    a.total_bands = 0;
    a.lk.npairs = npairs;
    int acc = 0, items = 0;
    for (int k = 0; k < npairs; k++) {
        if ((rc = check_pair(c, I[k], J[k], top[k])))
            return rc;
        if (I[k]->storage != VTG_STORE_Q6 || I[k]->levels < 2 || J[k]->levels != I[k]->levels || !frame0[k] ||
            !frame1[k] || pitch0[k] < I[k]->w || pitch1[k] < J[k]->w || offsets[k + 1] < offsets[k])
            return fail(c, VTG_EINVAL, "vtg_pyr_lk_fused: pair %d: Q6 pyramids of one level count >= 2, frames, "
                                       "pitches and offsets", k);
        a.lk.I[k] = I[k]->dview;
        a.lk.J[k] = J[k]->dview;
        a.lk.top[k] = top[k];
        a.lk.start[k] = offsets[k];
        a.lk.pre[k] = acc;
        acc += offsets[k + 1] - offsets[k];
        a.f0[k] = frame0[k];
        a.f1[k] = frame1[k];
        a.p0[k] = pitch0[k];
        a.p1[k] = pitch1[k];
        for (int l = 1; l < I[k]->levels; l++) {
            const int b = (I[k]->lv[l].h + FBAND - 1) / FBAND + (J[k]->lv[l].h + FBAND - 1) / FBAND;
            items += b;
            a.total_bands += b;
        }
        items += (offsets[k + 1] - offsets[k] + 3) / 4; // points, four waves to a workgroup
    }
    a.lk.pre[npairs] = acc;
    a.lk.n = acc;
    if (acc > 0 && (!xy || !out))
        return fail(c, VTG_EINVAL, "vtg_pyr_lk_fused: null points or output");
    a.lk.cfg = Cfg{cfg->w, cfg->h, cfg->max_iter, cfg->eps, cfg->min_eig, cfg->max_residual};
    a.lk.margin = 1;
    a.lk.lds_floats = 0;
    a.lk.xy = xy;
    a.lk.guess = guess;
    a.lk.fb_max = fb_max;
    a.lk.fb_top = -1;
    a.lk.out = out;
    a.lk.waves = 4;
    a.lk.per_xcd = 0;
    const int nctr = FCTR + FB * 2 * kMaxLevels;
    if (!c->fused)
        VTG_HIP(c, hipMalloc(&c->fused, nctr * sizeof(int) + 256));
    a.ctr = (int *)c->fused;
    // the counters cleared by a kernel (no memset node, so the launch can be captured in a graph)
    hipLaunchKernelGGL(fused_reset, dim3(1), dim3(256), 0, c->stream, a.ctr, nctr, c->fused_init ? 0 : 1);
    c->fused_init = true;
    VTG_HIP(c, hipGetLastError());
    const int groups = 64 / cfg->w, chunk = (cfg->h + groups - 1) / groups;
    return solve_mode == VTG_SOLVE_MIXED ? launch_fused_rows<VTG_SOLVE_MIXED>(c, a, items, chunk)
                                         : launch_fused_rows<VTG_SOLVE_F64>(c, a, items, chunk);
}

int vtg_ctx_device_error(vtg_ctx *c, int *code)
{
    if (!c || !code)
        return VTG_EINVAL;
    *code = 0;
    if (!c->fused)
        return VTG_OK;
    VTG_HIP(c, hipStreamSynchronize(c->stream));
    VTG_HIP(c, hipMemcpy(code, (int *)c->fused + FERR, sizeof(int), hipMemcpyDeviceToHost));
    if (*code)
        VTG_HIP(c, hipMemset((int *)c->fused + FERR, 0, sizeof(int)));
    return VTG_OK;
}

int vtg_lk_set_margin(vtg_ctx *c, int margin)
{
    if (!c || margin < 0 || margin > 16)
        return fail(c, VTG_EINVAL, "vtg_lk_set_margin: %d (0..16)", margin);
    c->lk_margin = margin;
    return VTG_OK;
}

}  // extern "C"

extern "C" int vtg_lk_track_planes(vtg_ctx *c, const struct vt_plane *I, const struct vt_plane *J, int top,
                                   const struct lk_cfg *cfg, int n, const float *x, const float *y, const float *gx,
                                   const float *gy, float fb_max, struct lk_out *out)
{
    if (!c || !I || !J || !cfg || n < 0 || (n > 0 && (!x || !y || !out)) || top < 0 || top >= VTG_MAX_LEVELS)
        return fail(c, VTG_EINVAL, "vtg_lk_track_planes: bad argument");
    if (n == 0)
        return VTG_OK;
    vtg_pyr *pi = nullptr, *pj = nullptr;
    int rc = pyr_from_planes(c, I, top + 1, &pi);
    if (!rc)
        rc = pyr_from_planes(c, J, top + 1, &pj);
    std::vector<float> xy((size_t)2 * n), gs;
    for (int i = 0; i < n; i++)
        xy[2 * i] = x[i], xy[2 * i + 1] = y[i];
    if (gx || gy) {
        gs.resize((size_t)2 * n);
        for (int i = 0; i < n; i++)
            gs[2 * i] = gx ? gx[i] : 0.0f, gs[2 * i + 1] = gy ? gy[i] : 0.0f;
    }
    void *dev = nullptr;
    const size_t xyb = xy.size() * sizeof(float), outb = (size_t)n * sizeof(lk_out);
    if (!rc) {
        hipError_t e = hipMalloc(&dev, 2 * xyb + outb);
        if (e != hipSuccess)
            rc = hip_fail(c, e, "vtg_lk_track_planes: hipMalloc");
    }
    if (!rc) {
        float *dxy = (float *)dev, *dg = gs.empty() ? nullptr : (float *)((char *)dev + xyb);
        lk_out *dout = (lk_out *)((char *)dev + 2 * xyb);
        hipError_t e = hipMemcpyAsync(dxy, xy.data(), xyb, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess && dg)
            e = hipMemcpyAsync(dg, gs.data(), xyb, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess)
            rc = hip_fail(c, e, "vtg_lk_track_planes: upload");
        if (!rc)
            rc = vtg_lk_track(c, pi, pj, top, cfg, n, dxy, dg, fb_max, dout);
        if (!rc) {
            e = hipMemcpyAsync(out, dout, outb, hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess)
                e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess)
                rc = hip_fail(c, e, "vtg_lk_track_planes: download");
        }
    }
    if (dev) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(dev);
    }
    if (pi)
        vtg_pyr_destroy(c, pi);
    if (pj)
        vtg_pyr_destroy(c, pj);
    return rc;
}
