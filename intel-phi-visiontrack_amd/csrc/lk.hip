// lk.hip: sparse pyramidal Lucas-Kanade on the GPU in the reference's own lane order (lk_driver.inc,
// src/kern/lk_driver.inc:7-121; patch operations lk_vx.c:36-175; the OP_LK batch loop, src/kern/track.c:275-292).
//
// Lane layout. A point owns one row of 16 lanes of a wavefront (DPP row), so a wavefront tracks 4 points. Lane j of
// the row is the reference's vector lane j: for each 16-column block k of the window it owns window column 16k + j
// and keeps that column's template values T, Tx, Ty (every window row) in registers for the whole level. Its three
// patch sums run exactly as the reference's lane j runs them: fused multiply-adds over the rows, and within a row over
// the blocks in order (lk_vx.c:82-94, 119-126, 160-168), and the 16 lane sums meet in the reference's butterfly
// (partners i^1, i^2, i^4, i^8; vx_hsum, src/common/vx.h:213-221) through four DPP adds inside the row. Every fp32
// value of the reference's patch arithmetic is produced by the same operations on the same inputs: the bilinear
// window S (x-interpolation per image row, y-interpolation between rows), the Sobel gradients / 512 (a context may
// choose the Scharr variant instead, vtg_lk_set_gradient, which has no reference counterpart), It = T - J, the
// residual's |It| / 64 and (It / 64)^2. So the structure tensor, the mismatch vector and the residual sums equal the
// reference's bit for bit. Lanes past the window's width (the last block of a window narrower than a whole block) get
// zero gradients as in the reference (lk_vx.c:82-85) and read a column inside the window, so every value they touch is
// finite.
// The driver (level walk, eigenvalue test, cached inverse, 2x2 solve, convergence, bounds, covariance) runs in fp64
// per lane on those sums; the reference's runs in x87 long double. The positions it produces are converted to the
// float bilinear weights and integer corners exactly as the reference converts its own, and the residual's means are
// rounded through a 64-bit significand first, as x87 rounds them (the exact path of ldnorm.hpp when a double
// quotient is within reach of a float rounding boundary). Where fp64 and x87 would round a driver step differently
// the records can differ; the tests compare every record with the reference's.
// All 16 lanes of a row hold identical sums, so a point's control flow is uniform across its row; the 4 points of a
// wavefront diverge by exec mask (a point done with a level waits for the others of its wavefront).
// Registers and LDS. The windows of the workloads here (15 x 15, 21 x 21) compile as exact shapes: straight-line code,
// the template of every full 16-column block in registers, the last, partly used block (5 columns at 21 x 21) in LDS,
// and each level's J neighbourhood (the window plus 2 pixels each side) staged in LDS once, so the iterations read
// LDS rather than the level in memory (a window that leaves the neighbourhood stages it again). Any other window runs
// a guarded form of its block count that reads the level in memory at every iteration. The loads of both forms are
// pipelined a fixed number of rows ahead of their use.
// Batches. A launch covers points of up to 64 frame pairs of any sizes; a wavefront's 4 points always belong to one
// pair (each pair's points are numbered from a multiple of 4), so its pyramid levels are wave-uniform and every level
// is read through one buffer resource with the row step in a scalar offset. Workgroups are mapped to points so that
// each of the 8 XCDs (which take workgroups in turn) works on one contiguous run of points, keeping neighbouring
// windows in one L2.
// Element types. Level values are read in their own units (u8: grey, u16 and f32: Q6, f16/bf16: grey) and carried at
// that scale through the interpolation; the Q6 factor (a power of two) is applied where it is exact, so the results
// equal those of interpolating Q6 values.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <climits>
#include <type_traits>

#include "ldnorm.hpp"
#include "reduce_tile.hpp"
#include "vtg_internal.hpp"

// No contraction of a product and a sum into one fused operation anywhere below: the driver's fp64 expressions round
// each operation as the reference's x87 does (fused operations only where the reference's vector code fuses them,
// written as fmaf).
#pragma clang fp contract(off)

namespace vtg {

namespace {

struct Cfg {
    int w, h, max_iter;
    float eps, min_eig, max_res;
    int grad; // VTG_GRAD_SOBEL or VTG_GRAD_SCHARR
};

// The gradient operator's weights: outer and centre rows of the smoothing across, and the factor after the sum
// (per stored unit: the storage scale times 1/512 for Sobel's [1 2 1] / 8, 1/2048 for Scharr's [3 10 3] / 32, both per
// Q6 unit). Sobel's outer weight 1 makes its product exact, so its values are the reference's bit for bit.
struct GradW {
    float we, wc, k;
};
__device__ __forceinline__ GradW grad_w(int grad, float kg)
{
    const bool sch = grad == VTG_GRAD_SCHARR;
    return GradW{sch ? 3.0f : 1.0f, sch ? 10.0f : 2.0f, sch ? kg * 0.25f : kg};
}

constexpr int BMAX = 64; // pairs per launch (inline in the kernel arguments)
constexpr int NXCD = 8;  // XCDs of the device; workgroup i runs on XCD i mod 8
constexpr int PPW = 4;   // points per wavefront: one per row of 16 lanes

// A launch's pairs: pair q's points are xy[start[q] .. start[q] + cnt[q]) (global indices), numbered
// pre[q] .. pre[q] + cnt[q] - 1 in the launch; pre[q] is a multiple of PPW and slots up to pre[q + 1] are empty.
struct LkArgs {
    const PyrView *I[BMAX], *J[BMAX];
    int top[BMAX];
    int start[BMAX];
    int cnt[BMAX];
    int pre[BMAX + 1];
    int npairs, slots, per_xcd, waves;
    Cfg cfg;
    int solve;                  // VTG_SOLVE_*
    const float *xy, *guess;
    float fb_max;
    int fb_top;                 // the flowback's top level (-1: the pair's own)
    lk_out *out;
    // the launch's points in spatial order (or null: in the caller's order): slot j of pair q tracks the point
    // order[obase[q] + j], a global index into xy, guess and out
    const int *order;
    int obase[BMAX];
};

// The global index of the point in slot `local` of pair q.
__device__ __forceinline__ int point_at(const LkArgs &a, int q, int local)
{
    return a.order ? a.order[a.obase[q] + local] : a.start[q] + local;
}

// A level element in its own units, and how many Q6 units one of them is.
template <typename T> struct Ld;
template <> struct Ld<uint8_t> {
    static constexpr float scale = 64.0f;
};
template <> struct Ld<uint16_t> {
    static constexpr float scale = 1.0f;
};
template <> struct Ld<float> {
    static constexpr float scale = 1.0f;
};
template <> struct Ld<__half> {
    static constexpr float scale = 64.0f;
    static __device__ __forceinline__ float get(__half v) { return __half2float(v); }
};
template <> struct Ld<__hip_bfloat16> {
    static constexpr float scale = 64.0f;
    static __device__ __forceinline__ float get(__hip_bfloat16 v) { return __bfloat162float(v); }
};

// A level read through a buffer resource: the base and size are wave-uniform (scalar registers), each load is the
// resource plus a lane's 32-bit byte offset o plus a uniform offset so (a row step, kept in a scalar register), and a
// read past the level's bytes returns 0 instead of touching memory. Two (P) or four (Q) neighbouring elements of a row
// come in one load; the raw bits are converted where they are used, so a load can be issued rows ahead of its use.
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
template <typename T> struct Row;
template <> struct Row<uint8_t> {
    typedef unsigned P;
    typedef unsigned Q;
    static __device__ __forceinline__ float l1(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so)
    {
        return (float)__builtin_amdgcn_raw_buffer_load_b8(r, o, so, 0);
    }
    static __device__ __forceinline__ P lp(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so)
    {
        return __builtin_amdgcn_raw_buffer_load_b16(r, o, so, 0);
    }
    static __device__ __forceinline__ void cp(P v, float &a0, float &a1)
    {
        a0 = (float)(v & 0xffu);
        a1 = (float)((v >> 8) & 0xffu);
    }
    static __device__ __forceinline__ Q lq(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so)
    {
        return __builtin_amdgcn_raw_buffer_load_b32(r, o, so, 0);
    }
    static __device__ __forceinline__ void cq(Q v, float *a)
    {
        a[0] = (float)(v & 0xffu);
        a[1] = (float)((v >> 8) & 0xffu);
        a[2] = (float)((v >> 16) & 0xffu);
        a[3] = (float)(v >> 24);
    }
};
template <> struct Row<uint16_t> {
    typedef unsigned P;
    typedef u32x2 Q;
    static __device__ __forceinline__ float l1(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so)
    {
        return (float)__builtin_amdgcn_raw_buffer_load_b16(r, o, so, 0);
    }
    static __device__ __forceinline__ P lp(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so)
    {
        return __builtin_amdgcn_raw_buffer_load_b32(r, o, so, 0);
    }
    static __device__ __forceinline__ void cp(P v, float &a0, float &a1)
    {
        a0 = (float)(v & 0xffffu);
        a1 = (float)(v >> 16);
    }
    static __device__ __forceinline__ Q lq(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so)
    {
        return __builtin_amdgcn_raw_buffer_load_b64(r, o, so, 0);
    }
    static __device__ __forceinline__ void cq(Q v, float *a)
    {
        a[0] = (float)(v[0] & 0xffffu);
        a[1] = (float)(v[0] >> 16);
        a[2] = (float)(v[1] & 0xffffu);
        a[3] = (float)(v[1] >> 16);
    }
};
template <> struct Row<float> {
    typedef u32x2 P;
    typedef u32x4 Q;
    static __device__ __forceinline__ float l1(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so)
    {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, o, so, 0));
    }
    static __device__ __forceinline__ P lp(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so)
    {
        return __builtin_amdgcn_raw_buffer_load_b64(r, o, so, 0);
    }
    static __device__ __forceinline__ void cp(P v, float &a0, float &a1)
    {
        a0 = __uint_as_float(v[0]);
        a1 = __uint_as_float(v[1]);
    }
    static __device__ __forceinline__ Q lq(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so)
    {
        return __builtin_amdgcn_raw_buffer_load_b128(r, o, so, 0);
    }
    static __device__ __forceinline__ void cq(Q v, float *a)
    {
        for (int i = 0; i < 4; i++)
            a[i] = __uint_as_float(v[i]);
    }
};
template <typename H> struct Row16 {
    typedef unsigned P;
    typedef u32x2 Q;
    static __device__ __forceinline__ float cvt(unsigned short b);
    static __device__ __forceinline__ float l1(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so)
    {
        return cvt((unsigned short)__builtin_amdgcn_raw_buffer_load_b16(r, o, so, 0));
    }
    static __device__ __forceinline__ P lp(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so)
    {
        return __builtin_amdgcn_raw_buffer_load_b32(r, o, so, 0);
    }
    static __device__ __forceinline__ void cp(P v, float &a0, float &a1)
    {
        a0 = cvt((unsigned short)(v & 0xffffu));
        a1 = cvt((unsigned short)(v >> 16));
    }
    static __device__ __forceinline__ Q lq(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned so)
    {
        return __builtin_amdgcn_raw_buffer_load_b64(r, o, so, 0);
    }
    static __device__ __forceinline__ void cq(Q v, float *a)
    {
        a[0] = cvt((unsigned short)(v[0] & 0xffffu));
        a[1] = cvt((unsigned short)(v[0] >> 16));
        a[2] = cvt((unsigned short)(v[1] & 0xffffu));
        a[3] = cvt((unsigned short)(v[1] >> 16));
    }
};
template <typename H> __device__ __forceinline__ float Row16<H>::cvt(unsigned short b)
{
    H h;
    __builtin_memcpy(&h, &b, 2);
    return Ld<H>::get(h);
}
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

template <int CTRL>
__device__ __forceinline__ float dpp_add(float v)
{
    const int moved = __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false);
    return v + __int_as_float(moved);
}

// The sum of the 16 lanes of the caller's row in the reference's butterfly order (vx_hsum): partners i^1 and i^2 by
// quad permutes; after those every lane of a quad holds the same value, so the half-row mirror (lane 7 - i) brings the
// value of lane i^4 and the row mirror (lane 15 - i) that of lane i^8. Each step adds the lane's own value and its
// partner's, as the reference does, and fp32 addition commutes, so every lane of the row ends with the same bits.
__device__ __forceinline__ float row_sum(float v)
{
    v = dpp_add<0xb1>(v);  // quad_perm [1,0,3,2]
    v = dpp_add<0x4e>(v);  // quad_perm [2,3,0,1]
    v = dpp_add<0x141>(v); // row_half_mirror
    v = dpp_add<0x140>(v); // row_mirror
    return v;
}

// A lane's columns: per block k, the window column it reads (16k + j, or the window's last column for a lane past the
// width) and whether it is a window column.
template <int NV> struct Lanes {
    unsigned col[NV]; // column offset in elements from the window's first column
    bool on[NV];
    int j;            // the lane in its row
};

template <int NV>
__device__ __forceinline__ Lanes<NV> lanes(const Cfg &c, int j)
{
    Lanes<NV> ln;
    ln.j = j;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        const int cc = 16 * k + j;
        ln.on[k] = cc < c.w;
        ln.col[k] = (unsigned)(cc < c.w ? cc : c.w - 1);
    }
    return ln;
}

// A compiled window shape: NV blocks of 16 columns, HMAX rows, WX the exact width (0: the launch's width and height,
// rows past h skipped by uniform branches). An exact shape of two or more blocks keeps its last block, whose lanes are
// only partly window columns (5 of 16 at 21 x 21), in LDS: a row of TAIL entries (T, Tx, Ty, 0) and one zero entry
// that the lanes past the width read, so the registers hold only the full blocks.
template <int NV, int HMAX, int WX> struct Shape {
    static constexpr bool EXACT = WX > 0;
    static constexpr bool SPLIT = WX > 0 && NV >= 2;
    static constexpr int NR = SPLIT ? NV - 1 : NV;                 // blocks in registers
    static constexpr int TAIL = WX > 0 ? WX - 16 * (NV - 1) : 16;  // window columns of the last block
    static constexpr int LROW = TAIL + 1;                          // LDS entries per row
    static constexpr int LDS_POINT = SPLIT ? HMAX * LROW : 0;      // LDS entries (16 B) per point
    // block k holds only window columns in every lane (known at compile time for an exact width)
    static constexpr bool full(int k) { return WX > 0 && 16 * k + 15 < WX; }
    // STG (exact shapes): each level's J neighbourhood is staged in LDS, the window (w + 1 by h + 1 pixels) plus a
    // margin of M on every side, as floats in the level's own units; the iterations read it there
    static constexpr bool STG = WX > 0;
    static constexpr int M = 2;
    static constexpr int RW = WX + 1 + 2 * M, RH = HMAX + 1 + 2 * M;
    static constexpr int STAGE = STG ? (RW * RH + 3) / 4 * 4 : 0;    // floats per point, a multiple of 4
    static constexpr int LDS_BYTES = LDS_POINT * 16 + STAGE * 4;     // LDS per point
};

// The template of one level: window rows 0 .. HMAX - 1 (the first h used) of the lane's register blocks, and (SPLIT)
// the point's LDS rows of the last block with the entry this lane reads in each.
template <int NV, int HMAX, int WX> struct Tmpl {
    float t[HMAX][Shape<NV, HMAX, WX>::NR], tx[HMAX][Shape<NV, HMAX, WX>::NR], ty[HMAX][Shape<NV, HMAX, WX>::NR];
    float4 *last;
    int lc;
    float *R; // STG: the staged neighbourhood of J, RH rows of RW floats
};

// v unchanged, but known to the compiler only after `after` exists: a load addressed by the result cannot be issued
// before the value it follows, which keeps a software pipeline's loads a fixed number of rows ahead of their use (the
// loads are free of side effects, so without this the compiler issues all of them at once and runs out of registers).
template <typename V, typename A>
__device__ __forceinline__ V after(V v, A after)
{
    asm volatile("" : "+v"(v) : "v"(after));
    return v;
}

// LDS written by one lane and read by another lane of the same wave: order the accesses, no workgroup barrier.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The template at (px, py) (build_vx, lk_vx.c:36-97): the bilinear window S over rows -1 .. h and, per lane and block,
// columns c - 1, c, c + 1 of its window column c (four taps of each image row, x-interpolated, then y-interpolated
// with the row above), the Sobel gradients / 512 (or the Scharr variant's / 2048) and T = S's middle; returns G = the
// butterfly totals of Tx^2, Tx Ty, Ty^2 accumulated over rows, then blocks.
template <int NV, int HMAX, int WX, typename T>
__device__ __forceinline__ void build(const Level &lv, double px, double py, const Cfg &c, const Lanes<NV> &ln,
                                      Tmpl<NV, HMAX, WX> &tp, float *G)
{
    typedef Shape<NV, HMAX, WX> SH;
    constexpr bool EXACT = SH::EXACT;
    constexpr float ks = Ld<T>::scale, kg = ks * (1.0f / 512.0f);
    constexpr int K = NV == 1 ? 8 : 4; // image rows whose loads are in flight ahead of their use
    const GradW gw = grad_w(c.grad, kg);
    const double flx = floor(px), fly = floor(py);
    const float fx = (float)(px - flx), fy = (float)(py - fly);
    unsigned pitch;
    const __amdgpu_buffer_rsrc_t rs = level_rsrc<T>(lv, pitch);
    // image row 0 is the level's row fly - h/2 - 1; tap 0 of window column c is level column flx - w/2 - 1 + c
    const unsigned o = (unsigned)((int)fly - c.h / 2 - 1) * pitch + (unsigned)((int)flx - c.w / 2 - 1) * sizeof(T);
    unsigned ob[NV];
#pragma unroll
    for (int k = 0; k < NV; k++)
        ob[k] = o + ln.col[k] * (unsigned)sizeof(T);
    // image rows 0 .. h + 2; a constant trip count, so the loops unroll and the template stays in registers; rows past
    // h (EXACT: h == HMAX, none) are skipped by a uniform branch
    typename Row<T>::Q raw[HMAX + 3][NV];
#pragma unroll
    for (int i = 0; i < K; i++)
        if (EXACT || i <= c.h + 2)
#pragma unroll
            for (int k = 0; k < NV; k++)
                raw[i][k] = Row<T>::lq(rs, ob[k], (unsigned)i * pitch);
    float X[NV][3], S0[NV][3], S1[NV][3], d0[NV], d1[NV];
    float sxx = 0.0f, sxy = 0.0f, syy = 0.0f;
#pragma unroll
    for (int i = 0; i <= HMAX + 2; i++) {
        if (EXACT || i <= c.h + 2) {
            const int r = i - 1; // the S row this image row completes
#pragma unroll
            for (int k = 0; k < NV; k++) {
                float a[4], S[3];
                Row<T>::cq(raw[i][k], a);
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    const float here = fmaf(a[q + 1] - a[q], fx, a[q]);
                    S[q] = i == 0 ? 0.0f : fmaf(here - X[k][q], fy, X[k][q]);
                    X[k][q] = here;
                }
                if (i == 0)
                    continue;
                const float dS = S[2] - S[0];
                if (r == 0) {
#pragma unroll
                    for (int q = 0; q < 3; q++)
                        S0[k][q] = S[q];
                    d0[k] = dS;
                } else if (r == 1) {
#pragma unroll
                    for (int q = 0; q < 3; q++)
                        S1[k][q] = S[q];
                    d1[k] = dS;
                } else {
                    // window row rr = r - 2 from S rows rr (S0), rr + 1 (S1), rr + 2 (S)
                    const int rr = r - 2;
                    float tx = fmaf(d1[k], gw.wc, (d0[k] + dS) * gw.we) * gw.k;
                    const float vl = S[0] - S0[k][0], vc = S[1] - S0[k][1], vr = S[2] - S0[k][2];
                    float ty = fmaf(vc, gw.wc, (vl + vr) * gw.we) * gw.k;
                    const bool on = SH::full(k) || ln.on[k];
                    tx = on ? tx : 0.0f;
                    ty = on ? ty : 0.0f;
                    if (SH::SPLIT && k == NV - 1) {
                        tp.last[rr * SH::LROW + tp.lc] =
                            on ? make_float4(S1[k][1] * ks, tx, ty, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    } else {
                        tp.t[rr][k] = S1[k][1] * ks;
                        tp.tx[rr][k] = tx;
                        tp.ty[rr][k] = ty;
                    }
                    sxx = fmaf(tx, tx, sxx);
                    sxy = fmaf(tx, ty, sxy);
                    syy = fmaf(ty, ty, syy);
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        S0[k][q] = S1[k][q];
                        S1[k][q] = S[q];
                    }
                    d0[k] = d1[k];
                    d1[k] = dS;
                }
            }
            // the next row's loads: K rows ahead, issued once this row is done
            if (i + K <= HMAX + 2 && (EXACT || i + K <= c.h + 2))
#pragma unroll
                for (int k = 0; k < NV; k++) {
                    ob[k] = after(ob[k], X[k][0]);
                    raw[i + K][k] = Row<T>::lq(rs, ob[k], (unsigned)(i + K) * pitch);
                }
        }
    }
    if (SH::SPLIT)
        wave_sync(); // the last block's entries before any lane reads them
    G[0] = row_sum(sxx);
    G[1] = row_sum(sxy);
    G[2] = row_sum(syy);
}

// J sampled at (qx, qy) over the window (iterate_vx, lk_vx.c:100-140; residual_vx, lk_vx.c:142-175): two taps of each
// of rows 0 .. h of the window at floor(q), x-interpolated, the row above carried; It = T - J. RES = false: b = the
// butterfly totals of It Tx and It Ty; RES = true: of |It| / 64 and (It / 64)^2, lanes past the width masked.
template <int NV, int HMAX, int WX, typename T, bool RES>
__device__ __forceinline__ void sample(__amdgpu_buffer_rsrc_t rs, unsigned pitch, double qx, double qy, const Cfg &c,
                                       const Lanes<NV> &ln, const Tmpl<NV, HMAX, WX> &tp, float *b)
{
    typedef Shape<NV, HMAX, WX> SH;
    constexpr bool EXACT = SH::EXACT;
    constexpr float ks = Ld<T>::scale;
    constexpr int K = NV == 1 ? 12 : 8; // window rows whose loads are in flight ahead of their use
    const double flx = floor(qx), fly = floor(qy);
    const float fx = (float)(qx - flx), fy = (float)(qy - fly);
    const unsigned o = (unsigned)((int)fly - c.h / 2) * pitch + (unsigned)((int)flx - c.w / 2) * sizeof(T);
    unsigned ob[NV];
#pragma unroll
    for (int k = 0; k < NV; k++)
        ob[k] = o + ln.col[k] * (unsigned)sizeof(T);
    typename Row<T>::P raw[HMAX + 1][NV]; // window rows 0 .. h
#pragma unroll
    for (int r = 0; r < K; r++)
        if (EXACT || r <= c.h)
#pragma unroll
            for (int k = 0; k < NV; k++)
                raw[r][k] = Row<T>::lp(rs, ob[k], (unsigned)r * pitch);
    // SPLIT: the last block's template entries, KL rows ahead
    constexpr int KL = 3;
    float4 e[HMAX];
    int li = tp.lc; // an index, not a pointer, goes through after(): the reads stay LDS reads
    if (SH::SPLIT)
#pragma unroll
        for (int r = 0; r < KL && r < HMAX; r++)
            e[r] = tp.last[li + r * SH::LROW];
    float prev[NV];
    float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
    for (int r = 0; r <= HMAX; r++) {
        if (EXACT || r <= c.h) {
#pragma unroll
            for (int k = 0; k < NV; k++) {
                float a0, a1;
                Row<T>::cp(raw[r][k], a0, a1);
                const float cur = fmaf(a1 - a0, fx, a0);
                if (r > 0) {
                    const float jv = fmaf(cur - prev[k], fy, prev[k]);
                    const bool lds = SH::SPLIT && k == NV - 1;
                    const int kr = lds ? 0 : k;
                    const float tt = lds ? e[r - 1].x : tp.t[r - 1][kr], ttx = lds ? e[r - 1].y : tp.tx[r - 1][kr];
                    const float tty = lds ? e[r - 1].z : tp.ty[r - 1][kr];
                    const float it = fmaf(jv, -ks, tt); // T - J, the product by ks exact
                    if (RES) {
                        const float q = SH::full(k) || ln.on[k] ? it * (1.0f / 64.0f) : 0.0f;
                        s0 = s0 + fabsf(q);
                        s1 = fmaf(q, q, s1);
                    } else {
                        s0 = fmaf(it, ttx, s0);
                        s1 = fmaf(it, tty, s1);
                    }
                }
                prev[k] = cur;
            }
            // the next loads, once this row is done: image rows K ahead, template entries KL ahead
            if (r + K <= HMAX && (EXACT || r + K <= c.h))
#pragma unroll
                for (int k = 0; k < NV; k++) {
                    ob[k] = after(ob[k], prev[k]);
                    raw[r + K][k] = Row<T>::lp(rs, ob[k], (unsigned)(r + K) * pitch);
                }
            if (SH::SPLIT && r >= 1 && (r - 1) % KL == 0 && r - 1 + KL < HMAX) {
                li = after(li, s0);
#pragma unroll
                for (int q = r - 1 + KL; q < r - 1 + 2 * KL && q < HMAX; q++)
                    e[q] = tp.last[li + q * SH::LROW];
            }
        }
    }
    b[0] = row_sum(s0);
    b[1] = row_sum(s1);
}

// STG: the neighbourhood of J with corner (rx, ry) into the point's LDS region: RH rows of RW elements as floats in the
// level's own units. Lane j of the point's row loads columns j, j + 16, ...; the loads go out CH rows at a time.
// Elements outside the level are never read by a window that passed the bounds test (the margin only), so they are
// left as the buffer resource returns them.
template <int NV, int HMAX, int WX, typename T>
__device__ __forceinline__ void stage(__amdgpu_buffer_rsrc_t rs, unsigned pitch, int rx, int ry, int j, float *R)
{
    typedef Shape<NV, HMAX, WX> SH;
    constexpr int NQ = (SH::RW + 15) / 16, CH = 7;
    unsigned o = (unsigned)ry * pitch + (unsigned)(rx + j) * (unsigned)sizeof(T);
    wave_sync(); // every earlier read of the region is done
#pragma unroll
    for (int r0 = 0; r0 < SH::RH; r0 += CH) {
        float v[CH][NQ];
#pragma unroll
        for (int r = r0; r < r0 + CH && r < SH::RH; r++)
#pragma unroll
            for (int q = 0; q < NQ; q++)
                if (q < NQ - 1 || j + 16 * q < SH::RW)
                    v[r - r0][q] = Row<T>::l1(rs, o + 16u * q * (unsigned)sizeof(T), (unsigned)r * pitch);
#pragma unroll
        for (int r = r0; r < r0 + CH && r < SH::RH; r++)
#pragma unroll
            for (int q = 0; q < NQ; q++)
                if (q < NQ - 1 || j + 16 * q < SH::RW)
                    R[r * SH::RW + j + 16 * q] = v[r - r0][q];
        o = after(o, v[0][0]); // the next chunk's loads after this chunk's
    }
    wave_sync(); // the region before any lane reads it
}

// iterate_vx (lk_vx.c:100-140) on the staged neighbourhood: the window's corner is (bx, by) in the region.
template <int NV, int HMAX, int WX>
__device__ __forceinline__ void sample_st(const float *R, int bx, int by, float fx, float fy, float ks,
                                          const Lanes<NV> &ln, const Tmpl<NV, HMAX, WX> &tp, float *b)
{
    typedef Shape<NV, HMAX, WX> SH;
    constexpr int K = 4; // rows read ahead of their use
    int base = by * SH::RW + bx;
    float raw0[HMAX + 1][NV], raw1[HMAX + 1][NV];
#pragma unroll
    for (int r = 0; r < K; r++)
#pragma unroll
        for (int k = 0; k < NV; k++) {
            raw0[r][k] = R[base + r * SH::RW + (int)ln.col[k]];
            raw1[r][k] = R[base + r * SH::RW + (int)ln.col[k] + 1];
        }
    constexpr int KL = 3;
    float4 e[HMAX];
    int li = tp.lc;
    if (SH::SPLIT)
#pragma unroll
        for (int r = 0; r < KL && r < HMAX; r++)
            e[r] = tp.last[li + r * SH::LROW];
    float prev[NV];
    float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
    for (int r = 0; r <= HMAX; r++) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            const float a0 = raw0[r][k], a1 = raw1[r][k];
            const float cur = fmaf(a1 - a0, fx, a0);
            if (r > 0) {
                const float jv = fmaf(cur - prev[k], fy, prev[k]);
                const bool lds = SH::SPLIT && k == NV - 1;
                const int kr = lds ? 0 : k;
                const float tt = lds ? e[r - 1].x : tp.t[r - 1][kr], ttx = lds ? e[r - 1].y : tp.tx[r - 1][kr];
                const float tty = lds ? e[r - 1].z : tp.ty[r - 1][kr];
                const float it = fmaf(jv, -ks, tt); // T - J, the product by ks exact
                s0 = fmaf(it, ttx, s0);
                s1 = fmaf(it, tty, s1);
            }
            prev[k] = cur;
        }
        if (r + K <= HMAX) {
            base = after(base, prev[0]);
#pragma unroll
            for (int k = 0; k < NV; k++) {
                raw0[r + K][k] = R[base + (r + K) * SH::RW + (int)ln.col[k]];
                raw1[r + K][k] = R[base + (r + K) * SH::RW + (int)ln.col[k] + 1];
            }
        }
        if (SH::SPLIT && r >= 1 && (r - 1) % KL == 0 && r - 1 + KL < HMAX) {
            li = after(li, s0);
#pragma unroll
            for (int q = r - 1 + KL; q < r - 1 + 2 * KL && q < HMAX; q++)
                e[q] = tp.last[li + q * SH::LROW];
        }
    }
    b[0] = row_sum(s0);
    b[1] = row_sum(s1);
}

// (float)((long double)s / n) for s >= 0: the quotient rounded to a 64-bit significand, then to float
// (residual_vx, lk_vx.c:171-174). The double quotient rounds to the same float unless it lies within one of its own
// ulps of a float rounding boundary; then the exact 128-bit path decides.
__device__ __forceinline__ float mean_ld(float s, int n)
{
    if (!(s > 0.0f))
        return 0.0f;
    const double d = (double)s / (double)n;
    const uint64_t bits = (uint64_t)__double_as_longlong(d);
    const int64_t low = (int64_t)(bits & ((1ull << 29) - 1)) - (1ll << 28); // distance to the float midpoint pattern
    if (d > 0x1p-120 && (low > 2 || low < -2))
        return (float)d;
    return xl_div_to_float(xl_add(XL{0, 0}, s), n);
}

struct Track {
    float x, y, min_eig, residual, cxx, cxy, cyy;
    int iters, status;
};

// The 2x2 system of one level: eigenvalue test and inverse. mixed: fp32 unless G's condition number passes 1e4.
struct Solve {
    double ia, ib, ic, det, eig;
};

__device__ __forceinline__ Solve solve2(double gxx, double gxy, double gyy, double area, bool mixed)
{
    Solve s;
    if (mixed) {
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
    s.ia = gyy / s.det;
    s.ib = -gxy / s.det;
    s.ic = gxx / s.det;
    return s;
}

template <int NV, int HMAX, int WX, typename T0, typename T1>
__device__ __forceinline__ void build_at(int L, const Level &lv, double px, double py, const Cfg &c,
                                         const Lanes<NV> &ln, Tmpl<NV, HMAX, WX> &tp, float *G)
{
    if (std::is_same<T0, T1>::value || L > 0)
        build<NV, HMAX, WX, T1>(lv, px, py, c, ln, tp, G);
    else
        build<NV, HMAX, WX, T0>(lv, px, py, c, ln, tp, G);
}

// The level J is read at: its buffer resource and pitch in bytes for the element type of level L.
template <typename T0, typename T1>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_at(int L, const Level &lv, unsigned &pitch)
{
    if (std::is_same<T0, T1>::value || L > 0)
        return level_rsrc<T1>(lv, pitch);
    return level_rsrc<T0>(lv, pitch);
}

template <int NV, int HMAX, int WX, typename T0, typename T1>
__device__ __forceinline__ void stage_at(int L, __amdgpu_buffer_rsrc_t rs, unsigned pitch, int rx, int ry, int j,
                                         float *R)
{
    if (std::is_same<T0, T1>::value || L > 0)
        stage<NV, HMAX, WX, T1>(rs, pitch, rx, ry, j, R);
    else
        stage<NV, HMAX, WX, T0>(rs, pitch, rx, ry, j, R);
}

template <int NV, int HMAX, int WX, typename T0, typename T1>
__device__ __forceinline__ void sample_at(int L, __amdgpu_buffer_rsrc_t rs, unsigned pitch, double qx, double qy,
                                          const Cfg &c, const Lanes<NV> &ln, const Tmpl<NV, HMAX, WX> &tp, float *b)
{
    if (std::is_same<T0, T1>::value || L > 0)
        sample<NV, HMAX, WX, T1, false>(rs, pitch, qx, qy, c, ln, tp, b);
    else
        sample<NV, HMAX, WX, T0, false>(rs, pitch, qx, qy, c, ln, tp, b);
}

// One point from pyramid I to J: the reference's level walk (lk_driver.inc:7-109). Every value is the same in the 16
// lanes of the point's row.
template <int NV, int HMAX, int WX, typename T0, typename T1>
__device__ void track(const PyrView &I, const PyrView &J, int top, const Cfg &c, bool mixed, const Lanes<NV> &ln,
                      char *lds, float x, float y, float gx0, float gy0, Track &o)
{
    Tmpl<NV, HMAX, WX> tp;
    tp.last = (float4 *)lds;
    tp.R = (float *)(lds + Shape<NV, HMAX, WX>::LDS_POINT * 16);
    tp.lc = ln.j < Shape<NV, HMAX, WX>::TAIL ? ln.j : Shape<NV, HMAX, WX>::TAIL;
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
        build_at<NV, HMAX, WX, T0, T1>(L, li, px, py, c, ln, tp, G);
        const double gxx = G[0], gxy = G[1], gyy = G[2];
        const Solve s = solve2(gxx, gxy, gyy, area, mixed);
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
        const double sx = px + gx, sy = py + gy; // the reference adds px + gx first, then the level's step
        unsigned pj;
        const __amdgpu_buffer_rsrc_t rsj = rsrc_at<T0, T1>(L, lj, pj); // once per level, not per iteration
        typedef Shape<NV, HMAX, WX> SH;
        int rx = 0, ry = 0; // STG: the staged region's corner
        if (SH::STG) {
            // the corner stays inside the level: a negative lane offset would wrap, and rows reached through the
            // scalar row step past it would be range-checked as out of the level
            rx = max((int)floor(sx) - c.w / 2 - SH::M, 0);
            ry = max((int)floor(sy) - c.h / 2 - SH::M, 0);
            stage_at<NV, HMAX, WX, T0, T1>(L, rsj, pj, rx, ry, ln.j, tp.R);
        }
        double vx = 0, vy = 0;
        for (int it = 0; it < c.max_iter; it++) {
            const double qx = sx + vx, qy = sy + vy;
            if (qx < mx || qy < my || qx >= lx || qy >= ly) {
                if (L == 0)
                    status |= LK_OUT, valid0 = false;
                break;
            }
            float b[2];
            if (SH::STG) {
                const double flx = floor(qx), fly = floor(qy);
                const int bx = (int)flx - c.w / 2 - rx, by = (int)fly - c.h / 2 - ry;
                if (bx < 0 || by < 0 || bx + c.w + 1 > SH::RW || by + c.h + 1 > SH::RH) {
                    // the window left the region: stage again around it
                    rx = max(rx + bx - SH::M, 0);
                    ry = max(ry + by - SH::M, 0);
                    stage_at<NV, HMAX, WX, T0, T1>(L, rsj, pj, rx, ry, ln.j, tp.R);
                }
                const float ks = (std::is_same<T0, T1>::value || L > 0) ? Ld<T1>::scale : Ld<T0>::scale;
                sample_st<NV, HMAX, WX>(tp.R, (int)flx - c.w / 2 - rx, (int)fly - c.h / 2 - ry, (float)(qx - flx),
                                        (float)(qy - fly), ks, ln, tp, b);
            } else {
                sample_at<NV, HMAX, WX, T0, T1>(L, rsj, pj, qx, qy, c, ln, tp, b);
            }
            const double bx = (double)b[0] / 64.0, by = (double)b[1] / 64.0;
            const double ex = s.ia * bx + s.ib * by, ey = s.ib * bx + s.ic * by;
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
            unsigned p0;
            const __amdgpu_buffer_rsrc_t rs0 = level_rsrc<T0>(l0, p0);
            sample<NV, HMAX, WX, T0, true>(rs0, p0, qx, qy, c, ln, tp, r);
            const int n = c.w * c.h;
            o.residual = mean_ld(r[0], n);
            if (o.residual > c.max_res)
                status |= LK_RESIDUAL;
            cov_ld(mean_ld(r[1], n), n, G0[0], G0[2], G0[1], &o.cxx, &o.cxy, &o.cyy);
        }
    }
    o.status = status;
}

// A point and its flowback (lk_driver.inc:111-121): the second pass tracks the estimate back from J to I, the template
// rebuilt from J. One call site of track for both passes.
template <int NV, int HMAX, int WX, typename T0, typename T1>
__device__ __forceinline__ void track_fb(const PyrView &I, const PyrView &J, int top, int fb_top, const Cfg &c,
                                         bool mixed, float fb_max, const Lanes<NV> &ln, char *lds, float x,
                                         float y, float gx, float gy, Track &o)
{
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1 && !(fb_max > 0.0f && o.status == 0))
            break;
        Track t;
        const float sx = pass ? o.x : x, sy = pass ? o.y : y;
        const float ux = pass ? (float)((double)x - (double)o.x) : gx;
        const float uy = pass ? (float)((double)y - (double)o.y) : gy;
        track<NV, HMAX, WX, T0, T1>(pass ? J : I, pass ? I : J, pass ? fb_top : top, c, mixed, ln, lds, sx, sy, ux, uy,
                                    t);
        if (pass == 0) {
            o = t;
        } else {
            const double ex = (double)t.x - (double)x, ey = (double)t.y - (double)y;
            if (t.status || ex * ex + ey * ey > (double)fb_max * (double)fb_max)
                o.status |= LK_FLOWBACK;
        }
    }
}

__device__ __forceinline__ void store(lk_out *out, const Track &o)
{
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
    *out = r;
}

// The last pair whose numbering starts at or before slot s (uniform when s is).
__device__ __forceinline__ int pair_of(const int *pre, int npairs, int s)
{
    int lo = 0, hi = npairs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s >= pre[mid])
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <int NV, int HMAX, int WX, typename T0, typename T1>
__global__ __launch_bounds__(512) void lk_kernel(LkArgs a)
{
    extern __shared__ float4 lk_lds[];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    // workgroup b runs on XCD b mod 8; XCD x takes the contiguous run of workgroups [x per_xcd, (x + 1) per_xcd)
    const int wg = (int)(blockIdx.x % NXCD) * a.per_xcd + (int)(blockIdx.x / NXCD);
    const int s0 = (wg * a.waves + wave) * PPW;
    if (s0 >= a.slots)
        return;
    const int q = __builtin_amdgcn_readfirstlane(pair_of(a.pre, a.npairs, s0));
    const int local = s0 - a.pre[q] + (lane >> 4);
    const bool live = local < a.cnt[q];
    const int pt = live ? point_at(a, q, local) : 0;
    // an empty slot starts far outside every level: it fits none, reads nothing and writes nothing
    float x = -1e30f, y = -1e30f, gx = 0.0f, gy = 0.0f;
    if (live) {
        x = a.xy[2 * pt];
        y = a.xy[2 * pt + 1];
        if (a.guess) {
            gx = a.guess[2 * pt];
            gy = a.guess[2 * pt + 1];
        }
    }
    const Lanes<NV> ln = lanes<NV>(a.cfg, lane & 15);
    const int top = a.top[q];
    char *lds = (char *)lk_lds + (wave * PPW + (lane >> 4)) * Shape<NV, HMAX, WX>::LDS_BYTES;
    Track o;
    track_fb<NV, HMAX, WX, T0, T1>(*a.I[q], *a.J[q], top, a.fb_top >= 0 ? a.fb_top : top, a.cfg,
                                   a.solve == VTG_SOLVE_MIXED, a.fb_max, ln, lds, x, y, gx, gy, o);
    if (live && (lane & 15) == 0)
        store(a.out + pt, o);
}

// ---------------------------------------------------------------- three points per wavefront (windows 17 .. 21 wide)
//
// A window wider than 16 columns and at most 21 (the 1080p workload's 21 x 21) leaves 16 - (w - 16) lanes of a
// point's second block idle in the layout above (11 of 16 at 21 x 21). Here a wavefront tracks 3 points and every lane
// works: lanes 16p .. 16p + 15 (p = 0, 1, 2) are point p's reference vector lanes, lane j owning window column j and
// keeping that column's template in registers; lanes 48 .. 48 + 3 TAIL - 1 take the second block's TAIL = w - 16
// columns of the three points (lane 48 + TAIL p + t: point p, column 16 + t; the zero lane follows them, and any lanes
// after that repeat point 2's last column and feed no one). A second-block column's gradients (at each build) and its
// T - J (at each sample) reach vector lane t of its point by ds_bpermute; that lane keeps the gradients in registers
// and adds the products to its sums right after its own column of the same row, where the reference's lane t adds its
// second block (lk_vx.c:82-94, 119-126, 160-168). So every sum, and every record, is the same bits as the reference's
// and as the layout above. A point's totals reach its second-block lanes the same way, so all lanes of a point run the
// same level walk on the same values.
// Per point and iteration both layouts issue about the same VALU work (the second block's idle lanes there, the
// hand-off and the extra chain products here); this one builds templates with about a fifth fewer instructions.
template <int HMAX, int WX> struct Shape3 {
    static_assert(WX > 16 && WX <= 21, "three points per wavefront: windows 17 .. 21 wide");
    static constexpr int TAIL = WX - 16;
    static constexpr int M = 2; // the staged neighbourhood's margin, as in Shape
    static constexpr int RW = WX + 1 + 2 * M, RH = HMAX + 1 + 2 * M;
    static constexpr int STAGE = (RW * RH + 3) / 4 * 4; // floats per point
    static constexpr int WAVE_BYTES = 3 * STAGE * 4;    // LDS per wavefront
};
constexpr int PPW3 = 3;

struct Lane3 {
    int p;     // the wavefront's point this lane works for
    int c;     // its window column
    int j;     // lanes 0 .. 47: the reference vector lane; lanes 48 .. 63: the column's place in the second block
    bool vec;  // a vector lane (lanes 0 .. 47)
    bool tail; // a vector lane whose reference lane also owns second-block column 16 + j
    bool zero; // the zero lane (lane 48 + 3 TAIL): its level reads fall outside the level and return 0
    int src;   // ds_bpermute address of the lane computing this point's column 16 + j; for a vector lane without a
               // second-block column, the zero lane, whose gradients and (read from the level) T - J are 0
    int head;  // ds_bpermute address of a vector lane of this point (the lanes holding its totals)
};

template <int WX>
__device__ __forceinline__ Lane3 lane3(int lane)
{
    constexpr int TAIL = WX - 16;
    static_assert(3 * TAIL < 16, "a spare lane for the zero lane");
    const int row = lane >> 4, j = lane & 15;
    Lane3 l;
    l.vec = row < 3;
    l.zero = lane == 48 + 3 * TAIL;
    if (l.vec) {
        l.p = row;
        l.j = j;
        l.c = j;
        l.tail = j < TAIL;
        l.src = 4 * (j < TAIL ? 48 + TAIL * row + j : 48 + 3 * TAIL);
    } else {
        l.p = min(j / TAIL, 2);
        l.j = min(j - TAIL * l.p, TAIL - 1);
        l.c = 16 + l.j;
        l.tail = false;
        l.src = 4 * lane;
    }
    l.head = 4 * 16 * l.p;
    return l;
}

// LDS reads by byte address: an address carried in a register (the loop's ordering goes through it) plus a constant
// offset, which the reads take as their immediate offset.
typedef __attribute__((address_space(3))) float lds_float;
__device__ __forceinline__ unsigned lds_addr(const void *p) { return (unsigned)(uintptr_t)(const lds_float *)p; }
__device__ __forceinline__ float lds_f(unsigned a) { return *(const lds_float *)(uintptr_t)a; }

// v of the lane at byte address a (ds_bpermute: every lane read here is active when its reader is, as both work for
// the same point).
__device__ __forceinline__ float bperm(int a, float v)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(a, __float_as_int(v)));
}

// A level's template in one lane: its own column's T, Tx, Ty and (vector lanes) the gradients of its second-block column,
// zero for a lane without one (zero gradients add nothing to a sum: the reference's zeroed lanes, lk_vx.c:82-85).
template <int HMAX> struct Tmpl3 {
    float t[HMAX], tx[HMAX], ty[HMAX], ux[HMAX], uy[HMAX];
};

// build_vx (lk_vx.c:36-97) in this layout: the same per-column arithmetic as build, one column per lane.
template <int HMAX, int WX, typename T>
__device__ __forceinline__ void build3(const Level &lv, double px, double py, int grad, const Lane3 &ln,
                                       Tmpl3<HMAX> &tp, float *G)
{
    constexpr float ks = Ld<T>::scale, kg = ks * (1.0f / 512.0f);
    constexpr int K = 4; // image rows whose loads are in flight ahead of their use
    const GradW gw = grad_w(grad, kg);
    const double flx = floor(px), fly = floor(py);
    const float fx = (float)(px - flx), fy = (float)(py - fly);
    unsigned pitch;
    const __amdgpu_buffer_rsrc_t rs = level_rsrc<T>(lv, pitch);
    // image row 0 is the level's row fly - h/2 - 1; tap 0 of window column c is level column flx - w/2 - 1 + c
    unsigned o = (unsigned)((int)fly - HMAX / 2 - 1) * pitch + (unsigned)((int)flx - WX / 2 - 1 + ln.c) * sizeof(T);
    o = ln.zero ? 0x80000000u : o; // past every level: the zero lane reads 0
    typename Row<T>::Q raw[HMAX + 3];
#pragma unroll
    for (int i = 0; i < K; i++)
        raw[i] = Row<T>::lq(rs, o, (unsigned)i * pitch);
    float X[3], S0[3], S1[3], d0 = 0.0f, d1 = 0.0f;
    float sxx = 0.0f, sxy = 0.0f, syy = 0.0f;
#pragma unroll
    for (int i = 0; i <= HMAX + 2; i++) {
        const int r = i - 1; // the S row this image row completes
        float a[4], S[3];
        Row<T>::cq(raw[i], a);
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const float here = fmaf(a[q + 1] - a[q], fx, a[q]);
            S[q] = i == 0 ? 0.0f : fmaf(here - X[q], fy, X[q]);
            X[q] = here;
        }
        if (i > 0) {
            const float dS = S[2] - S[0];
            if (r == 0) {
#pragma unroll
                for (int q = 0; q < 3; q++)
                    S0[q] = S[q];
                d0 = dS;
            } else if (r == 1) {
#pragma unroll
                for (int q = 0; q < 3; q++)
                    S1[q] = S[q];
                d1 = dS;
            } else {
                // window row rr = r - 2 from S rows rr (S0), rr + 1 (S1), rr + 2 (S)
                const int rr = r - 2;
                const float tx = fmaf(d1, gw.wc, (d0 + dS) * gw.we) * gw.k;
                const float vl = S[0] - S0[0], vc = S[1] - S0[1], vr = S[2] - S0[2];
                const float ty = fmaf(vc, gw.wc, (vl + vr) * gw.we) * gw.k;
                tp.t[rr] = S1[1] * ks;
                tp.tx[rr] = tx;
                tp.ty[rr] = ty;
                const float ux = bperm(ln.src, tx), uy = bperm(ln.src, ty); // 0 from the zero lane
                tp.ux[rr] = ux;
                tp.uy[rr] = uy;
                sxx = fmaf(tx, tx, sxx);
                sxy = fmaf(tx, ty, sxy);
                syy = fmaf(ty, ty, syy);
                sxx = fmaf(ux, ux, sxx);
                sxy = fmaf(ux, uy, sxy);
                syy = fmaf(uy, uy, syy);
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    S0[q] = S1[q];
                    S1[q] = S[q];
                }
                d0 = d1;
                d1 = dS;
            }
        }
        // the next row's loads: K rows ahead, issued once this row is done
        if (i + K <= HMAX + 2) {
            o = after(o, X[0]);
            raw[i + K] = Row<T>::lq(rs, o, (unsigned)(i + K) * pitch);
        }
    }
    G[0] = bperm(ln.head, row_sum(sxx));
    G[1] = bperm(ln.head, row_sum(sxy));
    G[2] = bperm(ln.head, row_sum(syy));
}

// iterate_vx (lk_vx.c:100-140; RES = false) and residual_vx (lk_vx.c:142-175; RES = true) in this layout, J read from
// the level: two taps of each of window rows 0 .. h at floor(q), x-interpolated, the row above carried; It = T - J.
template <int HMAX, int WX, typename T, bool RES>
__device__ __forceinline__ void sample3(__amdgpu_buffer_rsrc_t rs, unsigned pitch, double qx, double qy,
                                        const Lane3 &ln, const Tmpl3<HMAX> &tp, float *b)
{
    constexpr float ks = Ld<T>::scale;
    constexpr int K = 8; // window rows whose loads are in flight ahead of their use
    const double flx = floor(qx), fly = floor(qy);
    const float fx = (float)(qx - flx), fy = (float)(qy - fly);
    unsigned o = (unsigned)((int)fly - HMAX / 2) * pitch + (unsigned)((int)flx - WX / 2 + ln.c) * sizeof(T);
    o = ln.zero ? 0x80000000u : o; // past every level: the zero lane's T - J is 0
    typename Row<T>::P raw[HMAX + 1];
#pragma unroll
    for (int r = 0; r < K; r++)
        raw[r] = Row<T>::lp(rs, o, (unsigned)r * pitch);
    float prev = 0.0f, s0 = 0.0f, s1 = 0.0f;
#pragma unroll
    for (int r = 0; r <= HMAX; r++) {
        float a0, a1;
        Row<T>::cp(raw[r], a0, a1);
        const float cur = fmaf(a1 - a0, fx, a0);
        if (r > 0) {
            const float jv = fmaf(cur - prev, fy, prev);
            const float it = fmaf(jv, -ks, tp.t[r - 1]); // T - J, the product by ks exact
            const float iu = bperm(ln.src, it);
            if (RES) {
                const float q = it * (1.0f / 64.0f), qu = iu * (1.0f / 64.0f); // 0 from the zero lane
                s0 = s0 + fabsf(q);
                s1 = fmaf(q, q, s1);
                s0 = s0 + fabsf(qu);
                s1 = fmaf(qu, qu, s1);
            } else {
                s0 = fmaf(it, tp.tx[r - 1], s0);
                s1 = fmaf(it, tp.ty[r - 1], s1);
                s0 = fmaf(iu, tp.ux[r - 1], s0);
                s1 = fmaf(iu, tp.uy[r - 1], s1);
            }
        }
        prev = cur;
        if (r + K <= HMAX) {
            o = after(o, prev);
            raw[r + K] = Row<T>::lp(rs, o, (unsigned)(r + K) * pitch);
        }
    }
    b[0] = bperm(ln.head, row_sum(s0));
    b[1] = bperm(ln.head, row_sum(s1));
}

// The neighbourhood of J with corner (rx, ry) into the point's LDS region (as stage; the vector lanes load it).
template <int HMAX, int WX, typename T>
__device__ __forceinline__ void stage3(__amdgpu_buffer_rsrc_t rs, unsigned pitch, int rx, int ry, const Lane3 &ln,
                                       float *R)
{
    typedef Shape3<HMAX, WX> SH;
    constexpr int NQ = (SH::RW + 15) / 16, CH = 7;
    wave_sync(); // every earlier read of the region is done
    if (ln.vec) {
        const int j = ln.j;
        unsigned o = (unsigned)ry * pitch + (unsigned)(rx + j) * (unsigned)sizeof(T);
#pragma unroll
        for (int r0 = 0; r0 < SH::RH; r0 += CH) {
            float v[CH][NQ];
#pragma unroll
            for (int r = r0; r < r0 + CH && r < SH::RH; r++)
#pragma unroll
                for (int q = 0; q < NQ; q++)
                    if (q < NQ - 1 || j + 16 * q < SH::RW)
                        v[r - r0][q] = Row<T>::l1(rs, o + 16u * q * (unsigned)sizeof(T), (unsigned)r * pitch);
#pragma unroll
            for (int r = r0; r < r0 + CH && r < SH::RH; r++)
#pragma unroll
                for (int q = 0; q < NQ; q++)
                    if (q < NQ - 1 || j + 16 * q < SH::RW)
                        R[r * SH::RW + j + 16 * q] = v[r - r0][q];
            o = after(o, v[0][0]); // the next chunk's loads after this chunk's
        }
    }
    wave_sync(); // the region before any lane reads it
}

// iterate_vx (lk_vx.c:100-140) on the staged neighbourhood: the window's corner is (bx, by) in the region.
template <int HMAX, int WX>
__device__ __forceinline__ void sample_st3(const float *R, int bx, int by, float fx, float fy, float ks,
                                           const Lane3 &ln, const Tmpl3<HMAX> &tp, float *b)
{
    typedef Shape3<HMAX, WX> SH;
    constexpr int K = 4; // rows read ahead of their use
    // the region's rows from three byte addresses, rows 0-9, 10-19 and 20-, each within the reads' immediate offsets
    constexpr int RB = 10;
    static_assert(HMAX + 1 <= 3 * RB, "");
    unsigned A[3];
    A[0] = lds_addr(R) + 4u * (unsigned)(by * SH::RW + bx + ln.c);
    A[1] = A[0] + RB * SH::RW * 4;
    A[2] = A[1] + RB * SH::RW * 4;
    float raw0[HMAX + 1], raw1[HMAX + 1];
#pragma unroll
    for (int r = 0; r < K; r++) {
        raw0[r] = lds_f(A[r / RB] + (r % RB) * SH::RW * 4);
        raw1[r] = lds_f(A[r / RB] + (r % RB) * SH::RW * 4 + 4);
    }

    // the sums run D window rows behind the values they add, so a row's T - J has come back from the second-block
    // lane (ds_bpermute) by the time its sums need it
    constexpr int D = 3;
    float it[HMAX], iu[HMAX];
    float prev = 0.0f, s0 = 0.0f, s1 = 0.0f;
#pragma unroll
    for (int r = 0; r <= HMAX + D; r++) {
        if (r <= HMAX) {
            const float cur = fmaf(raw1[r] - raw0[r], fx, raw0[r]);
            if (r > 0) {
                const float jv = fmaf(cur - prev, fy, prev);
                it[r - 1] = fmaf(jv, -ks, tp.t[r - 1]); // T - J, the product by ks exact
                iu[r - 1] = bperm(ln.src, it[r - 1]);
            }
            prev = cur;
            if (r + K <= HMAX) {
                const int n = r + K;
                A[n / RB] = after(A[n / RB], prev);
                raw0[n] = lds_f(A[n / RB] + (n % RB) * SH::RW * 4);
                raw1[n] = lds_f(A[n / RB] + (n % RB) * SH::RW * 4 + 4);
            }
        }
        const int w = r - 1 - D; // the window row whose sums this step adds
        if (w >= 0 && w < HMAX) {
            s0 = fmaf(it[w], tp.tx[w], s0);
            s1 = fmaf(it[w], tp.ty[w], s1);
            s0 = fmaf(iu[w], tp.ux[w], s0);
            s1 = fmaf(iu[w], tp.uy[w], s1);
        }
    }
    b[0] = bperm(ln.head, row_sum(s0));
    b[1] = bperm(ln.head, row_sum(s1));
}

// One point from pyramid I to J in this layout: the level walk of track (lk_driver.inc:7-109), every value the same in
// all lanes of the point.
template <int HMAX, int WX, typename T0, typename T1>
__device__ __forceinline__ void track3(const PyrView &I, const PyrView &J, int top, const Cfg c, bool mixed,
                                       const Lane3 &ln, float *R, float x, float y, float gx0, float gy0,
                                       Track &o)
{
    typedef Shape3<HMAX, WX> SH;
    Tmpl3<HMAX> tp;
    double gx = ldexp((double)gx0, -top), gy = ldexp((double)gy0, -top);
    const double area = (double)(WX * HMAX), eps2 = (double)c.eps * (double)c.eps;
    constexpr int mx = WX / 2 + 1, my = HMAX / 2 + 1;
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
        if (std::is_same<T0, T1>::value || L > 0)
            build3<HMAX, WX, T1>(li, px, py, c.grad, ln, tp, G);
        else
            build3<HMAX, WX, T0>(li, px, py, c.grad, ln, tp, G);
        const double gxx = G[0], gxy = G[1], gyy = G[2];
        const Solve s = solve2(gxx, gxy, gyy, area, mixed);
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
        const double sx = px + gx, sy = py + gy; // the reference adds px + gx first, then the level's step
        unsigned pj;
        const __amdgpu_buffer_rsrc_t rsj = rsrc_at<T0, T1>(L, lj, pj); // once per level, not per iteration
        const float ks = (std::is_same<T0, T1>::value || L > 0) ? Ld<T1>::scale : Ld<T0>::scale;
        // the staged region's corner stays inside the level: a negative lane offset would wrap, and rows reached
        // through the scalar row step past it would be range-checked as out of the level
        int rx = max((int)floor(sx) - WX / 2 - SH::M, 0);
        int ry = max((int)floor(sy) - HMAX / 2 - SH::M, 0);
        if (std::is_same<T0, T1>::value || L > 0)
            stage3<HMAX, WX, T1>(rsj, pj, rx, ry, ln, R);
        else
            stage3<HMAX, WX, T0>(rsj, pj, rx, ry, ln, R);
        double vx = 0, vy = 0;
        for (int it = 0; it < c.max_iter; it++) {
            const double qx = sx + vx, qy = sy + vy;
            if (qx < mx || qy < my || qx >= lx || qy >= ly) {
                if (L == 0)
                    status |= LK_OUT, valid0 = false;
                break;
            }
            const double flx = floor(qx), fly = floor(qy);
            int bx = (int)flx - WX / 2 - rx, by = (int)fly - HMAX / 2 - ry;
            if (bx < 0 || by < 0 || bx + WX + 1 > SH::RW || by + HMAX + 1 > SH::RH) {
                // the window left the region: stage again around it
                rx = max(rx + bx - SH::M, 0);
                ry = max(ry + by - SH::M, 0);
                if (std::is_same<T0, T1>::value || L > 0)
                    stage3<HMAX, WX, T1>(rsj, pj, rx, ry, ln, R);
                else
                    stage3<HMAX, WX, T0>(rsj, pj, rx, ry, ln, R);
                bx = (int)flx - WX / 2 - rx;
                by = (int)fly - HMAX / 2 - ry;
            }
            float b[2];
            sample_st3<HMAX, WX>(R, bx, by, (float)(qx - flx), (float)(qy - fly), ks, ln, tp, b);
            const double bxx = (double)b[0] / 64.0, byy = (double)b[1] / 64.0;
            const double ex = s.ia * bxx + s.ib * byy, ey = s.ib * bxx + s.ic * byy;
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
            unsigned p0;
            const __amdgpu_buffer_rsrc_t rs0 = level_rsrc<T0>(l0, p0);
            sample3<HMAX, WX, T0, true>(rs0, p0, qx, qy, ln, tp, r);
            constexpr int n = WX * HMAX;
            o.residual = mean_ld(r[0], n);
            if (o.residual > c.max_res)
                status |= LK_RESIDUAL;
            cov_ld(mean_ld(r[1], n), n, G0[0], G0[2], G0[1], &o.cxx, &o.cxy, &o.cyy);
        }
    }
    o.status = status;
}

template <int HMAX, int WX, typename T0, typename T1>
__global__ __launch_bounds__(512) void lk3_kernel(LkArgs a)
{
    extern __shared__ float4 lk_lds[];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    // workgroup b runs on XCD b mod 8; XCD x takes the contiguous run of workgroups [x per_xcd, (x + 1) per_xcd)
    const int wg = (int)(blockIdx.x % NXCD) * a.per_xcd + (int)(blockIdx.x / NXCD);
    const int s0 = (wg * a.waves + wave) * PPW3;
    if (s0 >= a.slots)
        return;
    const int q = __builtin_amdgcn_readfirstlane(pair_of(a.pre, a.npairs, s0));
    const Lane3 ln = lane3<WX>(lane);
    const int local = s0 - a.pre[q] + ln.p;
    const bool live = local < a.cnt[q];
    const int pt = live ? point_at(a, q, local) : 0;
    // an empty slot starts far outside every level: it fits none, reads nothing and writes nothing
    float x = -1e30f, y = -1e30f, gx = 0.0f, gy = 0.0f;
    if (live) {
        x = a.xy[2 * pt];
        y = a.xy[2 * pt + 1];
        if (a.guess) {
            gx = a.guess[2 * pt];
            gy = a.guess[2 * pt + 1];
        }
    }
    const int top = a.top[q], fb_top = a.fb_top >= 0 ? a.fb_top : top;
    typedef Shape3<HMAX, WX> SH;
    float *R = (float *)((char *)lk_lds + wave * SH::WAVE_BYTES) + ln.p * SH::STAGE;
    const bool mixed = a.solve == VTG_SOLVE_MIXED;
    const Cfg cfg = a.cfg;
    const float fb_max = a.fb_max;
    const PyrView *pI = a.I[q], *pJ = a.J[q];
    // the point, then its flowback (lk_driver.inc:111-121): one call site of track3 for both passes
    Track o;
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1 && !(fb_max > 0.0f && o.status == 0))
            break;
        Track t;
        const float sx = pass ? o.x : x, sy = pass ? o.y : y;
        const float ux = pass ? (float)((double)x - (double)o.x) : gx;
        const float uy = pass ? (float)((double)y - (double)o.y) : gy;
        track3<HMAX, WX, T0, T1>(pass ? *pJ : *pI, pass ? *pI : *pJ, pass ? fb_top : top, cfg, mixed, ln, R, sx, sy,
                                 ux, uy, t);
        if (pass == 0) {
            o = t;
        } else {
            const double ex = (double)t.x - (double)x, ey = (double)t.y - (double)y;
            if (t.status || ex * ex + ey * ey > (double)fb_max * (double)fb_max)
                o.status |= LK_FLOWBACK;
        }
    }
    if (live && ln.vec && ln.j == 0)
        store(a.out + pt, o);
}

template <int HMAX, int WX>
int dispatch3(vtg_ctx *c, const LkArgs &a, int storage, dim3 grid, dim3 block)
{
    const size_t lds = (size_t)a.waves * Shape3<HMAX, WX>::WAVE_BYTES;
    switch (storage) {
    case VTG_STORE_Q6:
        hipLaunchKernelGGL((lk3_kernel<HMAX, WX, uint8_t, uint16_t>), grid, block, lds, c->stream, a);
        break;
    case VTG_STORE_F32:
    case VTG_STORE_F32_U8:
        hipLaunchKernelGGL((lk3_kernel<HMAX, WX, float, float>), grid, block, lds, c->stream, a);
        break;
    case VTG_STORE_Q6W:
        hipLaunchKernelGGL((lk3_kernel<HMAX, WX, uint16_t, uint16_t>), grid, block, lds, c->stream, a);
        break;
    case VTG_STORE_Q6F:
        hipLaunchKernelGGL((lk3_kernel<HMAX, WX, uint8_t, float>), grid, block, lds, c->stream, a);
        break;
    case VTG_STORE_F16:
        hipLaunchKernelGGL((lk3_kernel<HMAX, WX, uint8_t, __half>), grid, block, lds, c->stream, a);
        break;
    case VTG_STORE_BF16:
        hipLaunchKernelGGL((lk3_kernel<HMAX, WX, uint8_t, __hip_bfloat16>), grid, block, lds, c->stream, a);
        break;
    default:
        return fail(c, VTG_EINVAL, "storage %d", storage);
    }
    VTG_HIP(c, hipGetLastError());
    return VTG_OK;
}

// The compiled window shapes: the windows of every workload here, 15 x 15 and 21 x 21, have their own exact shapes
// (straight-line code); any other window runs the guarded form of its block count with 31 rows.
template <int NV, int HMAX, int WX>
int dispatch_storage(vtg_ctx *c, const LkArgs &a, int storage, dim3 grid, dim3 block)
{
    const size_t lds = (size_t)a.waves * PPW * Shape<NV, HMAX, WX>::LDS_BYTES;
    switch (storage) {
    case VTG_STORE_Q6:
        hipLaunchKernelGGL((lk_kernel<NV, HMAX, WX, uint8_t, uint16_t>), grid, block, lds, c->stream, a);
        break;
    case VTG_STORE_F32:
    case VTG_STORE_F32_U8:
        hipLaunchKernelGGL((lk_kernel<NV, HMAX, WX, float, float>), grid, block, lds, c->stream, a);
        break;
    case VTG_STORE_Q6W:
        hipLaunchKernelGGL((lk_kernel<NV, HMAX, WX, uint16_t, uint16_t>), grid, block, lds, c->stream, a);
        break;
    case VTG_STORE_Q6F:
        hipLaunchKernelGGL((lk_kernel<NV, HMAX, WX, uint8_t, float>), grid, block, lds, c->stream, a);
        break;
    case VTG_STORE_F16:
        hipLaunchKernelGGL((lk_kernel<NV, HMAX, WX, uint8_t, __half>), grid, block, lds, c->stream, a);
        break;
    case VTG_STORE_BF16:
        hipLaunchKernelGGL((lk_kernel<NV, HMAX, WX, uint8_t, __hip_bfloat16>), grid, block, lds, c->stream, a);
        break;
    default:
        return fail(c, VTG_EINVAL, "storage %d", storage);
    }
    VTG_HIP(c, hipGetLastError());
    return VTG_OK;
}

// LDS per point of the shape a window runs in.
int shape_lds_point(const Cfg &cfg)
{
    if (cfg.w == 15 && cfg.h == 15)
        return Shape<1, 15, 15>::LDS_BYTES;
    if (cfg.w == 21 && cfg.h == 21)
        return Shape<2, 21, 21>::LDS_BYTES;
    return 0;
}

// The window shapes compiled in the three-points-per-wavefront layout.
bool three_shape(const Cfg &cfg) { return cfg.w == 21 && cfg.h == 21; }
bool three_shape(const lk_cfg &cfg) { return cfg.w == 21 && cfg.h == 21; }

int dispatch_shape(vtg_ctx *c, const LkArgs &a, int storage, dim3 grid, dim3 block)
{
    const int nv = (a.cfg.w + 15) / 16;
    if (a.cfg.w == 15 && a.cfg.h == 15)
        return dispatch_storage<1, 15, 15>(c, a, storage, grid, block);
    if (a.cfg.w == 21 && a.cfg.h == 21)
        return dispatch_storage<2, 21, 21>(c, a, storage, grid, block);
    if (nv == 1)
        return dispatch_storage<1, 31, 0>(c, a, storage, grid, block);
    if (nv == 2)
        return dispatch_storage<2, 31, 0>(c, a, storage, grid, block);
    return dispatch_storage<3, 31, 0>(c, a, storage, grid, block);
}

// ---------------------------------------------------------------- fused pyramid + LK, one persistent launch

// The batched tk_build + OP_LK (track.c:584-629, 275-292) as one launch: work items in the order of their
// dependencies: bands of 16 output rows of every pair's level 1 (which also write the frames into level 0), then of
// every pair's level 2, ..., then the points in blocks of 4 (one wavefront each). Workgroups take items from a counter
// in that order (a workgroup's first point block turns its four waves into independent takers of one block each, as
// no band follows). A level-l band waits until its pyramid's level l - 1 is complete; a wave's first block waits
// until every pyramid is complete, then the wave acquires once for all its blocks (an agent-scope acquire invalidates
// the XCD's L2, which LK's overlapping windows live in, so one per pair change costs far more than the wait). Every
// item waited for was taken before the waiter's own, by a workgroup that is running and waits only for earlier items,
// so the earliest unfinished item always progresses. There is one queue per XCD (workgroup w serves queue w mod 8, the
// XCD it runs on), each in that same order of classes (level 1, level 2, ..., then the points): a level's bands are
// dealt round-robin over the queues and the queue's point blocks are the q-th of 8 contiguous runs of all blocks, so
// an XCD's LK windows stay in a few pairs. The earliest unfinished class has all its dependencies finished and its
// items in every queue either finished or held by running workgroups (the grid is at most what is resident at once),
// so the queues cannot deadlock either.
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

template <int NV, int HMAX, int WX>
__global__ __launch_bounds__(256) void fused_kernel(FusedArgs a)
{
    __shared__ int s_item;
    // this workgroup's queue: first the bands of every pair's level 1, then of level 2, ... (qa[(l - 1) np + k]: the
    // first item of level l of pair k, both sides), then the q-th of 8 contiguous runs of all point blocks
    __shared__ int qa[(kMaxLevels - 1) * FB + 1];
    __shared__ int lds[rt::tile_lds_elems<FBAND>()];
    extern __shared__ float4 lk_lds[];
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
    const int blocks = a.lk.slots / PPW, run = (blocks + FQ - 1) / FQ, bfirst = q * run;
    const int blast = bfirst + run < blocks ? bfirst + run : blocks;
    const int total = nbands + (blast > bfirst ? blast - bfirst : 0);
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
        // Every later item of this queue is a block of 4 points: the waves now take them one at a time, each on its
        // own (wave 0 starts with the item just taken), and no longer meet at barriers.
        int it = wave == 0 ? item : -1;
        bool ready = false;
        const Lanes<NV> ln = lanes<NV>(a.lk.cfg, lane & 15);
        for (;;) {
            if (it < 0)
                it = wave_fetch_add(a.ctr + q, 1, lane);
            if (it >= total)
                return;
            const int s0 = (bfirst + (it - nbands)) * PPW; // the block's first slot in the numbering of all pairs
            const int k = __builtin_amdgcn_readfirstlane(pair_of(a.lk.pre, np, s0));
            if (!ready) {
                // every pyramid complete, then one acquire for all the wave's blocks (an agent-scope acquire
                // invalidates the XCD's L2, which LK's overlapping windows live in)
                wait_count(a.ctr + FTOT, a.total_bands, err);
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                ready = true;
            }
            it = -1;
            const int local = s0 - a.lk.pre[k] + (lane >> 4);
            const bool live = local < a.lk.cnt[k];
            const int pt = a.lk.start[k] + (live ? local : 0);
            float x = -1e30f, y = -1e30f, gx = 0.0f, gy = 0.0f;
            if (live) {
                x = a.lk.xy[2 * pt];
                y = a.lk.xy[2 * pt + 1];
                if (a.lk.guess) {
                    gx = a.lk.guess[2 * pt];
                    gy = a.lk.guess[2 * pt + 1];
                }
            }
            const int top = a.lk.top[k];
            char *pl = (char *)lk_lds + (wave * PPW + (lane >> 4)) * Shape<NV, HMAX, WX>::LDS_BYTES;
            Track o;
            track_fb<NV, HMAX, WX, uint8_t, uint16_t>(*a.lk.I[k], *a.lk.J[k], top, top, a.lk.cfg,
                                                      a.lk.solve == VTG_SOLVE_MIXED, a.lk.fb_max, ln, pl, x, y, gx,
                                                      gy, o);
            if (live && (lane & 15) == 0)
                store(a.lk.out + pt, o);
        }
    }
}

template <int NV, int HMAX, int WX>
int launch_fused_t(vtg_ctx *c, const FusedArgs &a, int total)
{
    static int per_cu = 0, cus = 0; // resident workgroups per CU of this kernel, CUs of device 0 class
    const size_t lds = (size_t)4 * PPW * Shape<NV, HMAX, WX>::LDS_BYTES;
    if (!per_cu) {
        VTG_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fused_kernel<NV, HMAX, WX>, 256, lds));
        VTG_HIP(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
        if (per_cu < 1)
            per_cu = 1;
    }
    // every queue needs workgroups: a multiple of FQ, at least FQ
    const int grid = std::max(FQ, std::min((total + FQ - 1) / FQ * FQ, per_cu * cus / FQ * FQ));
    hipEvent_t ev;
    timer_begin(c, "pyr_lk_fused", &ev);
    hipLaunchKernelGGL((fused_kernel<NV, HMAX, WX>), dim3(grid), dim3(256), lds, c->stream, a);
    VTG_HIP(c, hipGetLastError());
    timer_end(c, "pyr_lk_fused", ev);
    return VTG_OK;
}

int launch_fused_shape(vtg_ctx *c, const FusedArgs &a, int total)
{
    const int nv = (a.lk.cfg.w + 15) / 16;
    if (a.lk.cfg.w == 15 && a.lk.cfg.h == 15)
        return launch_fused_t<1, 15, 15>(c, a, total);
    if (a.lk.cfg.w == 21 && a.lk.cfg.h == 21)
        return launch_fused_t<2, 21, 21>(c, a, total);
    if (nv == 1)
        return launch_fused_t<1, 31, 0>(c, a, total);
    if (nv == 2)
        return launch_fused_t<2, 31, 0>(c, a, total);
    return launch_fused_t<3, 31, 0>(c, a, total);
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

Cfg to_cfg(const vtg_ctx *c, const lk_cfg *cfg)
{
    return Cfg{cfg->w, cfg->h, cfg->max_iter, cfg->eps, cfg->min_eig, cfg->max_residual, c->lk_grad};
}

// ---------------------------------------------------------------- a launch's points in spatial order
// The points of each pair are tracked in the order of the 32 x 32-pixel cell that holds their level-0 position (cells
// row by row; points of one cell in any order), so a wavefront's points are neighbours: their windows share cache lines
// and their iteration counts are alike. A record depends on its own point alone, so the order changes time only. A
// counting sort: clear the cells' counts, count (each point's rank in its cell from the atomic add), scan each pair's
// counts, scatter. Every kernel is a plain launch (no memset), so the launch can be captured in a graph.
constexpr int OCELL = 5; // log2 of the cell size in pixels

struct OrderArgs {
    int npairs, total;
    int start[BMAX], pbase[BMAX + 1], cbase[BMAX + 1], cw[BMAX], ch[BMAX];
    const float *xy;
    int *cells, *rank, *order;
};

__device__ __forceinline__ int order_pair(const OrderArgs &a, int i)
{
    int lo = 0, hi = a.npairs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (i >= a.pbase[mid])
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// The cell of point g of pair q (positions outside the frame, or not numbers, go to its border cells).
__device__ __forceinline__ int order_cell(const OrderArgs &a, int q, int g)
{
    const float x = fminf(fmaxf(a.xy[2 * g], 0.0f), (float)((a.cw[q] << OCELL) - 1));
    const float y = fminf(fmaxf(a.xy[2 * g + 1], 0.0f), (float)((a.ch[q] << OCELL) - 1));
    return a.cbase[q] + ((int)y >> OCELL) * a.cw[q] + ((int)x >> OCELL);
}

__global__ __launch_bounds__(256) void order_clear(int *cells, int n)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        cells[i] = 0;
}

__global__ __launch_bounds__(256) void order_count(OrderArgs a)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.total)
        return;
    const int q = order_pair(a, i), g = a.start[q] + (i - a.pbase[q]);
    a.rank[i] = atomicAdd(&a.cells[order_cell(a, q, g)], 1);
}

// The exclusive scan of each pair's cell counts in place: one workgroup per pair, 8 cells a thread per round.
__global__ __launch_bounds__(256) void order_scan(OrderArgs a)
{
    __shared__ int part[256];
    const int q = blockIdx.x, t = threadIdx.x;
    int *c = a.cells + a.cbase[q];
    const int n = a.cbase[q + 1] - a.cbase[q];
    int carry = 0;
    for (int b = 0; b < n; b += 256 * 8) {
        int v[8], s = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int i = b + t * 8 + k;
            v[k] = i < n ? c[i] : 0;
            s += v[k];
        }
        part[t] = s;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const int u = t >= off ? part[t - off] : 0;
            __syncthreads();
            part[t] += u;
            __syncthreads();
        }
        int run = carry + part[t] - s;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int i = b + t * 8 + k;
            if (i < n)
                c[i] = run;
            run += v[k];
        }
        carry += part[255];
        __syncthreads(); // part is rewritten next round
    }
}

__global__ __launch_bounds__(256) void order_scatter(OrderArgs a)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.total)
        return;
    const int q = order_pair(a, i), g = a.start[q] + (i - a.pbase[q]);
    a.order[a.pbase[q] + a.cells[order_cell(a, q, g)] + a.rank[i]] = g;
}

// The order of launch a's points (pairs sel, level-0 sizes from I) into the context's LK scratch; sets a.order and
// a.obase.
int order_points(vtg_ctx *c, LkArgs &a, const std::vector<int> &sel, const vtg_pyr *const *I)
{
    OrderArgs o;
    o.npairs = a.npairs;
    int total = 0, cells = 0;
    for (int q = 0; q < a.npairs; q++) {
        const vtg_pyr *p = I[sel[q]];
        o.start[q] = a.start[q];
        o.pbase[q] = total;
        o.cbase[q] = cells;
        o.cw[q] = (p->w + (1 << OCELL) - 1) >> OCELL;
        o.ch[q] = (p->h + (1 << OCELL) - 1) >> OCELL;
        total += a.cnt[q];
        cells += o.cw[q] * o.ch[q];
        a.obase[q] = o.pbase[q];
    }
    o.pbase[a.npairs] = total;
    o.cbase[a.npairs] = cells;
    o.total = total;
    o.xy = a.xy;
    void *buf = nullptr;
    const int rc = lk_scratch(c, ((size_t)cells + 2 * (size_t)total) * sizeof(int), &buf);
    if (rc)
        return rc;
    o.cells = (int *)buf;
    o.rank = o.cells + cells;
    o.order = o.rank + total;
    const int gp = (total + 255) / 256;
    hipEvent_t ev;
    timer_begin(c, "lk_order", &ev);
    hipLaunchKernelGGL(order_clear, dim3(std::min((cells + 255) / 256, 1024)), dim3(256), 0, c->stream, o.cells,
                       cells);
    hipLaunchKernelGGL(order_count, dim3(gp), dim3(256), 0, c->stream, o);
    hipLaunchKernelGGL(order_scan, dim3(o.npairs), dim3(256), 0, c->stream, o);
    hipLaunchKernelGGL(order_scatter, dim3(gp), dim3(256), 0, c->stream, o);
    timer_end(c, "lk_order", ev);
    VTG_HIP(c, hipGetLastError());
    a.order = o.order;
    return VTG_OK;
}

}  // namespace

// One launch for pairs sel[0 .. m) (m <= BMAX; indices into the caller's arrays) with `waves` wavefronts per workgroup.
static int launch_group(vtg_ctx *c, const std::vector<int> &sel, const vtg_pyr *const *I, const vtg_pyr *const *J,
                        const int *top, const int *off, const lk_cfg *cfg, const float *xy, const float *guess,
                        float fb_max, int solve_mode, lk_out *out, int waves, int ppw, int fb_top)
{
    LkArgs a;
    a.npairs = (int)sel.size();
    // points per wavefront: 3 (one lane per window column) where the window's shape is compiled in that layout, else 4
    const bool three = ppw != PPW && three_shape(*cfg);
    const int per = three ? PPW3 : PPW;
    int acc = 0;
    for (int q = 0; q < a.npairs; q++) {
        const int k = sel[q];
        a.I[q] = I[k]->dview;
        a.J[q] = J[k]->dview;
        a.top[q] = top[k];
        a.start[q] = off[k];
        a.cnt[q] = off[k + 1] - off[k];
        a.pre[q] = acc;
        acc += (a.cnt[q] + per - 1) / per * per;
    }
    a.pre[a.npairs] = acc;
    a.slots = acc;
    if (a.slots <= 0)
        return VTG_OK;
    a.cfg = to_cfg(c, cfg);
    a.solve = solve_mode;
    a.xy = xy;
    a.guess = guess;
    a.fb_max = fb_max;
    a.fb_top = fb_top;
    a.out = out;
    a.order = nullptr;
    if (c->lk_order) {
        const int rc = order_points(c, a, sel, I);
        if (rc)
            return rc;
    }
    if (waves < 1 || waves > 8)
        return fail(c, VTG_EINVAL, "vtg_lk_track: %d waves per workgroup (1..8)", waves);
    // fewer wavefronts per workgroup where the shape's LDS would pass the device's per-workgroup limit (the schedule
    // changes, the records do not)
    static int max_lds = 0;
    if (!max_lds)
        VTG_HIP(c, hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
    const int per_wave = three ? Shape3<21, 21>::WAVE_BYTES : PPW * shape_lds_point(a.cfg);
    while (waves > 1 && per_wave * waves > max_lds)
        waves--;
    a.waves = waves;
    const int wgs = (a.slots / per + a.waves - 1) / a.waves;
    a.per_xcd = (wgs + NXCD - 1) / NXCD;
    const dim3 grid(a.per_xcd * NXCD), block(64 * a.waves);
    hipEvent_t ev;
    timer_begin(c, "lk_track", &ev);
    const int rc = three ? dispatch3<21, 21>(c, a, I[sel[0]]->storage, grid, block)
                         : dispatch_shape(c, a, I[sel[0]]->storage, grid, block);
    timer_end(c, "lk_track", ev);
    return rc;
}

int lk_default_ppw(const lk_cfg *cfg) { return three_shape(*cfg) ? PPW3 : PPW; }

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
    return LkSched{c->lk_waves, c->lk_ppw};
}

// Pairs [0, np) with points off[k] .. off[k+1] - 1 of xy, one launch per BMAX pairs. The workgroup size is `sched`'s
// when given; otherwise the size-class pick of the pair holding the most points (the context default when its class
// was not tuned) serves the whole batch, so a batch is never split by schedule.
int launch_lk_many(vtg_ctx *c, int np, const vtg_pyr *const *I, const vtg_pyr *const *J, const int *top,
                   const int *off, const lk_cfg *cfg, const float *xy, const float *guess, float fb_max,
                   int solve_mode, lk_out *out, const LkSched *sched, int fb_top)
{
    LkSched s{c->lk_waves, c->lk_ppw};
    if (sched) {
        s = *sched;
    } else {
        int most = -1;
        for (int k = 0; k < np; k++)
            if (off[k + 1] - off[k] > most) {
                most = off[k + 1] - off[k];
                s = tuned_lk(c, I[k], top[k], cfg);
            }
    }
    for (int b = 0; b < np; b += BMAX) {
        std::vector<int> part;
        for (int k = b; k < std::min(np, b + BMAX); k++)
            part.push_back(k);
        const int rc =
            launch_group(c, part, I, J, top, off, cfg, xy, guess, fb_max, solve_mode, out, s.waves, s.ppw, fb_top);
        if (rc)
            return rc;
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
        a.lk.cnt[k] = offsets[k + 1] - offsets[k];
        a.lk.pre[k] = acc;
        acc += (a.lk.cnt[k] + PPW - 1) / PPW * PPW;
        a.f0[k] = frame0[k];
        a.f1[k] = frame1[k];
        a.p0[k] = pitch0[k];
        a.p1[k] = pitch1[k];
        for (int l = 1; l < I[k]->levels; l++) {
            const int b = (I[k]->lv[l].h + FBAND - 1) / FBAND + (J[k]->lv[l].h + FBAND - 1) / FBAND;
            items += b;
            a.total_bands += b;
        }
    }
    items += acc / PPW; // point blocks, one wavefront each
    a.lk.pre[npairs] = acc;
    a.lk.slots = acc;
    if (acc > 0 && (!xy || !out))
        return fail(c, VTG_EINVAL, "vtg_pyr_lk_fused: null points or output");
    a.lk.cfg = to_cfg(c, cfg);
    a.lk.solve = solve_mode;
    a.lk.xy = xy;
    a.lk.guess = guess;
    a.lk.fb_max = fb_max;
    a.lk.fb_top = -1;
    a.lk.out = out;
    a.lk.order = nullptr;
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
    return launch_fused_shape(c, a, items);
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

int vtg_lk_set_waves(vtg_ctx *c, int waves)
{
    if (!c || waves < 1 || waves > 8)
        return fail(c, VTG_EINVAL, "vtg_lk_set_waves: %d (1..8)", waves);
    c->lk_waves = waves;
    return VTG_OK;
}

int vtg_lk_set_order(vtg_ctx *c, int on)
{
    if (!c || (on != 0 && on != 1))
        return fail(c, VTG_EINVAL, "vtg_lk_set_order: %d (0 or 1)", on);
    c->lk_order = on;
    return VTG_OK;
}

int vtg_lk_set_gradient(vtg_ctx *c, int grad)
{
    if (!c || (grad != VTG_GRAD_SOBEL && grad != VTG_GRAD_SCHARR))
        return fail(c, VTG_EINVAL, "vtg_lk_set_gradient: %d (VTG_GRAD_SOBEL or VTG_GRAD_SCHARR)", grad);
    c->lk_grad = grad;
    return VTG_OK;
}

int vtg_lk_set_ppw(vtg_ctx *c, int ppw)
{
    if (!c || (ppw != 0 && ppw != 3 && ppw != 4))
        return fail(c, VTG_EINVAL, "vtg_lk_set_ppw: %d (0, 3 or 4)", ppw);
    c->lk_ppw = ppw;
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
