// filters.hip: the tracker's optional level-0 filters on the GPU, the contrast stretch (img_minmax + img_norm,
// src/kern/img_ref.c:42-71, run by tk_build when b_contrast is set, src/kern/track.c:606-617) and the pre-blur
// (img_gauss5, img_ref.c:83-96, b_pre_gauss, track.c:618-622): the pyramid build with them, and the reference-shaped
// calls over host planes.
//
// With either filter level 0 is no longer 64 x a byte, so it is stored in 16 bits (VTG_STORE_Q6W: every level u16 Q6,
// exact) or as fp32 (VTG_STORE_F32). The stretch of a frame follows tk_build: min and max of the frame's Q6 values,
// gain = (float)(16320 / span) and offset = (float)(-min * gain) in long double (gain 1 for a flat frame), then every
// value clamp(rint(s * gain + offset), 0, 16320) with the reference's roundings (ldnorm.hpp). A u8 frame has 256 values,
// so its stretch is a 256-entry table: one workgroup per frame reduces the min / max partials and fills the table, and
// the level-0 kernel looks every byte up. The gains and offsets for every (span, min) of u8 frames are computed once
// per context on the host in long double (x86-64 long double is the x87 80-bit type the reference computes in) and
// kept in device memory. The blur is (sum of (1 4 6 4 1)^2 x value over the 5 x 5 + 128) >> 8 in exact integers with
// the border replicated (the reference's padding).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "ldnorm.hpp"
#include "vtg_internal.hpp"

namespace vtg {

namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ---------------------------------------------------------------- u8 frames: min / max and the stretch table

constexpr int MM_BLOCKS = 64; // partial min / max per frame

struct MJob {
    const uint8_t *frame;
    int pitch, w, h;
    uint16_t *lut;   // 256 entries: the stretched Q6 value of each byte
    uint8_t *part;   // MM_BLOCKS (min, max) byte pairs
};
constexpr int FMAX = 64; // frames per launch
struct MJobs {
    MJob j[FMAX];
    int n;
};

// Grid: (MM_BLOCKS, frames). Block b takes rows b, b + MM_BLOCKS, ...; 16 bytes per thread per step.
__global__ __launch_bounds__(256) void minmax_u8(MJobs js)
{
    const MJob &J = js.j[blockIdx.y];
    __shared__ unsigned smn[256], smx[256];
    unsigned mn = 255, mx = 0;
    for (int y = blockIdx.x; y < J.h; y += MM_BLOCKS) {
        const uint8_t *row = J.frame + (size_t)y * J.pitch;
        for (int x = threadIdx.x; x < J.w; x += 256) {
            const unsigned v = row[x];
            mn = v < mn ? v : mn;
            mx = v > mx ? v : mx;
        }
    }
    smn[threadIdx.x] = mn;
    smx[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smn[threadIdx.x] = min(smn[threadIdx.x], smn[threadIdx.x + s]);
            smx[threadIdx.x] = max(smx[threadIdx.x], smx[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        J.part[2 * blockIdx.x] = (uint8_t)smn[0];
        J.part[2 * blockIdx.x + 1] = (uint8_t)smx[0];
    }
}

// Grid: frames, 256 threads. gain_tab[k] and off_tab[k * 256 + m]: the stretch of a frame whose bytes span m .. m + k.
__global__ __launch_bounds__(256) void stretch_table(MJobs js, const float *gain_tab, const float *off_tab)
{
    const MJob &J = js.j[blockIdx.x];
    __shared__ unsigned mn, mx;
    if (threadIdx.x == 0) {
        unsigned a = 255, b = 0;
        for (int i = 0; i < MM_BLOCKS; i++) {
            a = min(a, (unsigned)J.part[2 * i]);
            b = max(b, (unsigned)J.part[2 * i + 1]);
        }
        mn = a;
        mx = b;
    }
    __syncthreads();
    const unsigned k = mx - mn, m = mn;
    const float g = gain_tab[k], o = off_tab[k * 256 + m];
    J.lut[threadIdx.x] = (uint16_t)norm_exact(64.0f * (float)threadIdx.x, g, o);
}

// ---------------------------------------------------------------- level 0 from u8 frames, filtered

constexpr int LT_W = 64, LT_H = 16; // output tile

struct LJob {
    const uint8_t *frame;
    int pitch, w, h;
    void *dst;
    int dpitch;
    const uint16_t *lut; // null: 64 x byte
    int tiles_x, t0;
};
struct LJobs {
    LJob j[FMAX];
    int n;
};

template <typename TD> struct Put;
template <> struct Put<uint16_t> {
    static __device__ __forceinline__ uint16_t v(int q) { return (uint16_t)q; }
};
template <> struct Put<float> {
    static __device__ __forceinline__ float v(int q) { return (float)q; }
};

// One 64 x 16 tile of level 0 per workgroup: the (16 + 4) x (64 + 4) bytes around it (coordinates clamped), each
// through the stretch table (or x 64) into LDS as integers; GAUSS: 5 rows then 5 columns, (h + 128) >> 8.
template <typename TD, bool GAUSS>
__global__ __launch_bounds__(256) void level0_filter(LJobs js)
{
    constexpr int IW = LT_W + 4, IH = LT_H + 4;
    __shared__ int tin[IH][IW + 1];
    __shared__ int tv[LT_H][IW + 1];
    __shared__ uint16_t lut[256];
    int jn = 0;
    for (int k = 1; k < js.n; k++)
        if ((int)blockIdx.x >= js.j[k].t0)
            jn = k;
    const LJob &J = js.j[jn];
    const int tile = blockIdx.x - J.t0, tyi = tile / J.tiles_x, txi = tile - tyi * J.tiles_x;
    const int ox0 = txi * LT_W, oy0 = tyi * LT_H;
    lut[threadIdx.x] = J.lut ? J.lut[threadIdx.x] : (uint16_t)(64 * threadIdx.x);
    constexpr int NIN = IH * IW, NLD = (NIN + 255) / 256;
    uint8_t raw[NLD];
#pragma unroll
    for (int k = 0; k < NLD; k++) {
        const int i = threadIdx.x + 256 * k, ic = i < NIN ? i : NIN - 1;
        const int r = ic / IW, cc = ic - r * IW;
        raw[k] = J.frame[(size_t)clampi(oy0 - 2 + r, 0, J.h - 1) * J.pitch + clampi(ox0 - 2 + cc, 0, J.w - 1)];
    }
    __syncthreads(); // the table is in LDS
#pragma unroll
    for (int k = 0; k < NLD; k++) {
        const int i = threadIdx.x + 256 * k;
        if (i < NIN)
            tin[i / IW][i % IW] = lut[raw[k]];
    }
    __syncthreads();
    TD *dst = (TD *)J.dst;
    if (GAUSS) {
        for (int i = threadIdx.x; i < LT_H * IW; i += 256) {
            const int j = i / IW, cc = i - j * IW;
            tv[j][cc] = tin[j][cc] + tin[j + 4][cc] + 4 * (tin[j + 1][cc] + tin[j + 3][cc]) + 6 * tin[j + 2][cc];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < LT_H * LT_W / 256; k++) {
        const int i = threadIdx.x + 256 * k;
        const int j = i / LT_W, kk = i - j * LT_W;
        const int ox = ox0 + kk, oy = oy0 + j;
        if (ox < J.w && oy < J.h) {
            int q;
            if (GAUSS) {
                const int h = tv[j][kk] + tv[j][kk + 4] + 4 * (tv[j][kk + 1] + tv[j][kk + 3]) + 6 * tv[j][kk + 2];
                q = (h + 128) >> 8;
            } else {
                q = tin[j + 2][kk + 2];
            }
            dst[(size_t)oy * J.dpitch + ox] = Put<TD>::v(q);
        }
    }
}

// ---------------------------------------------------------------- fp32 planes (the host-plane calls)

// The 5 x 5 binomial of rows y0 .. y1 - 1 of an fp32 Q6 level in device memory (pitch in floats; coordinates clamped:
// the replicated padding), in the reference's order with every sum and product rounded on its own (the reference is
// built without contraction): per column r0 + r4 + 4 (r1 + r3) + 6 r2, then the same across, (lrintf(h) + 128) >> 8.
__global__ __launch_bounds__(256) void gauss5_f32(const float *src, int sp, int w, int h, float *dst, int dp, int y0)
{
    const int x = blockIdx.x * 256 + threadIdx.x, y = y0 + blockIdx.y;
    if (x >= w)
        return;
    const float *r[5];
#pragma unroll
    for (int i = 0; i < 5; i++)
        r[i] = src + (size_t)clampi(y + i - 2, 0, h - 1) * sp;
    float v[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const int xx = clampi(x + i - 2, 0, w - 1);
        v[i] = __fadd_rn(__fadd_rn(__fadd_rn(r[0][xx], r[4][xx]), __fmul_rn(4.0f, __fadd_rn(r[1][xx], r[3][xx]))),
                         __fmul_rn(6.0f, r[2][xx]));
    }
    const float s = __fadd_rn(__fadd_rn(__fadd_rn(v[0], v[4]), __fmul_rn(4.0f, __fadd_rn(v[1], v[3]))),
                              __fmul_rn(6.0f, v[2]));
    dst[(size_t)y * dp + x] = (float)(((long long)__float2ll_rn(s) + 128) >> 8);
}

__global__ __launch_bounds__(256) void norm_f32(const float *src, int sp, int w, float *dst, int dp, int y0, float g,
                                                float o)
{
    const int x = blockIdx.x * 256 + threadIdx.x, y = y0 + blockIdx.y;
    if (x < w)
        dst[(size_t)y * dp + x] = norm_exact(src[(size_t)y * sp + x], g, o);
}

// min / max of rows y0 .. y1 - 1: grid MM_BLOCKS blocks of partials, then one block reduces them.
__global__ __launch_bounds__(256) void minmax_f32(const float *src, int sp, int w, int y0, int y1, float *part)
{
    __shared__ float smn[256], smx[256];
    float mn = src[(size_t)y0 * sp], mx = mn;
    for (int y = y0 + blockIdx.x; y < y1; y += MM_BLOCKS)
        for (int x = threadIdx.x; x < w; x += 256) {
            const float v = src[(size_t)y * sp + x];
            mn = v < mn ? v : mn;
            mx = v > mx ? v : mx;
        }
    smn[threadIdx.x] = mn;
    smx[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smn[threadIdx.x] = fminf(smn[threadIdx.x], smn[threadIdx.x + s]);
            smx[threadIdx.x] = fmaxf(smx[threadIdx.x], smx[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = smn[0];
        part[2 * blockIdx.x + 1] = smx[0];
    }
}

__global__ void minmax_final(const float *part, float *out)
{
    float mn = part[0], mx = part[1];
    for (int i = 1; i < MM_BLOCKS; i++) {
        mn = fminf(mn, part[2 * i]);
        mx = fmaxf(mx, part[2 * i + 1]);
    }
    out[0] = mn;
    out[1] = mx;
}

// The context's stretch tables (gain per span, offset per span and min, for u8 frames) and its filter buffer: built
// on first use. Long double here is the host's x87 80-bit type, the reference's `real`.
int filter_tables(vtg_ctx *c)
{
    if (c->filt)
        return VTG_OK;
    std::vector<float> tab(256 + 256 * 256);
    for (int k = 0; k < 256; k++) {
        const long double span = (long double)(64.0f * (float)k);
        const long double gain = span > 0.0L ? 16320.0L / span : 1.0L;
        tab[k] = (float)gain;
        for (int m = 0; m < 256; m++)
            tab[256 + k * 256 + m] = (float)(-(long double)(64.0f * (float)m) * gain);
    }
    const size_t tab_bytes = tab.size() * sizeof(float);
    const size_t bytes = tab_bytes + (size_t)FMAX * (256 * sizeof(uint16_t) + 2 * MM_BLOCKS);
    void *p = nullptr;
    VTG_HIP(c, hipMalloc(&p, bytes));
    hipError_t e = hipMemcpy(p, tab.data(), tab_bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(p);
        return hip_fail(c, e, "filter tables");
    }
    c->filt = p;
    return VTG_OK;
}

template <typename TD>
int launch_level0(vtg_ctx *c, const std::vector<LJob> &jobs, bool gauss)
{
    for (size_t b = 0; b < jobs.size(); b += FMAX) {
        LJobs js{};
        int tiles = 0;
        js.n = (int)std::min<size_t>(FMAX, jobs.size() - b);
        for (int k = 0; k < js.n; k++) {
            LJob j = jobs[b + k];
            j.tiles_x = (j.w + LT_W - 1) / LT_W;
            j.t0 = tiles;
            tiles += j.tiles_x * ((j.h + LT_H - 1) / LT_H);
            js.j[k] = j;
        }
        hipEvent_t ev;
        timer_begin(c, "pyr_level0", &ev);
        if (gauss)
            hipLaunchKernelGGL((level0_filter<TD, true>), dim3(tiles), dim3(256), 0, c->stream, js);
        else
            hipLaunchKernelGGL((level0_filter<TD, false>), dim3(tiles), dim3(256), 0, c->stream, js);
        VTG_HIP(c, hipGetLastError());
        timer_end(c, "pyr_level0", ev);
    }
    return VTG_OK;
}

}  // namespace

// Level 0 of pyramids ps[0 .. n) (storage Q6W or F32) from their device u8 frames, with the filters `flags` asks for
// (VTG_BUILD_CONTRAST, VTG_BUILD_PRE_GAUSS). Stream-ordered; no allocation once the context's tables exist.
int launch_level0_many(vtg_ctx *c, vtg_pyr *const *ps, int n, const uint8_t *const *frames, const int *pitch,
                       int flags)
{
    const bool contrast = flags & VTG_BUILD_CONTRAST, gauss = flags & VTG_BUILD_PRE_GAUSS;
    int rc = contrast ? filter_tables(c) : VTG_OK;
    if (rc)
        return rc;
    const float *gain_tab = (const float *)c->filt, *off_tab = gain_tab + 256;
    uint16_t *luts = (uint16_t *)(gain_tab + 256 + 256 * 256);
    uint8_t *parts = (uint8_t *)(luts + (size_t)FMAX * 256);
    for (int b = 0; b < n; b += FMAX) {
        const int m = std::min(FMAX, n - b);
        std::vector<LJob> jobs;
        MJobs mj{};
        mj.n = m;
        for (int k = 0; k < m; k++) {
            const vtg_pyr *p = ps[b + k];
            mj.j[k] = MJob{frames[b + k], pitch[b + k], p->w, p->h, luts + (size_t)k * 256,
                           parts + (size_t)k * 2 * MM_BLOCKS};
            jobs.push_back(LJob{frames[b + k], pitch[b + k], p->w, p->h, (void *)p->lv[0].p, p->lv[0].pitch,
                                contrast ? luts + (size_t)k * 256 : nullptr, 0, 0});
        }
        if (contrast) {
            hipEvent_t ev;
            timer_begin(c, "pyr_stretch", &ev);
            hipLaunchKernelGGL(minmax_u8, dim3(MM_BLOCKS, m), dim3(256), 0, c->stream, mj);
            hipLaunchKernelGGL(stretch_table, dim3(m), dim3(256), 0, c->stream, mj, gain_tab, off_tab);
            VTG_HIP(c, hipGetLastError());
            timer_end(c, "pyr_stretch", ev);
        }
        rc = ps[b]->storage == VTG_STORE_F32 ? launch_level0<float>(c, jobs, gauss)
                                             : launch_level0<uint16_t>(c, jobs, gauss);
        if (rc)
            return rc;
    }
    return VTG_OK;
}

}  // namespace vtg

using namespace vtg;

// ---- reference-shaped calls over host planes: copy in, run, copy the written rows (with their padding) out

namespace {

// The padded host plane `pl` uploaded into device memory at `dev`; returns the device address of pixel (0, 0).
float *upload_plane(vtg_ctx *c, const vt_plane *pl, void *dev, hipError_t *e)
{
    const float *base = pl->p - (size_t)pl->stride * VT_PAD_Y - VT_PAD_X;
    *e = hipMemcpyAsync(dev, base, plane_bytes(pl->w, pl->h), hipMemcpyHostToDevice, c->stream);
    return (float *)dev + (size_t)pl->stride * VT_PAD_Y + VT_PAD_X;
}

__global__ __launch_bounds__(256) void pad_rows_f32(float *pix, int stride, int w, int y0)
{
    // the left and right padding lanes of row y0 + blockIdx.y from its first and last pixel (img_pad_rows)
    const int y = y0 + blockIdx.y;
    float *r = pix + (size_t)y * stride;
    const int end = ((w + 15) & ~15) + VT_PAD_X;
    for (int x = (int)threadIdx.x - VT_PAD_X; x < end; x += 256)
        if (x < 0 || x >= w)
            r[x] = r[x < 0 ? 0 : w - 1];
}

}  // namespace

extern "C" {

int vtg_img_gauss5(vtg_ctx *c, const struct vt_plane *dst, const struct vt_plane *src, int y0, int y1)
{
    if (!c || !dst || !src || dst->w != src->w || dst->h != src->h || y0 < 0 || y1 > dst->h || y0 > y1)
        return fail(c, VTG_EINVAL, "vtg_img_gauss5: bad argument");
    if (y0 == y1)
        return VTG_OK;
    const size_t pb = (plane_bytes(src->w, src->h) + 255) & ~(size_t)255;
    void *tmp = nullptr;
    int rc = scratch(c, 2 * pb, &tmp);
    if (rc)
        return rc;
    hipError_t e;
    const float *s = upload_plane(c, src, tmp, &e);
    float *d = (float *)((char *)tmp + pb) + (size_t)dst->stride * VT_PAD_Y + VT_PAD_X;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(gauss5_f32, dim3((src->w + 255) / 256, y1 - y0), dim3(256), 0, c->stream, s, src->stride,
                           src->w, src->h, d, dst->stride, y0);
        hipLaunchKernelGGL(pad_rows_f32, dim3(1, y1 - y0), dim3(256), 0, c->stream, d, dst->stride, dst->w, y0);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(plane_row(dst, y0) - VT_PAD_X, d + (size_t)y0 * dst->stride - VT_PAD_X,
                           (size_t)(y1 - y0) * dst->stride * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? VTG_OK : hip_fail(c, e, "vtg_img_gauss5");
}

int vtg_img_norm(vtg_ctx *c, const struct vt_plane *dst, const struct vt_plane *src, const float *gain_offset, int y0,
                 int y1)
{
    if (!c || !dst || !src || !gain_offset || dst->w != src->w || dst->h != src->h || y0 < 0 || y1 > dst->h ||
        y0 > y1)
        return fail(c, VTG_EINVAL, "vtg_img_norm: bad argument");
    if (y0 == y1)
        return VTG_OK;
    const size_t pb = (plane_bytes(src->w, src->h) + 255) & ~(size_t)255;
    void *tmp = nullptr;
    int rc = scratch(c, 2 * pb, &tmp);
    if (rc)
        return rc;
    hipError_t e;
    const float *s = upload_plane(c, src, tmp, &e);
    float *d = (float *)((char *)tmp + pb) + (size_t)dst->stride * VT_PAD_Y + VT_PAD_X;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(norm_f32, dim3((src->w + 255) / 256, y1 - y0), dim3(256), 0, c->stream, s, src->stride,
                           src->w, d, dst->stride, y0, gain_offset[0], gain_offset[1]);
        hipLaunchKernelGGL(pad_rows_f32, dim3(1, y1 - y0), dim3(256), 0, c->stream, d, dst->stride, dst->w, y0);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(plane_row(dst, y0) - VT_PAD_X, d + (size_t)y0 * dst->stride - VT_PAD_X,
                           (size_t)(y1 - y0) * dst->stride * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? VTG_OK : hip_fail(c, e, "vtg_img_norm");
}

int vtg_img_minmax(vtg_ctx *c, const struct vt_plane *src, float *out, int y0, int y1)
{
    if (!c || !src || !out || y0 < 0 || y1 > src->h || y0 >= y1)
        return fail(c, VTG_EINVAL, "vtg_img_minmax: bad argument");
    const size_t pb = (plane_bytes(src->w, src->h) + 255) & ~(size_t)255;
    void *tmp = nullptr;
    int rc = scratch(c, pb + 2 * MM_BLOCKS * sizeof(float) + 256, &tmp);
    if (rc)
        return rc;
    hipError_t e;
    const float *s = upload_plane(c, src, tmp, &e);
    float *part = (float *)((char *)tmp + pb), *res = part + 2 * MM_BLOCKS;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(minmax_f32, dim3(MM_BLOCKS), dim3(256), 0, c->stream, s, src->stride, src->w, y0, y1,
                           part);
        hipLaunchKernelGGL(minmax_final, dim3(1), dim3(1), 0, c->stream, (const float *)part, res);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, res, 2 * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? VTG_OK : hip_fail(c, e, "vtg_img_minmax");
}

}  // extern "C"
