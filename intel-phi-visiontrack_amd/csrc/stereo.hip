// stereo.hip: correspondence along a rectified row (src/kern/stereo.c:1-80, stereo.h:9-20): a 16 x 11 zero-mean
// template of the left level at a feature against every integer disparity of the right row, the best peak and the best
// one at least three candidates away from it. One wavefront per feature; each lane scores every 64th candidate.
//
// The sums follow the reference's 16-lane model: per lane k a running fp32 sum down the 11 rows (fmaf where the model
// fuses), then the i^1, i^2, i^4, i^8 butterfly, so the three totals of each candidate (and the template's mean and
// norm: the mean's long double divide emulated exactly) are the reference's bits. The correlation from the three totals
// is long double in the reference and double here: the same argmax, the score within 1e-5.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <vector>

#include "ldnorm.hpp"
#include "vtg_internal.hpp"

namespace vtg {
namespace {

constexpr int SW = 16, SH = 11, SMAX = 512, SPL = SMAX / 64; // template, candidate cap, candidates per lane

template <typename T> struct Px;
template <> struct Px<uint8_t> {
    static __device__ float get(uint8_t v) { return 64.0f * (float)v; }
};
template <> struct Px<uint16_t> {
    static __device__ float get(uint16_t v) { return (float)v; }
};
template <> struct Px<float> {
    static __device__ float get(float v) { return v; }
};
template <> struct Px<__half> {
    static __device__ float get(__half v) { return 64.0f * __half2float(v); }
};
template <> struct Px<__hip_bfloat16> {
    static __device__ float get(__hip_bfloat16 v) { return 64.0f * __bfloat162float(v); }
};

struct ScanArgs {
    const void *left, *right; // the level of each side: w x h, pitch in elements
    int w, h, pl, pr;
    int n;
    const float *xy;
    int dmin, count;          // candidates dmin .. dmin + count - 1
    vtg_stereo_scan *out;
};

__device__ __forceinline__ float hsum16(const float (&v)[SW])
{
    return (((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]))) +
           (((v[8] + v[9]) + (v[10] + v[11])) + ((v[12] + v[13]) + (v[14] + v[15])));
}

__device__ __forceinline__ double shfl_d(double v, int m)
{
    const int lo = __shfl_xor(__double2loint(v), m), hi = __shfl_xor(__double2hiint(v), m);
    return __hiloint2double(hi, lo);
}

template <typename T>
__global__ __launch_bounds__(256) void stereo_scan(ScanArgs a)
{
    __shared__ float tpl[4][SH * SW];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + wave;
    const bool live = i < a.n; // a wave past the last feature still meets the barrier
    const int ii = live ? i : a.n - 1;
    const T *L = (const T *)a.left, *R = (const T *)a.right;
    const int x0 = (int)a.xy[2 * ii] - SW / 2, y0 = (int)a.xy[2 * ii + 1] - SH / 2;
    float *t = tpl[wave];
    for (int e = lane; e < SH * SW; e += 64) {
        const int r = e / SW, k = e % SW;
        const int yy = min(max(y0 + r, 0), a.h - 1), xx = min(max(x0 + k, 0), a.w - 1);
        t[e] = Px<T>::get(L[(size_t)yy * a.pl + xx]);
    }
    __syncthreads();
    if (!live)
        return;
    // the template's mean and norm (every lane computes them, from LDS)
    float col[SW];
#pragma unroll
    for (int k = 0; k < SW; k++) {
        float s = 0.0f;
        for (int r = 0; r < SH; r++)
            s += t[r * SW + k];
        col[k] = s;
    }
    const float mean = xl_div_to_float(xl_add(XL{0, 0}, hsum16(col)), SW * SH);
#pragma unroll
    for (int k = 0; k < SW; k++) {
        float s = 0.0f;
        for (int r = 0; r < SH; r++) {
            const float z = t[r * SW + k] - mean;
            s = fmaf(z, z, s);
        }
        col[k] = s;
    }
    const double norm_t = (double)hsum16(col);
    double cv[SPL];
#pragma unroll
    for (int m = 0; m < SPL; m++) {
        const int c = lane + 64 * m;
        cv[m] = -1.0;
        if (c >= a.count)
            continue;
        const int xr = x0 - (a.dmin + c);
        if (xr < 1 || xr + SW >= a.w - 1)
            continue;
        float st[SW], sj[SW], sjj[SW];
#pragma unroll
        for (int k = 0; k < SW; k++)
            st[k] = sj[k] = sjj[k] = 0.0f;
        for (int r = 0; r < SH; r++) {
            const T *row = R + (size_t)min(max(y0 + r, 0), a.h - 1) * a.pr + xr;
#pragma unroll
            for (int k = 0; k < SW; k++) {
                const float j = Px<T>::get(row[k]), z = t[r * SW + k] - mean;
                st[k] = fmaf(j, z, st[k]);
                sj[k] = sj[k] + j;
                sjj[k] = fmaf(j, j, sjj[k]);
            }
        }
        const double s_t = hsum16(st), s_j = hsum16(sj), s_jj = hsum16(sjj);
        const double var = s_jj - s_j * s_j / (double)(SW * SH);
        cv[m] = (var <= 0.0 || norm_t <= 0.0) ? 0.0 : s_t / sqrt(norm_t * var);
    }
    // the first highest candidate: each lane's own (in candidate order), then across the wave, ties to the lower index
    double bv = -2.0;
    int bi = -1;
#pragma unroll
    for (int m = 0; m < SPL; m++) {
        const int c = lane + 64 * m;
        if (c < a.count && cv[m] > bv) {
            bv = cv[m];
            bi = c;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = shfl_d(bv, o);
        const int oi = __shfl_xor(bi, o);
        if (ov > bv || (ov == bv && oi >= 0 && (bi < 0 || oi < bi))) {
            bv = ov;
            bi = oi;
        }
    }
    double sv = -1.0;
#pragma unroll
    for (int m = 0; m < SPL; m++) {
        const int c = lane + 64 * m;
        if (c < a.count && (c < bi - 2 || c > bi + 2) && cv[m] > sv)
            sv = cv[m];
    }
    for (int o = 32; o > 0; o >>= 1)
        sv = fmax(sv, shfl_d(sv, o));
    if (lane == 0) {
        vtg_stereo_scan r;
        r.disparity = bi < 0 ? -1.0f : (float)(a.dmin + bi);
        r.ncc = bi < 0 ? 0.0f : (float)bv;
        r.second = (float)sv;
        a.out[i] = r;
    }
}

template <typename T> void launch_t(vtg_ctx *c, const ScanArgs &a)
{
    hipLaunchKernelGGL(stereo_scan<T>, dim3((a.n + 3) / 4), dim3(256), 0, c->stream, a);
}

int launch_scan(vtg_ctx *c, const vtg_pyr *left, const vtg_pyr *right, int level, int n, const float *xy, int dmin,
                int dmax, vtg_stereo_scan *out)
{
    if (!left || !right || level < 0 || level >= left->levels || level >= right->levels || n < 0 ||
        (n > 0 && (!xy || !out)))
        return fail(c, VTG_EINVAL, "vtg_stereo_scan_many: bad argument");
    const Level &Ll = left->lv[level], &Rl = right->lv[level];
    if (left->storage != right->storage || Ll.w != Rl.w || Ll.h != Rl.h)
        return fail(c, VTG_EINVAL, "vtg_stereo_scan_many: the two levels differ (storage or size)");
    if (n == 0)
        return VTG_OK;
    const long long span = (long long)dmax - (long long)dmin + 1;
    const int count = span <= 0 ? 0 : (int)std::min(span, (long long)SMAX);
    ScanArgs a{Ll.p, Rl.p, Ll.w, Ll.h, Ll.pitch, Rl.pitch, n, xy, dmin, count, out};
    hipEvent_t ev;
    timer_begin(c, "stereo_scan", &ev);
    const int st = left->storage;
    if (st == VTG_STORE_F32)
        launch_t<float>(c, a);
    else if (level == 0 && st != VTG_STORE_Q6W)
        launch_t<uint8_t>(c, a);
    else if (st == VTG_STORE_Q6 || st == VTG_STORE_Q6W)
        launch_t<uint16_t>(c, a);
    else if (st == VTG_STORE_F16)
        launch_t<__half>(c, a);
    else if (st == VTG_STORE_BF16)
        launch_t<__hip_bfloat16>(c, a);
    else
        return fail(c, VTG_EINVAL, "vtg_stereo_scan_many: storage %d", st);
    VTG_HIP(c, hipGetLastError());
    timer_end(c, "stereo_scan", ev);
    return VTG_OK;
}

}  // namespace
}  // namespace vtg

using namespace vtg;

extern "C" {

int vtg_stereo_scan_many(vtg_ctx *c, const vtg_pyr *left, const vtg_pyr *right, int level, int n, const float *xy,
                         int dmin, int dmax, struct vtg_stereo_scan *out)
{
    if (!c)
        return VTG_EINVAL;
    return launch_scan(c, left, right, level, n, xy, dmin, dmax, out);
}

int vtg_stereo_scan_planes(vtg_ctx *c, const struct vt_plane *left, const struct vt_plane *right, int n,
                           const float *x, const float *y, int dmin, int dmax, struct vtg_stereo_scan *out)
{
    if (!c || !left || !right || n < 0 || (n > 0 && (!x || !y || !out)))
        return VTG_EINVAL;
    if (n == 0)
        return VTG_OK;
    vtg_pyr *pl = nullptr, *pr = nullptr;
    int rc = pyr_from_planes(c, left, 1, &pl);
    if (!rc)
        rc = pyr_from_planes(c, right, 1, &pr);
    void *buf = nullptr;
    const size_t xy_bytes = sizeof(float) * 2 * (size_t)n, out_bytes = sizeof(vtg_stereo_scan) * (size_t)n;
    if (!rc) {
        hipError_t e = hipMalloc(&buf, xy_bytes + out_bytes);
        if (e != hipSuccess)
            rc = hip_fail(c, e, "vtg_stereo_scan_planes: hipMalloc");
    }
    if (!rc) {
        std::vector<float> xy(2 * (size_t)n);
        for (int i = 0; i < n; i++) {
            xy[2 * i] = x[i];
            xy[2 * i + 1] = y[i];
        }
        hipError_t e = hipMemcpy(buf, xy.data(), xy_bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess)
            rc = hip_fail(c, e, "vtg_stereo_scan_planes: upload");
    }
    if (!rc)
        rc = launch_scan(c, pl, pr, 0, n, (const float *)buf, dmin, dmax,
                         (vtg_stereo_scan *)((char *)buf + xy_bytes));
    if (!rc) {
        hipError_t e = hipMemcpyAsync(out, (char *)buf + xy_bytes, out_bytes, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess)
            rc = hip_fail(c, e, "vtg_stereo_scan_planes: download");
    }
    (void)hipStreamSynchronize(c->stream);
    if (buf)
        (void)hipFree(buf);
    if (pl)
        vtg_pyr_destroy(c, pl);
    if (pr)
        vtg_pyr_destroy(c, pr);
    return rc;
}

}  // extern "C"
