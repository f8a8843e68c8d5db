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
#include <cmath>
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

size_t up(size_t v) { return (v + 255) / 256 * 256; }

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
    if (st == VTG_STORE_F32 || st == VTG_STORE_F32_U8 || (st == VTG_STORE_Q6F && level > 0))
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

// stereo.c:124-141, in long double on this thread as the reference computes them on one.
int triangulate(const vtg_stereo_geom &g, long double xl, long double yl, long double xr, long double yr,
                long double *p)
{
    const long double d = xl - xr;
    if (d <= 0.0L)
        return 0;
    const long double z = g.fx * g.baseline / d;
    p[0] = (xl - g.cx) * z / g.fx;
    p[1] = (0.5L * (yl + yr) - g.cy) * z / g.fy;
    p[2] = z;
    return 1;
}

long double reprojection(const vtg_stereo_geom &g, const long double *p, long double xl, long double yl,
                         long double xr, long double yr)
{
    const long double ul = g.fx * p[0] / p[2] + g.cx, vl = g.fy * p[1] / p[2] + g.cy;
    const long double ur = g.fx * (p[0] - g.baseline) / p[2] + g.cx;
    const long double el = ul - xl, fl = vl - yl, er = ur - xr, fr = vl - yr;
    return sqrtl((el * el + fl * fl + er * er + fr * fr) * 0.5L);
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

// The tracker's stereo step (stereo_one, track.c:1859-1914) for n features: the scan on the GPU for features without a
// prior, LK on the GPU for the rest (the disparity as x guess; from the stereo level after a scan, level 0 with a
// prior; the flowback from level 0), then each feature's decision and 3D point here, in long double.
int vtg_stereo_match(vtg_ctx *c, const vtg_pyr *left, const vtg_pyr *right, const struct vtg_stereo_cfg *cfg, int n,
                     struct vtg_stereo_feature *f)
{
    if (!c || !left || !right || !cfg || n < 0 || (n > 0 && !f))
        return VTG_EINVAL;
    const int sl = cfg->stereo_level;
    const lk_cfg &lk = cfg->lk;
    if (sl < 0 || sl >= left->levels || sl >= right->levels || sl > 30 || lk.w < 5 || lk.w > LK_MAX_W || lk.h < 5 ||
        lk.h > LK_MAX_H || lk.max_iter < 0)
        return fail(c, VTG_EINVAL, "vtg_stereo_match: stereo level %d of %d / %d levels, window %dx%d", sl,
                    left->levels, right->levels, lk.w, lk.h);
    if (left->storage != right->storage)
        return fail(c, VTG_EINVAL, "vtg_stereo_match: pyramids of different storage");
    for (int l = 0; l <= sl; l++)
        if (left->lv[l].w != right->lv[l].w || left->lv[l].h != right->lv[l].h)
            return fail(c, VTG_EINVAL, "vtg_stereo_match: level %d sizes differ", l);
    if (n == 0)
        return VTG_OK;
    // device scratch: scan points, scans, LK points, guesses, LK records
    const size_t o_sc = up(sizeof(float) * 2 * n), o_lxy = o_sc + up(sizeof(vtg_stereo_scan) * n),
                 o_g = o_lxy + up(sizeof(float) * 2 * n), o_out = o_g + up(sizeof(float) * 2 * n),
                 bytes = o_out + up(sizeof(lk_out) * n);
    void *buf = nullptr;
    int rc = scratch(c, bytes, &buf);
    if (rc)
        return rc;
    char *d = (char *)buf;
    const long double s = 1.0L / (long double)(1 << sl);
    std::vector<int> scanned;
    std::vector<float> sxy;
    for (int i = 0; i < n; i++) {
        f[i].xr = f[i].yr = 0.0f;
        f[i].z = 0.0f;
        if (f[i].disparity <= 0.0f) {
            scanned.push_back(i);
            sxy.push_back((float)((long double)f[i].x * s));
            sxy.push_back((float)((long double)f[i].y * s));
        }
    }
    std::vector<vtg_stereo_scan> sc(scanned.size());
    if (!scanned.empty()) {
        const int dmin = (int)((long double)cfg->min_disp * s), dmax = (int)((long double)cfg->max_disp * s);
        VTG_HIP(c, hipMemcpyAsync(d, sxy.data(), sizeof(float) * sxy.size(), hipMemcpyHostToDevice, c->stream));
        rc = launch_scan(c, left, right, sl, (int)scanned.size(), (const float *)d, dmin, dmax,
                         (vtg_stereo_scan *)(d + o_sc));
        if (rc)
            return rc;
        VTG_HIP(c, hipMemcpyAsync(sc.data(), d + o_sc, sizeof(vtg_stereo_scan) * sc.size(), hipMemcpyDeviceToHost,
                                  c->stream));
        VTG_HIP(c, hipStreamSynchronize(c->stream));
    }
    // the points LK takes: those with a prior (from level 0), then the scanned ones that passed (from the stereo level)
    std::vector<int> idx0, idx1;
    std::vector<float> guess(n, 0.0f);
    for (int i = 0; i < n; i++)
        if (f[i].disparity > 0.0f) {
            guess[i] = f[i].disparity;
            idx0.push_back(i);
        }
    for (size_t k = 0; k < scanned.size(); k++) {
        vtg_stereo_feature &q = f[scanned[k]];
        const uint32_t fl = q.flags & VTG_FF_NEW;
        if (sc[k].disparity < 0.0f || sc[k].ncc < cfg->min_ncc) {
            q.flags = fl;
            q.disparity = -1.0f;
        } else if ((long double)sc[k].second > (long double)sc[k].ncc * (long double)cfg->ambiguity) {
            q.flags = fl | VTG_FF_AMBIGUOUS;
            q.disparity = -1.0f;
        } else {
            guess[scanned[k]] = (float)((long double)sc[k].disparity * (long double)(1 << sl));
            idx1.push_back(scanned[k]);
        }
    }
    const int m0 = (int)idx0.size(), m1 = (int)idx1.size(), m = m0 + m1;
    if (m > 0) {
        std::vector<float> lxy(2 * (size_t)m), lg(2 * (size_t)m, 0.0f);
        for (int k = 0; k < m; k++) {
            const int i = k < m0 ? idx0[k] : idx1[k - m0];
            lxy[2 * k] = f[i].x;
            lxy[2 * k + 1] = f[i].y;
            lg[2 * k] = (float)(-(long double)guess[i]);
        }
        VTG_HIP(c, hipMemcpyAsync(d + o_lxy, lxy.data(), sizeof(float) * lxy.size(), hipMemcpyHostToDevice,
                                  c->stream));
        VTG_HIP(c, hipMemcpyAsync(d + o_g, lg.data(), sizeof(float) * lg.size(), hipMemcpyHostToDevice, c->stream));
        const vtg_pyr *I[2] = {left, left}, *J[2] = {right, right};
        const int top[2] = {0, sl}, off[3] = {0, m0, m};
        rc = launch_lk_many(c, 2, I, J, top, off, &lk, (const float *)(d + o_lxy), (const float *)(d + o_g),
                            cfg->fb_max, VTG_SOLVE_F64, (lk_out *)(d + o_out), nullptr, 0);
        if (rc)
            return rc;
        std::vector<lk_out> out(m);
        VTG_HIP(c, hipMemcpyAsync(out.data(), d + o_out, sizeof(lk_out) * m, hipMemcpyDeviceToHost, c->stream));
        VTG_HIP(c, hipStreamSynchronize(c->stream));
        for (int k = 0; k < m; k++) {
            vtg_stereo_feature &q = f[k < m0 ? idx0[k] : idx1[k - m0]];
            uint32_t fl = q.flags & VTG_FF_NEW;
            const lk_out &o = out[k];
            if (o.status) {
                q.flags = fl;
                q.disparity = -1.0f;
                continue;
            }
            q.xr = o.x;
            q.yr = o.y;
            const long double disparity = (long double)q.x - (long double)o.x;
            q.disparity = (float)disparity;
            fl |= VTG_FF_STEREO;
            if (disparity >= (long double)cfg->min_disp && disparity <= (long double)cfg->max_disp &&
                triangulate(cfg->geom, q.x, q.y, o.x, o.y, q.p) &&
                reprojection(cfg->geom, q.p, q.x, q.y, o.x, o.y) <= (long double)cfg->max_reproj) {
                fl |= VTG_FF_3D;
                q.z = (float)q.p[2];
            }
            q.flags = fl;
        }
    }
    return VTG_OK;
}

}  // extern "C"
