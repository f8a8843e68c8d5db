// still.hip: still-cell vectors (src/kern/still.c, still.h:18-53): one camera's running-mean picture of the level it
// reads, and per 16 x 8 cell of that level, LK from the mean to each new frame at the cell's centre.
//
// A frame is the reference's four phases (still.h:11): vectors (per-cell LK with the flowback check, the cell's
// mean |frame - mean| and LK's sigma) on the GPU; decide (the camera's noise median, each cell's evidence, the 3 x 3
// vote, counts, weights, drift) on one host thread, as the reference runs it on one thread, in x87 long double like
// the reference; update (the weighted mean) on the GPU; finish. The mean and the staged frame stay in HBM as 1-level
// F32 pyramids; per frame only the cells (24 B each) come down and their weights (4 B each) go up.
//
// Bit-exactness: a cell's u, v are fp32 differences (a long double difference of two floats rounded to float is the
// fp32 difference); its mean difference sums each row's 16 lanes in the reference's lane order (vx.h:213-221), adds
// the rows as long double and divides as long double, all emulated exactly with 128-bit integers (round to 64 bits,
// then to 24, as x87 does); sigma is sqrtl of a long double sum, emulated the same way. The mean update is one fmaf per
// pixel. LK itself is the path's LK kernel on F32 levels (15 x 15, 6 iterations, one level).
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "ldnorm.hpp"
#include "vtg_internal.hpp"

namespace vtg {
namespace {

// still.c:12-30
constexpr int ST_PATCH = 15, ST_ITER = 6;
constexpr float ST_EPS = 0.01f, ST_MIN_EIG = 0.1f, ST_MAX_RES = 16.0f, ST_FB = 0.25f, ST_MARGIN = 9.0f;
constexpr long double ST_VEC_SIGMA = 12.0L, ST_DIFF_NOISE = 3.0L, ST_NOISE_FRAMES = 32.0L;
constexpr int ST_EVIDENCE = 2, ST_BINS = 256, ST_PER_GREY = 16;

// ---- kernels

template <typename T> struct Q6Of;
template <> struct Q6Of<uint8_t> {
    static __device__ float get(uint8_t v) { return 64.0f * (float)v; }
};
template <> struct Q6Of<uint16_t> {
    static __device__ float get(uint16_t v) { return (float)v; }
};
template <> struct Q6Of<float> {
    static __device__ float get(float v) { return v; }
};
template <> struct Q6Of<__half> {
    static __device__ float get(__half v) { return 64.0f * __half2float(v); }
};
template <> struct Q6Of<__hip_bfloat16> {
    static __device__ float get(__hip_bfloat16 v) { return 64.0f * __bfloat162float(v); }
};

// A pyramid level in Q6 fp32 (the frame the cells read).
template <typename T>
__global__ __launch_bounds__(256) void stage_level(const T *__restrict__ src, int sp, int w, int h,
                                                   float *__restrict__ dst, int dp)
{
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x < w && y < h)
        dst[(size_t)y * dp + x] = Q6Of<T>::get(src[(size_t)y * sp + x]);
}

// The visible pixels of a padded fp32 host plane (uploaded, `stride` floats a row).
__global__ __launch_bounds__(256) void stage_plane(const float *__restrict__ src, int stride, int w, int h,
                                                   float *__restrict__ dst, int dp)
{
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x < w && y < h)
        dst[(size_t)y * dp + x] = src[(size_t)(y + VT_PAD_Y) * stride + VT_PAD_X + x];
}

struct CellArgs {
    const float *cur, *mean; // w x h, pitch p (floats)
    int w, h, p, cols, n;
    const float *xy;         // the cells' centres
    const lk_out *lk;
    vtg_still_cell *cell;
};

// One thread per cell (still.c:99-111, 125-136).
__global__ __launch_bounds__(256) void still_cells(CellArgs a)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n)
        return;
    const int col = i % a.cols, row = i / a.cols;
    const int x0 = col * VTG_CELL_W, y0 = row * VTG_CELL_H, y1 = min(y0 + VTG_CELL_H, a.h);
    XL sum{0, 0};
    for (int y = y0; y < y1; y++) {
        const float *c = a.cur + (size_t)y * a.p, *m = a.mean + (size_t)y * a.p;
        float d[16];
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const int x = min(x0 + k, a.w - 1); // lanes past the edge read the replicated border
            d[k] = fabsf(c[x] - m[x]);
        }
        // the lane-0 result of the i^1, i^2, i^4, i^8 butterfly
        const float s = (((d[0] + d[1]) + (d[2] + d[3])) + ((d[4] + d[5]) + (d[6] + d[7]))) +
                        (((d[8] + d[9]) + (d[10] + d[11])) + ((d[12] + d[13]) + (d[14] + d[15])));
        sum = xl_add(sum, s);
    }
    const float diff = xl_div_to_float(sum, (y1 - y0) * VTG_CELL_W);
    const lk_out o = a.lk[i];
    vtg_still_cell r;
    r.u = o.x - a.xy[2 * i];
    r.v = o.y - a.xy[2 * i + 1];
    r.diff = (float)((double)diff / 64.0);
    r.state = o.status;
    r.sigma = o.status ? 0.0f : sqrt_sum(o.cov_xx, o.cov_yy);
    r.pad = 0.0f;
    a.cell[i] = r;
}

// still.c:239-262: a weight of 1 copies the frame, otherwise mean + (frame - mean) * weight, fused.
__global__ __launch_bounds__(256) void still_update(const float *__restrict__ cur, float *__restrict__ mean, int w,
                                                    int h, int p, int cols, const float *__restrict__ weight)
{
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w || y >= h)
        return;
    const float wt = weight[(y / VTG_CELL_H) * cols + x / VTG_CELL_W];
    const size_t o = (size_t)y * p + x;
    const float c = cur[o], m = mean[o];
    mean[o] = wt == 1.0f ? c : fmaf(c - m, wt, m);
}

}  // namespace
}  // namespace vtg

using namespace vtg;

struct vtg_still {
    vtg_still_cfg cfg{};
    int cols = 0, rows = 0, ncells = 0;
    vtg_pyr *mean = nullptr, *cur = nullptr; // 1-level F32
    void *block = nullptr;                   // device: centres (2 n floats), lk_out n, cells n, weights n
    float *xy = nullptr, *weight = nullptr;
    lk_out *lk = nullptr;
    vtg_still_cell *dcell = nullptr;
    std::vector<vtg_still_cell> cell;
    std::vector<uint16_t> count;
    std::vector<float> wt;
    std::vector<uint8_t> raw;
    bool cells_pending = false; // cells coming down from the last vectors call
    uint64_t frames = 0;
    uint32_t tracked = 0, moving = 0;
    float count_mean = 0, drift_u = 0, drift_v = 0, drift_rms = 0, diff_median = 0, diff_limit = 0;
    long double noise_usual = 0.0L;
};

namespace vtg {
namespace {

int stage_from_pyr(vtg_ctx *c, vtg_still *s, const vtg_pyr *p, int level)
{
    if (!p || level < 0 || level >= p->levels)
        return fail(c, VTG_EINVAL, "vtg_still: no level %d", level);
    const Level &L = p->lv[level];
    if (L.w != s->cfg.w || L.h != s->cfg.h)
        return fail(c, VTG_EINVAL, "vtg_still: level %d is %dx%d, the camera's level is %dx%d", level, L.w, L.h,
                    s->cfg.w, s->cfg.h);
    const Level &D = s->cur->lv[0];
    const dim3 grid((L.w + 255) / 256, L.h), block(256);
    float *dst = (float *)D.p;
    const bool l0 = level == 0;
    switch (p->storage) {
    case VTG_STORE_Q6:
        if (l0)
            hipLaunchKernelGGL(stage_level<uint8_t>, grid, block, 0, c->stream, (const uint8_t *)L.p, L.pitch, L.w,
                               L.h, dst, D.pitch);
        else
            hipLaunchKernelGGL(stage_level<uint16_t>, grid, block, 0, c->stream, (const uint16_t *)L.p, L.pitch,
                               L.w, L.h, dst, D.pitch);
        break;
    case VTG_STORE_Q6W:
        hipLaunchKernelGGL(stage_level<uint16_t>, grid, block, 0, c->stream, (const uint16_t *)L.p, L.pitch, L.w, L.h,
                           dst, D.pitch);
        break;
    case VTG_STORE_Q6F:
        if (l0)
            hipLaunchKernelGGL(stage_level<uint8_t>, grid, block, 0, c->stream, (const uint8_t *)L.p, L.pitch, L.w,
                               L.h, dst, D.pitch);
        else
            hipLaunchKernelGGL(stage_level<float>, grid, block, 0, c->stream, (const float *)L.p, L.pitch, L.w, L.h,
                               dst, D.pitch);
        break;
    case VTG_STORE_F32:
    case VTG_STORE_F32_U8:
        hipLaunchKernelGGL(stage_level<float>, grid, block, 0, c->stream, (const float *)L.p, L.pitch, L.w, L.h, dst,
                           D.pitch);
        break;
    case VTG_STORE_F16:
        if (l0)
            hipLaunchKernelGGL(stage_level<uint8_t>, grid, block, 0, c->stream, (const uint8_t *)L.p, L.pitch, L.w,
                               L.h, dst, D.pitch);
        else
            hipLaunchKernelGGL(stage_level<__half>, grid, block, 0, c->stream, (const __half *)L.p, L.pitch, L.w,
                               L.h, dst, D.pitch);
        break;
    case VTG_STORE_BF16:
        if (l0)
            hipLaunchKernelGGL(stage_level<uint8_t>, grid, block, 0, c->stream, (const uint8_t *)L.p, L.pitch, L.w,
                               L.h, dst, D.pitch);
        else
            hipLaunchKernelGGL(stage_level<__hip_bfloat16>, grid, block, 0, c->stream, (const __hip_bfloat16 *)L.p,
                               L.pitch, L.w, L.h, dst, D.pitch);
        break;
    default:
        return fail(c, VTG_EINVAL, "vtg_still: storage %d", p->storage);
    }
    VTG_HIP(c, hipGetLastError());
    return VTG_OK;
}

int stage_from_plane(vtg_ctx *c, vtg_still *s, const vt_plane *pl)
{
    if (!pl || !pl->p || pl->w != s->cfg.w || pl->h != s->cfg.h || pl->stride < pl->w + 2 * VT_PAD_X)
        return fail(c, VTG_EINVAL, "vtg_still: the plane is not the camera's %dx%d", s->cfg.w, s->cfg.h);
    const size_t bytes = (size_t)pl->stride * (size_t)(pl->h + 2 * VT_PAD_Y) * sizeof(float);
    void *tmp = nullptr;
    int rc = scratch(c, bytes, &tmp);
    if (rc)
        return rc;
    const float *base = pl->p - (size_t)pl->stride * VT_PAD_Y - VT_PAD_X;
    VTG_HIP(c, hipMemcpyAsync(tmp, base, bytes, hipMemcpyHostToDevice, c->stream));
    const Level &D = s->cur->lv[0];
    hipLaunchKernelGGL(stage_plane, dim3((pl->w + 255) / 256, pl->h), dim3(256), 0, c->stream, (const float *)tmp,
                       pl->stride, pl->w, pl->h, (float *)D.p, D.pitch);
    VTG_HIP(c, hipGetLastError());
    return VTG_OK;
}

int run_vectors(vtg_ctx *c, vtg_still *s)
{
    const lk_cfg cfg{ST_PATCH, ST_PATCH, ST_ITER, ST_EPS, ST_MIN_EIG, ST_MAX_RES, 0};
    const int top = 0, off[2] = {0, s->ncells};
    const vtg_pyr *I = s->mean, *J = s->cur;
    int rc = launch_lk_many(c, 1, &I, &J, &top, off, &cfg, s->xy, nullptr, ST_FB, VTG_SOLVE_F64, s->lk, nullptr);
    if (rc)
        return rc;
    const Level &M = s->mean->lv[0], &C = s->cur->lv[0];
    CellArgs a{(const float *)C.p, (const float *)M.p, s->cfg.w, s->cfg.h, C.pitch, s->cols, s->ncells, s->xy, s->lk,
               s->dcell};
    hipEvent_t ev;
    timer_begin(c, "still_cells", &ev);
    hipLaunchKernelGGL(still_cells, dim3((s->ncells + 255) / 256), dim3(256), 0, c->stream, a);
    VTG_HIP(c, hipGetLastError());
    timer_end(c, "still_cells", ev);
    VTG_HIP(c, hipMemcpyAsync(s->cell.data(), s->dcell, sizeof(vtg_still_cell) * s->ncells, hipMemcpyDeviceToHost,
                              c->stream));
    s->cells_pending = true;
    return VTG_OK;
}

int wait_cells(vtg_ctx *c, vtg_still *s)
{
    if (s->cells_pending) {
        VTG_HIP(c, hipStreamSynchronize(c->stream));
        s->cells_pending = false;
    }
    return VTG_OK;
}

// still_decide (still.c:140-237), on this thread in long double.
void decide(vtg_still *s)
{
    const int cols = s->cols, rows = s->rows, wide = cols + 2, n = s->ncells;
    uint32_t hist[ST_BINS];
    std::memset(hist, 0, sizeof hist);
    for (int i = 0; i < n; i++) {
        const long b = (long)((long double)s->cell[i].diff * (long double)ST_PER_GREY);
        hist[b < ST_BINS - 1 ? b : ST_BINS - 1]++;
    }
    // the median bin: the first whose cumulative count reaches half the cells
    int med = 0;
    uint32_t below = 0;
    for (; med < ST_BINS - 1; med++) {
        if (below + hist[med] >= (uint32_t)n / 2u)
            break;
        below += hist[med];
    }
    const long double grey = (long double)s->cfg.motion_grey;
    const long double noise = (long double)med / (long double)ST_PER_GREY;
    const long double limit = std::max(noise * ST_DIFF_NOISE, grey);
    s->diff_median = (float)noise;
    s->diff_limit = (float)limit;
    const long double usual = std::max(s->noise_usual * ST_DIFF_NOISE, grey);
    const bool everything = s->frames >= 2 && noise > usual;
    if (s->frames >= 1 && !everything)
        s->noise_usual = s->frames == 1 ? noise : s->noise_usual + (noise - s->noise_usual) / ST_NOISE_FRAMES;
    const long double px2 = (long double)s->cfg.motion_px * (long double)s->cfg.motion_px;
    for (int i = 0; i < n; i++) {
        const vtg_still_cell &c = s->cell[i];
        const uint32_t lk = c.state & VTG_CELL_LK;
        const long double u = c.u, v = c.v, sg = (long double)c.sigma * ST_VEC_SIGMA;
        const long double need = std::max(sg * sg, px2);
        const bool evidence = (lk == 0 && u * u + v * v > need) || (lk & LK_RESIDUAL) != 0 || (long double)c.diff > limit;
        s->raw[(size_t)(i / cols + 1) * wide + (i % cols + 1)] = evidence ? 1 : 0;
    }
    uint32_t tracked = 0, moving = 0, quiet = 0;
    long double su = 0.0L, sv = 0.0L, suu = 0.0L, counts = 0.0L;
    for (int r = 0; r < rows; r++) {
        const uint8_t *up = &s->raw[(size_t)r * wide], *mid = up + wide, *dn = mid + wide;
        for (int q = 0; q < cols; q++) {
            const int i = r * cols + q;
            const uint32_t votes = up[q] + up[q + 1] + up[q + 2] + mid[q] + mid[q + 1] + mid[q + 2] + dn[q] +
                                   dn[q + 1] + dn[q + 2];
            const bool any = votes >= (uint32_t)ST_EVIDENCE || everything;
            const uint32_t k = any ? 1u : std::min((uint32_t)s->count[i] + 1u, (uint32_t)s->cfg.avg);
            s->count[i] = (uint16_t)k;
            s->wt[i] = (float)(1.0L / (long double)k);
            vtg_still_cell &c = s->cell[i];
            const uint32_t lk = c.state & VTG_CELL_LK;
            c.state = lk | (mid[q + 1] ? VTG_CELL_CHANGED : 0u) | (any ? VTG_CELL_MOVING : 0u) |
                      (k << VTG_CELL_COUNT_SHIFT);
            tracked += lk == 0;
            moving += any;
            counts += (long double)k;
            if (lk == 0 && !any) {
                const long double u = c.u, v = c.v;
                su += u;
                sv += v;
                suu += u * u + v * v;
                quiet++;
            }
        }
    }
    s->tracked = tracked;
    s->moving = moving;
    s->count_mean = (float)(counts / (long double)n);
    const long double mu = quiet ? su / (long double)quiet : 0.0L, mv = quiet ? sv / (long double)quiet : 0.0L;
    s->drift_u = (float)mu;
    s->drift_v = (float)mv;
    const long double var = quiet ? suu / (long double)quiet - mu * mu - mv * mv : 0.0L;
    s->drift_rms = var > 0.0L ? (float)sqrtl(var) : 0.0f;
}

int run_update(vtg_ctx *c, vtg_still *s)
{
    const Level &M = s->mean->lv[0], &C = s->cur->lv[0];
    hipEvent_t ev;
    timer_begin(c, "still_update", &ev);
    hipLaunchKernelGGL(still_update, dim3((s->cfg.w + 255) / 256, s->cfg.h), dim3(256), 0, c->stream,
                       (const float *)C.p, (float *)M.p, s->cfg.w, s->cfg.h, M.pitch, s->cols, s->weight);
    VTG_HIP(c, hipGetLastError());
    timer_end(c, "still_update", ev);
    return VTG_OK;
}

size_t up256(size_t v) { return (v + 255) / 256 * 256; }

uint64_t fnv(uint64_t h, const void *p, size_t n)
{
    const uint8_t *b = (const uint8_t *)p;
    for (size_t i = 0; i < n; i++) {
        h ^= b[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

}  // namespace
}  // namespace vtg

extern "C" {

int vtg_still_create(vtg_ctx *c, const struct vtg_still_cfg *cfg, vtg_still **out)
{
    if (!c || !cfg || !out)
        return VTG_EINVAL;
    *out = nullptr;
    if (cfg->w < 2 * (int)ST_MARGIN + 2 || cfg->h < 2 * (int)ST_MARGIN + 2 || cfg->w > 16384 || cfg->h > 16384 ||
        cfg->avg < 1 || cfg->avg > 65535)
        return fail(c, VTG_EINVAL, "vtg_still_create: %dx%d (20..16384), avg %d (1..65535)", cfg->w, cfg->h,
                    cfg->avg);
    vtg_still *s = new vtg_still();
    s->cfg = *cfg;
    s->cols = (cfg->w + VTG_CELL_W - 1) / VTG_CELL_W;
    s->rows = (cfg->h + VTG_CELL_H - 1) / VTG_CELL_H;
    s->ncells = s->cols * s->rows;
    const int n = s->ncells;
    s->cell.assign(n, vtg_still_cell{});
    s->count.assign(n, 0);
    s->wt.assign(n, 0.0f);
    s->raw.assign((size_t)(s->cols + 2) * (s->rows + 2), 0);
    int rc = vtg_pyr_create(c, cfg->w, cfg->h, 1, VTG_STORE_F32, &s->mean);
    if (!rc)
        rc = vtg_pyr_create(c, cfg->w, cfg->h, 1, VTG_STORE_F32, &s->cur);
    const size_t o_lk = up256(sizeof(float) * 2 * n), o_cell = o_lk + up256(sizeof(lk_out) * n),
                 o_wt = o_cell + up256(sizeof(vtg_still_cell) * n), bytes = o_wt + up256(sizeof(float) * n);
    if (!rc) {
        hipError_t e = hipMalloc(&s->block, bytes);
        if (e != hipSuccess)
            rc = hip_fail(c, e, "vtg_still_create: hipMalloc");
    }
    if (!rc) {
        s->xy = (float *)s->block;
        s->lk = (lk_out *)((char *)s->block + o_lk);
        s->dcell = (vtg_still_cell *)((char *)s->block + o_cell);
        s->weight = (float *)((char *)s->block + o_wt);
        // the cells' centres, moved in to where a patch fits (still.c:117-124)
        std::vector<float> xy(2 * (size_t)n);
        const float xm = (float)cfg->w - 1.0f - ST_MARGIN, ym = (float)cfg->h - 1.0f - ST_MARGIN;
        for (int i = 0; i < n; i++) {
            const float x = (float)((i % s->cols) * VTG_CELL_W + VTG_CELL_W / 2);
            const float y = (float)((i / s->cols) * VTG_CELL_H + VTG_CELL_H / 2);
            xy[2 * i] = x < ST_MARGIN ? ST_MARGIN : x > xm ? xm : x;
            xy[2 * i + 1] = y < ST_MARGIN ? ST_MARGIN : y > ym ? ym : y;
        }
        const Level &M = s->mean->lv[0];
        hipError_t e = hipMemcpy(s->xy, xy.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) // the mean starts at zero (still.c:84-89)
            e = hipMemset((void *)M.p, 0, (size_t)M.pitch * sizeof(float) * M.h);
        if (e != hipSuccess)
            rc = hip_fail(c, e, "vtg_still_create: initial state");
    }
    if (rc) {
        vtg_still_destroy(c, s);
        return rc;
    }
    *out = s;
    return VTG_OK;
}

int vtg_still_destroy(vtg_ctx *c, vtg_still *s)
{
    if (!c || !s)
        return VTG_EINVAL;
    (void)hipStreamSynchronize(c->stream);
    if (s->mean)
        vtg_pyr_destroy(c, s->mean);
    if (s->cur)
        vtg_pyr_destroy(c, s->cur);
    if (s->block)
        (void)hipFree(s->block);
    delete s;
    return VTG_OK;
}

int vtg_still_first(const vtg_still *s) { return s && s->frames == 0; }

int vtg_still_vectors(vtg_ctx *c, vtg_still *s, const vtg_pyr *p, int level)
{
    if (!c || !s)
        return VTG_EINVAL;
    int rc = stage_from_pyr(c, s, p, level);
    return rc ? rc : run_vectors(c, s);
}

int vtg_still_vectors_plane(vtg_ctx *c, vtg_still *s, const struct vt_plane *cur)
{
    if (!c || !s)
        return VTG_EINVAL;
    int rc = stage_from_plane(c, s, cur);
    return rc ? rc : run_vectors(c, s);
}

int vtg_still_decide(vtg_ctx *c, vtg_still *s)
{
    if (!c || !s)
        return VTG_EINVAL;
    int rc = wait_cells(c, s);
    if (rc)
        return rc;
    decide(s);
    VTG_HIP(c, hipMemcpy(s->weight, s->wt.data(), sizeof(float) * s->ncells, hipMemcpyHostToDevice));
    return VTG_OK;
}

int vtg_still_update(vtg_ctx *c, vtg_still *s, const vtg_pyr *p, int level)
{
    if (!c || !s)
        return VTG_EINVAL;
    int rc = stage_from_pyr(c, s, p, level);
    return rc ? rc : run_update(c, s);
}

int vtg_still_update_plane(vtg_ctx *c, vtg_still *s, const struct vt_plane *cur)
{
    if (!c || !s)
        return VTG_EINVAL;
    int rc = stage_from_plane(c, s, cur);
    return rc ? rc : run_update(c, s);
}

int vtg_still_finish(vtg_ctx *c, vtg_still *s)
{
    if (!c || !s)
        return VTG_EINVAL;
    s->frames++;
    return VTG_OK;
}

int vtg_still_frame(vtg_ctx *c, vtg_still *s, const vtg_pyr *p, int level)
{
    if (!c || !s)
        return VTG_EINVAL;
    int rc = stage_from_pyr(c, s, p, level); // staged once: the vectors and the update read the same frame
    if (!rc && s->frames > 0)
        rc = run_vectors(c, s);
    if (!rc)
        rc = vtg_still_decide(c, s);
    if (!rc)
        rc = run_update(c, s);
    if (!rc)
        s->frames++;
    return rc;
}

// Several cameras' frames at once: each camera's level staged, every camera's cells through ONE LK launch (their
// centres gathered into one array), then each camera's cell kernel; one wait for all the cells; the decisions here, one
// camera after another; the weights up and each camera's mean update. Each camera is as after vtg_still_frame.
int vtg_still_frame_many(vtg_ctx *c, int n, vtg_still *const *s, const vtg_pyr *const *p, const int *level)
{
    if (!c || n < 0 || (n > 0 && (!s || !p || !level)))
        return VTG_EINVAL;
    for (int i = 0; i < n; i++) {
        if (!s[i])
            return fail(c, VTG_EINVAL, "vtg_still_frame_many: camera %d is null", i);
        for (int j = 0; j < i; j++)
            if (s[j] == s[i])
                return fail(c, VTG_EINVAL, "vtg_still_frame_many: camera %d given twice", i);
    }
    int rc = VTG_OK;
    for (int i = 0; i < n && !rc; i++)
        rc = stage_from_pyr(c, s[i], p[i], level[i]);
    if (rc)
        return rc;
    std::vector<const vtg_pyr *> I, J;
    std::vector<int> top, off{0}, who;
    for (int i = 0; i < n; i++)
        if (s[i]->frames > 0) {
            I.push_back(s[i]->mean);
            J.push_back(s[i]->cur);
            top.push_back(0);
            off.push_back(off.back() + s[i]->ncells);
            who.push_back(i);
        }
    const int m = (int)who.size(), total = off.back();
    if (m > 0) {
        void *buf = nullptr;
        const size_t o_lk = up256(sizeof(float) * 2 * total);
        rc = scratch(c, o_lk + up256(sizeof(lk_out) * total), &buf);
        if (rc)
            return rc;
        float *xy = (float *)buf;
        lk_out *lk = (lk_out *)((char *)buf + o_lk);
        for (int k = 0; k < m; k++)
            VTG_HIP(c, hipMemcpyAsync(xy + 2 * (size_t)off[k], s[who[k]]->xy, sizeof(float) * 2 * s[who[k]]->ncells,
                                      hipMemcpyDeviceToDevice, c->stream));
        const lk_cfg cfg{ST_PATCH, ST_PATCH, ST_ITER, ST_EPS, ST_MIN_EIG, ST_MAX_RES, 0};
        rc = launch_lk_many(c, m, I.data(), J.data(), top.data(), off.data(), &cfg, xy, nullptr, ST_FB,
                            VTG_SOLVE_F64, lk, nullptr);
        if (rc)
            return rc;
        for (int k = 0; k < m; k++) {
            vtg_still *t = s[who[k]];
            const Level &M = t->mean->lv[0], &C = t->cur->lv[0];
            CellArgs a{(const float *)C.p, (const float *)M.p, t->cfg.w, t->cfg.h, C.pitch, t->cols, t->ncells, t->xy,
                       lk + off[k], t->dcell};
            hipLaunchKernelGGL(still_cells, dim3((t->ncells + 255) / 256), dim3(256), 0, c->stream, a);
            VTG_HIP(c, hipGetLastError());
            VTG_HIP(c, hipMemcpyAsync(t->cell.data(), t->dcell, sizeof(vtg_still_cell) * t->ncells,
                                      hipMemcpyDeviceToHost, c->stream));
        }
    }
    VTG_HIP(c, hipStreamSynchronize(c->stream));
    // the cameras' decisions are independent (each reference camera decides on its own thread): one host thread each
    {
        std::vector<std::thread> th;
        for (int i = 1; i < n; i++)
            th.emplace_back([&, i] { decide(s[i]); });
        if (n > 0)
            decide(s[0]);
        for (auto &t : th)
            t.join();
    }
    for (int i = 0; i < n; i++) {
        s[i]->cells_pending = false;
        VTG_HIP(c, hipMemcpyAsync(s[i]->weight, s[i]->wt.data(), sizeof(float) * s[i]->ncells, hipMemcpyHostToDevice,
                                  c->stream));
    }
    for (int i = 0; i < n && !rc; i++) {
        rc = run_update(c, s[i]);
        if (!rc)
            s[i]->frames++;
    }
    return rc;
}

int vtg_still_cells(vtg_ctx *c, vtg_still *s, struct vtg_still_cell *out)
{
    if (!c || !s || !out)
        return VTG_EINVAL;
    int rc = wait_cells(c, s);
    if (rc)
        return rc;
    std::memcpy(out, s->cell.data(), sizeof(vtg_still_cell) * s->ncells);
    return VTG_OK;
}

int vtg_still_mean(vtg_ctx *c, const vtg_still *s, const struct vt_plane *out)
{
    if (!c || !s || !out)
        return VTG_EINVAL;
    return vtg_pyr_read_level(c, s->mean, 0, out);
}

int vtg_still_motion(const vtg_still *s, struct vtg_motion *m)
{
    if (!s || !m)
        return VTG_EINVAL;
    std::memset(m, 0, sizeof *m);
    m->cols = (uint32_t)s->cols;
    m->rows = (uint32_t)s->rows;
    m->tracked = s->tracked;
    m->moving = s->moving;
    m->count_mean = s->count_mean;
    m->drift_u = s->drift_u;
    m->drift_v = s->drift_v;
    m->drift_rms = s->drift_rms;
    m->diff_median = s->diff_median;
    m->diff_limit = s->diff_limit;
    m->frames = s->frames;
    return VTG_OK;
}

int vtg_still_hash(vtg_ctx *c, vtg_still *s, uint64_t h, uint64_t *out)
{
    if (!c || !s || !out)
        return VTG_EINVAL;
    int rc = wait_cells(c, s);
    if (rc)
        return rc;
    const Level &M = s->mean->lv[0];
    std::vector<float> rows((size_t)M.pitch * M.h);
    VTG_HIP(c, hipMemcpyAsync(rows.data(), M.p, rows.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    VTG_HIP(c, hipStreamSynchronize(c->stream));
    h = fnv(h, s->cell.data(), sizeof(vtg_still_cell) * s->ncells);
    for (int y = 0; y < M.h; y++)
        h = fnv(h, &rows[(size_t)y * M.pitch], sizeof(float) * M.w);
    *out = h;
    return VTG_OK;
}

int vtg_still_size(const vtg_still *s, int *cols, int *rows)
{
    if (!s || !cols || !rows)
        return VTG_EINVAL;
    *cols = s->cols;
    *rows = s->rows;
    return VTG_OK;
}

}  // extern "C"
