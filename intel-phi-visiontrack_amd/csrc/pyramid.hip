// pyramid.hip: the Gaussian pyramid on the GPU. Level 0 is the u8 frame (Q6 = 64 x byte); level l is the (1 4 6 4 1)^2
// blur of level l-1 sampled at even pixels, (sum + 128) >> 8 in exact integers (img_reduce_ref, src/kern/img_ref.c:98-114),
// with the border replicated (the reference's padding, src/kern/img_ref.c:9-30, is a clamp of the coordinates).
//
// One launch per level. A workgroup makes a 64 x TH tile of the output (TH = 8, 16 or 32 rows, picked per size class by
// the autotuner; 16 by default): it stages the 131 x (2 TH + 3) input tile (with the 2-pixel halo, coordinates clamped)
// in LDS as integers, sums the 5 rows at each even row, then the 5 columns at each even column. The level-0 launch also
// writes the frame into the pyramid's level 0 from the same tile, so the frame is read once.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>

#include "reduce_tile.hpp"
#include "vtg_internal.hpp"

namespace vtg {

namespace {

using namespace rt;

// Reduce for a batch of jobs, one 64 x TH output tile per workgroup (reduce_tile.hpp).
template <typename TS, typename TD, typename TC, bool COPY, bool EXACT, int TH>
__global__ __launch_bounds__(256) void reduce_batch(RJobs jobs)
{
    int jn = 0;
#pragma unroll
    for (int k = 1; k < JMAX; k++)
        jn = (k < jobs.n && (int)blockIdx.x >= jobs.t0s[k]) ? k : jn;
    const RJob &J = jobs.j[jn];
    __shared__ typename AccT<EXACT>::type lds[tile_lds_elems<TH>()];
    reduce_tile<TS, TD, TC, COPY, EXACT, TH>(J, (int)blockIdx.x - J.t0, lds);
}

// Level 0 of an F32 pyramid from the u8 frame: 64 x byte (img_u8_to_q6, src/kern/img_ref.c:32-40).
__global__ __launch_bounds__(256) void u8_to_f32(const uint8_t *__restrict__ src, int sp, float *__restrict__ dst,
                                                 int dp, int w, int h)
{
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x < w && y < h)
        dst[(size_t)y * dp + x] = (float)src[(size_t)y * sp + x] * 64.0f;
}

// A level as the reference's padded fp32 plane: every lane of rows -4 .. h+3 and columns -32 .. stride-33, the
// border replicated (img_pad_rows + img_pad_tb).
template <typename T>
__global__ __launch_bounds__(256) void to_plane(const T *__restrict__ src, int sp, int w, int h, float *__restrict__ dst,
                                                int stride)
{
    const int x = blockIdx.x * 256 + threadIdx.x, r = blockIdx.y; // r = padded row 0 .. h + 7
    if (x >= stride)
        return;
    const int sx = clampi(x - VT_PAD_X, 0, w - 1), sy = clampi(r - VT_PAD_Y, 0, h - 1);
    dst[(size_t)r * stride + x] = Q6<T>::get(src[(size_t)sy * sp + sx]);
}

// The visible pixels of a padded fp32 host plane (already uploaded, `stride` floats per row) into a device F32 level.
__global__ __launch_bounds__(256) void from_plane(const float *__restrict__ src, int stride, int w, int h,
                                                  float *__restrict__ dst, int dp)
{
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x < w && y < h)
        dst[(size_t)y * dp + x] = src[(size_t)(y + VT_PAD_Y) * stride + VT_PAD_X + x];
}

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

template <typename TS, typename TD, typename TC, bool COPY, bool EXACT, int TH>
int launch_jobs_th(vtg_ctx *c, const std::vector<RJob> &all)
{
    for (size_t b = 0; b < all.size(); b += JMAX) {
        RJobs js{};
        int tiles = 0;
        js.n = (int)std::min<size_t>(JMAX, all.size() - b);
        for (int k = 0; k < js.n; k++) {
            RJob j = all[b + k];
            j.tiles_x = (j.dw + TW - 1) / TW;
            j.t0 = tiles;
            tiles += j.tiles_x * ((j.dh + TH - 1) / TH);
            js.j[k] = j;
            js.t0s[k] = j.t0;
        }
        hipEvent_t ev;
        timer_begin(c, "pyr_reduce", &ev);
        hipLaunchKernelGGL((reduce_batch<TS, TD, TC, COPY, EXACT, TH>), dim3(tiles), dim3(256), 0, c->stream, js);
        VTG_HIP(c, hipGetLastError());
        timer_end(c, "pyr_reduce", ev);
    }
    return VTG_OK;
}

// jobs[i] with tile rows rows[i]: one launch (per JMAX jobs) for each tile height present.
template <typename TS, typename TD, typename TC, bool COPY, bool EXACT>
int launch_jobs(vtg_ctx *c, const std::vector<RJob> &jobs, const std::vector<int> &rows)
{
    for (int th : {8, 16, 32}) {
        std::vector<RJob> sel;
        for (size_t i = 0; i < jobs.size(); i++)
            if (rows[i] == th)
                sel.push_back(jobs[i]);
        if (sel.empty())
            continue;
        int rc = th == 8    ? launch_jobs_th<TS, TD, TC, COPY, EXACT, 8>(c, sel)
                 : th == 16 ? launch_jobs_th<TS, TD, TC, COPY, EXACT, 16>(c, sel)
                            : launch_jobs_th<TS, TD, TC, COPY, EXACT, 32>(c, sel);
        if (rc)
            return rc;
    }
    return VTG_OK;
}

// Level 1 of every pyramid from its frame (writing level 0 on the way), then each higher level from the one below,
// one launch per level for all pyramids together.
template <typename T1, typename TC, bool EXACT>
int build_levels(vtg_ctx *c, vtg_pyr *const *ps, int n, const uint8_t *const *frames, const int *pitch,
                 const int *rows)
{
    int maxl = 0;
    for (int i = 0; i < n; i++)
        maxl = std::max(maxl, ps[i]->levels);
    std::vector<RJob> jobs;
    std::vector<int> th;
    for (int i = 0; i < n; i++) {
        const Level *L = ps[i]->lv;
        if (ps[i]->levels < 2)
            continue;
        // F32_U8: the frame's bytes kept beside level 0 from the same tile (unless the frame already lives there)
        const bool keep8 = ps[i]->storage == VTG_STORE_F32_U8 && frames[i] != ps[i]->byte0.p;
        jobs.push_back(RJob{frames[i], (void *)L[1].p, (void *)L[0].p, ps[i]->w, ps[i]->h, pitch[i], L[1].w, L[1].h,
                            L[1].pitch, L[0].pitch, 0, 0, keep8 ? (void *)ps[i]->byte0.p : nullptr,
                            keep8 ? ps[i]->byte0.pitch : 0});
        th.push_back(rows[i]);
    }
    int rc = jobs.empty() ? VTG_OK : launch_jobs<uint8_t, T1, TC, true, EXACT>(c, jobs, th);
    for (int l = 2; l < maxl && !rc; l++) {
        jobs.clear();
        th.clear();
        for (int i = 0; i < n; i++) {
            if (ps[i]->levels <= l)
                continue;
            const Level *L = ps[i]->lv;
            jobs.push_back(RJob{L[l - 1].p, (void *)L[l].p, nullptr, L[l - 1].w, L[l - 1].h, L[l - 1].pitch, L[l].w,
                                L[l].h, L[l].pitch, 0, 0, 0});
            th.push_back(rows[i]);
        }
        rc = launch_jobs<T1, T1, T1, false, EXACT>(c, jobs, th);
    }
    return rc;
}

}  // namespace

PyrView view(const vtg_pyr *p)
{
    PyrView v{};
    v.levels = p->levels;
    v.storage = p->storage;
    for (int l = 0; l < p->levels; l++)
        v.l[l] = p->lv[l];
    return v;
}

static int elem_bytes(int storage, int level)
{
    switch (storage) {
    case VTG_STORE_Q6: return level == 0 ? 1 : 2;
    case VTG_STORE_F32: return 4;
    case VTG_STORE_F16:
    case VTG_STORE_BF16: return level == 0 ? 1 : 2;
    case VTG_STORE_Q6W: return 2;
    case VTG_STORE_Q6F: return level == 0 ? 1 : 4;
    case VTG_STORE_F32_U8: return 4;
    }
    return 0;
}

// The level layout of a pyramid (pitches rounded to 256 B, levels to 4 KiB, then the device descriptor).
static size_t pyr_layout(vtg_pyr *p)
{
    size_t off = 0;
    int lw = p->w, lh = p->h;
    for (int l = 0; l < p->levels; l++) {
        const int eb = elem_bytes(p->storage, l);
        const size_t pitch_b = round_up((size_t)lw * eb, 256);
        p->lv[l].w = lw;
        p->lv[l].h = lh;
        p->lv[l].pitch = (int)(pitch_b / eb);
        p->offset[l] = off;
        p->elem[l] = eb;
        off += round_up(pitch_b * lh, 4096);
        lw = (lw + 1) / 2;
        lh = (lh + 1) / 2;
    }
    if (p->storage == VTG_STORE_F32_U8) { // the u8 frame beside level 0
        const size_t pitch_b = round_up((size_t)p->w, 256);
        p->byte0 = Level{nullptr, p->w, p->h, (int)pitch_b};
        p->byte0_off = off;
        off += round_up(pitch_b * p->h, 4096);
    }
    return off;
}

size_t pyr_bytes(int w, int h, int levels, int storage)
{
    vtg_pyr p;
    p.w = w;
    p.h = h;
    p.levels = levels;
    p.storage = storage;
    return pyr_layout(&p) + round_up(sizeof(PyrView), 256);
}

int pyr_new(vtg_ctx *c, int w, int h, int levels, int storage, void *mem, vtg_pyr **out)
{
    *out = nullptr;
    if (w < 8 || h < 8 || w > 16384 || h > 16384)
        return fail(c, VTG_ERANGE, "vtg_pyr_create: frame %dx%d outside 8..16384", w, h);
    if (levels < 1 || levels > VTG_MAX_LEVELS)
        return fail(c, VTG_EINVAL, "vtg_pyr_create: %d levels (1..%d)", levels, VTG_MAX_LEVELS);
    if (storage < VTG_STORE_Q6 || storage > VTG_STORE_F32_U8)
        return fail(c, VTG_EINVAL, "vtg_pyr_create: storage %d", storage);
    vtg_pyr *p = new vtg_pyr();
    p->w = w;
    p->h = h;
    p->levels = levels;
    p->storage = storage;
    const size_t view_off = pyr_layout(p);
    p->bytes = view_off + round_up(sizeof(PyrView), 256);
    p->owned = mem == nullptr;
    if (mem) {
        p->block = mem;
    } else {
        hipError_t e = hipMalloc(&p->block, p->bytes);
        if (e != hipSuccess) {
            delete p;
            return hip_fail(c, e, "vtg_pyr_create: hipMalloc");
        }
    }
    for (int l = 0; l < levels; l++)
        p->lv[l].p = (char *)p->block + p->offset[l];
    if (storage == VTG_STORE_F32_U8)
        p->byte0.p = (char *)p->block + p->byte0_off;
    // the descriptor batched kernels read, written once
    p->dview = (PyrView *)((char *)p->block + view_off);
    const PyrView v = view(p);
    hipError_t e = hipMemcpy(p->dview, &v, sizeof v, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (p->owned)
            (void)hipFree(p->block);
        delete p;
        return hip_fail(c, e, "vtg_pyr_create: descriptor");
    }
    *out = p;
    return VTG_OK;
}

// Levels 1 .. of pyramids whose level 0 is written: each from the one below, one launch per level for all.
template <typename T>
static int levels_from_l0(vtg_ctx *c, vtg_pyr *const *ps, int n, const int *rows)
{
    int maxl = 0;
    for (int i = 0; i < n; i++)
        maxl = std::max(maxl, ps[i]->levels);
    std::vector<RJob> jobs;
    std::vector<int> th;
    int rc = VTG_OK;
    for (int l = 1; l < maxl && !rc; l++) {
        jobs.clear();
        th.clear();
        for (int i = 0; i < n; i++) {
            if (ps[i]->levels <= l)
                continue;
            const Level *L = ps[i]->lv;
            jobs.push_back(RJob{L[l - 1].p, (void *)L[l].p, nullptr, L[l - 1].w, L[l - 1].h, L[l - 1].pitch, L[l].w,
                                L[l].h, L[l].pitch, 0, 0, 0});
            th.push_back(rows[i]);
        }
        rc = launch_jobs<T, T, T, false, true>(c, jobs, th);
    }
    return rc;
}

// frames[i]: pyramid i's u8 frame on the device, pitch[i] bytes per row. All pyramids of one storage. rows: each
// pyramid's reduce tile height (8, 16 or 32), or null for the context's choice for its size class. flags: the level-0
// filters (storage Q6W or F32 only; Q6W always builds level 0 through the filter kernel, 64 x byte without flags).
int launch_pyr_build_many(vtg_ctx *c, vtg_pyr *const *ps, int n, const uint8_t *const *frames, const int *pitch,
                          const int *rows_in, int flags)
{
    std::vector<int> rows(n);
    for (int i = 0; i < n; i++)
        rows[i] = rows_in ? rows_in[i] : tuned_pyr_rows(c, ps[i]);
    if (flags || ps[0]->storage == VTG_STORE_Q6W) {
        int rc = launch_level0_many(c, ps, n, frames, pitch, flags);
        if (rc)
            return rc;
        return ps[0]->storage == VTG_STORE_F32 ? levels_from_l0<float>(c, ps, n, rows.data())
                                               : levels_from_l0<uint16_t>(c, ps, n, rows.data());
    }
    // single-level pyramids: level 0 is the frame (converted for F32)
    for (int i = 0; i < n; i++) {
        if (ps[i]->levels != 1)
            continue;
        const Level &L0 = ps[i]->lv[0];
        if (ps[i]->storage == VTG_STORE_F32 || ps[i]->storage == VTG_STORE_F32_U8) {
            hipLaunchKernelGGL(u8_to_f32, dim3((ps[i]->w + 255) / 256, ps[i]->h), dim3(256), 0, c->stream, frames[i],
                               pitch[i], (float *)L0.p, L0.pitch, ps[i]->w, ps[i]->h);
            VTG_HIP(c, hipGetLastError());
        } else {
            VTG_HIP(c, hipMemcpy2DAsync((void *)L0.p, L0.pitch, frames[i], pitch[i], ps[i]->w, ps[i]->h,
                                        hipMemcpyDeviceToDevice, c->stream));
        }
    }
    switch (ps[0]->storage) {
    case VTG_STORE_Q6: return build_levels<uint16_t, uint8_t, true>(c, ps, n, frames, pitch, rows.data());
    case VTG_STORE_F32: return build_levels<float, float, true>(c, ps, n, frames, pitch, rows.data());
    case VTG_STORE_Q6F: return build_levels<float, uint8_t, true>(c, ps, n, frames, pitch, rows.data());
    case VTG_STORE_F32_U8:
        // the frame kept beside level 0 (unless it was uploaded there): by the level-1 reduce, which reads it anyway,
        // or by a copy for a single-level pyramid; the fp32 levels as VTG_STORE_F32
        for (int i = 0; i < n; i++)
            if (ps[i]->levels == 1 && frames[i] != ps[i]->byte0.p)
                VTG_HIP(c, hipMemcpy2DAsync((void *)ps[i]->byte0.p, ps[i]->byte0.pitch, frames[i], pitch[i], ps[i]->w,
                                            ps[i]->h, hipMemcpyDeviceToDevice, c->stream));
        return build_levels<float, float, true>(c, ps, n, frames, pitch, rows.data());
    case VTG_STORE_F16: return build_levels<__half, uint8_t, false>(c, ps, n, frames, pitch, rows.data());
    case VTG_STORE_BF16: return build_levels<__hip_bfloat16, uint8_t, false>(c, ps, n, frames, pitch, rows.data());
    }
    return fail(c, VTG_EINVAL, "pyramid storage %d", ps[0]->storage);
}

int launch_pyr_build(vtg_ctx *c, vtg_pyr *p, const uint8_t *frame, int pitch)
{
    return launch_pyr_build_many(c, &p, 1, &frame, &pitch, nullptr);
}

int pyr_from_planes(vtg_ctx *c, const vt_plane *levels, int n, vtg_pyr **out)
{
    *out = nullptr;
    for (int l = 1; l < n; l++)
        if (levels[l].w != (levels[l - 1].w + 1) / 2 || levels[l].h != (levels[l - 1].h + 1) / 2)
            return fail(c, VTG_EINVAL, "level %d is %dx%d, not half of %dx%d", l, levels[l].w, levels[l].h,
                        levels[l - 1].w, levels[l - 1].h);
    vtg_pyr *p = nullptr;
    int rc = vtg_pyr_create(c, levels[0].w, levels[0].h, n, VTG_STORE_F32, &p);
    if (rc)
        return rc;
    for (int l = 0; l < n && !rc; l++) {
        const vt_plane &pl = levels[l];
        if (pl.stride != plane_stride(pl.w)) {
            rc = fail(c, VTG_EINVAL, "level %d stride %d, expected %d", l, pl.stride, plane_stride(pl.w));
            break;
        }
        void *tmp = nullptr;
        const size_t bytes = plane_bytes(pl.w, pl.h);
        rc = scratch(c, bytes, &tmp);
        if (rc)
            break;
        const float *base = pl.p - (size_t)pl.stride * VT_PAD_Y - VT_PAD_X;
        hipError_t e = hipMemcpyAsync(tmp, base, bytes, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(from_plane, dim3((pl.w + 255) / 256, pl.h), dim3(256), 0, c->stream,
                               (const float *)tmp, pl.stride, pl.w, pl.h, (float *)p->lv[l].p, p->lv[l].pitch);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipStreamSynchronize(c->stream); // the scratch is reused by the next level
        if (e != hipSuccess)
            rc = hip_fail(c, e, "pyr_from_planes");
    }
    if (rc) {
        vtg_pyr_destroy(c, p);
        return rc;
    }
    *out = p;
    return VTG_OK;
}

}  // namespace vtg

using namespace vtg;

extern "C" {

int vtg_pyr_create(vtg_ctx *c, int w, int h, int levels, int storage, vtg_pyr **out)
{
    if (!c || !out)
        return VTG_EINVAL;
    return pyr_new(c, w, h, levels, storage, nullptr, out);
}

int vtg_pyr_wrap(vtg_ctx *c, int w, int h, const uint8_t *frame, int pitch, vtg_pyr **out)
{
    if (!c || !out)
        return VTG_EINVAL;
    *out = nullptr;
    if (!frame || pitch < w || (pitch & 3) || ((uintptr_t)frame & 3))
        return fail(c, VTG_EINVAL, "vtg_pyr_wrap: a device frame with a word-aligned start and pitch >= width, a "
                                   "multiple of 4");
    if (w < 8 || h < 8 || w > 16384 || h > 16384)
        return fail(c, VTG_ERANGE, "vtg_pyr_wrap: frame %dx%d outside 8..16384", w, h);
    vtg_pyr *p = new vtg_pyr();
    p->w = w;
    p->h = h;
    p->levels = 1;
    p->storage = VTG_STORE_Q6;
    p->wrapped = true;
    p->lv[0] = Level{frame, w, h, pitch};
    p->offset[0] = 0;
    p->elem[0] = 1;
    p->bytes = round_up(sizeof(PyrView), 256); // the block holds the descriptor only
    hipError_t e = hipMalloc(&p->block, p->bytes);
    if (e != hipSuccess) {
        delete p;
        return hip_fail(c, e, "vtg_pyr_wrap: hipMalloc");
    }
    p->dview = (PyrView *)p->block;
    const PyrView v = view(p);
    e = hipMemcpy(p->dview, &v, sizeof v, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(p->block);
        delete p;
        return hip_fail(c, e, "vtg_pyr_wrap: descriptor");
    }
    *out = p;
    return VTG_OK;
}

int vtg_pyr_destroy(vtg_ctx *c, vtg_pyr *p)
{
    if (!c || !p)
        return VTG_EINVAL;
    (void)hipStreamSynchronize(c->stream);
    if (p->owned)
        (void)hipFree(p->block);
    delete p;
    return VTG_OK;
}

static int check_flags(vtg_ctx *c, const vtg_pyr *p, int flags)
{
    if (p->wrapped)
        return fail(c, VTG_EINVAL, "vtg_pyr_build: a wrapped frame is read, never built");
    if (flags & ~(VTG_BUILD_CONTRAST | VTG_BUILD_PRE_GAUSS))
        return fail(c, VTG_EINVAL, "vtg_pyr_build: unknown flags 0x%x", flags);
    if (flags && p->storage != VTG_STORE_Q6W && p->storage != VTG_STORE_F32)
        return fail(c, VTG_EINVAL, "vtg_pyr_build: the level-0 filters need storage VTG_STORE_Q6W or VTG_STORE_F32 "
                                   "(storage %d keeps level 0 as the u8 frame)", p->storage);
    return VTG_OK;
}

int vtg_pyr_build(vtg_ctx *c, vtg_pyr *p, const uint8_t *frame, int pitch, int on_device)
{
    return vtg_pyr_build_ex(c, p, frame, pitch, on_device, 0);
}

int vtg_pyr_build_ex(vtg_ctx *c, vtg_pyr *p, const uint8_t *frame, int pitch, int on_device, int flags)
{
    if (!c || !p || !frame || pitch < p->w)
        return fail(c, VTG_EINVAL, "vtg_pyr_build: bad argument");
    int rc = check_flags(c, p, flags);
    if (rc)
        return rc;
    if (!on_device) {
        // the host frame goes straight into level 0's bytes when that level is u8, else into scratch
        void *dst = nullptr;
        size_t dp = (size_t)p->w;
        if (p->storage == VTG_STORE_F32 || p->storage == VTG_STORE_Q6W) {
            int rc = scratch(c, (size_t)p->w * p->h, &dst);
            if (rc)
                return rc;
        } else {
            const Level &L0 = p->storage == VTG_STORE_F32_U8 ? p->byte0 : p->lv[0]; // where the u8 frame lives
            dst = (void *)L0.p;
            dp = (size_t)L0.pitch;
        }
        VTG_HIP(c, hipMemcpy2DAsync(dst, dp, frame, pitch, p->w, p->h, hipMemcpyHostToDevice, c->stream));
        frame = (const uint8_t *)dst;
        pitch = (int)dp;
    }
    return launch_pyr_build_many(c, &p, 1, &frame, &pitch, nullptr, flags);
}

int vtg_pyr_build_many(vtg_ctx *c, int n, vtg_pyr *const *pyrs, const uint8_t *const *frames, const int *pitches,
                       int on_device)
{
    return vtg_pyr_build_many_ex(c, n, pyrs, frames, pitches, on_device, 0);
}

int vtg_pyr_build_many_ex(vtg_ctx *c, int n, vtg_pyr *const *pyrs, const uint8_t *const *frames, const int *pitches,
                          int on_device, int flags)
{
    if (!c || n < 0 || (n > 0 && (!pyrs || !frames || !pitches)))
        return fail(c, VTG_EINVAL, "vtg_pyr_build_many: bad argument");
    if (n == 0)
        return VTG_OK;
    for (int i = 0; i < n; i++) {
        if (!pyrs[i] || !frames[i] || pitches[i] < pyrs[i]->w || pyrs[i]->storage != pyrs[0]->storage)
            return fail(c, VTG_EINVAL, "vtg_pyr_build_many: pyramid %d: bad frame, pitch or storage", i);
        int rc = check_flags(c, pyrs[i], flags);
        if (rc)
            return rc;
    }
    if (on_device)
        return launch_pyr_build_many(c, pyrs, n, frames, pitches, nullptr, flags);
    // host frames: into level 0 where it is u8, else into scratch, then the same batch
    std::vector<const uint8_t *> dev(n);
    std::vector<int> dp(n);
    size_t need = 0;
    const bool wide = pyrs[0]->storage == VTG_STORE_F32 || pyrs[0]->storage == VTG_STORE_Q6W;
    for (int i = 0; i < n; i++)
        if (wide)
            need += round_up((size_t)pyrs[i]->w * pyrs[i]->h, 256);
    void *tmp = nullptr;
    if (need) {
        int rc = scratch(c, need, &tmp);
        if (rc)
            return rc;
    }
    size_t off = 0;
    for (int i = 0; i < n; i++) {
        vtg_pyr *p = pyrs[i];
        if (wide) {
            dev[i] = (const uint8_t *)tmp + off;
            dp[i] = p->w;
            off += round_up((size_t)p->w * p->h, 256);
        } else {
            const Level &L0 = p->storage == VTG_STORE_F32_U8 ? p->byte0 : p->lv[0]; // where the u8 frame lives
            dev[i] = (const uint8_t *)L0.p;
            dp[i] = L0.pitch;
        }
        VTG_HIP(c, hipMemcpy2DAsync((void *)dev[i], dp[i], frames[i], pitches[i], p->w, p->h, hipMemcpyHostToDevice,
                                    c->stream));
    }
    return launch_pyr_build_many(c, pyrs, n, dev.data(), dp.data(), nullptr, flags);
}

int vtg_pyr_level(const vtg_pyr *p, int level, void **dptr, int *w, int *h, size_t *pitch_bytes, int *elem)
{
    if (!p || level < 0 || level >= p->levels)
        return VTG_EINVAL;
    if (dptr)
        *dptr = (void *)p->lv[level].p;
    if (w)
        *w = p->lv[level].w;
    if (h)
        *h = p->lv[level].h;
    if (pitch_bytes)
        *pitch_bytes = (size_t)p->lv[level].pitch * p->elem[level];
    if (elem)
        *elem = p->elem[level];
    return VTG_OK;
}

int vtg_pyr_read_level(vtg_ctx *c, const vtg_pyr *p, int level, const struct vt_plane *dst)
{
    if (!c || !p || !dst || level < 0 || level >= p->levels)
        return VTG_EINVAL;
    const Level &L = p->lv[level];
    if (dst->w != L.w || dst->h != L.h || dst->stride != plane_stride(L.w))
        return fail(c, VTG_EINVAL, "vtg_pyr_read_level: plane %dx%d stride %d, level %dx%d", dst->w, dst->h,
                    dst->stride, L.w, L.h);
    const size_t bytes = plane_bytes(L.w, L.h);
    void *tmp = nullptr;
    int rc = scratch(c, bytes, &tmp);
    if (rc)
        return rc;
    const dim3 grid((dst->stride + 255) / 256, L.h + 2 * VT_PAD_Y);
    const int es = p->elem[level];
    if (p->storage == VTG_STORE_F32 || p->storage == VTG_STORE_F32_U8 || (p->storage == VTG_STORE_Q6F && es == 4))
        hipLaunchKernelGGL(to_plane<float>, grid, dim3(256), 0, c->stream, (const float *)L.p, L.pitch, L.w, L.h,
                           (float *)tmp, dst->stride);
    else if (es == 1)
        hipLaunchKernelGGL(to_plane<uint8_t>, grid, dim3(256), 0, c->stream, (const uint8_t *)L.p, L.pitch, L.w, L.h,
                           (float *)tmp, dst->stride);
    else if (p->storage == VTG_STORE_Q6 || p->storage == VTG_STORE_Q6W)
        hipLaunchKernelGGL(to_plane<uint16_t>, grid, dim3(256), 0, c->stream, (const uint16_t *)L.p, L.pitch, L.w,
                           L.h, (float *)tmp, dst->stride);
    else if (p->storage == VTG_STORE_F16)
        hipLaunchKernelGGL(to_plane<__half>, grid, dim3(256), 0, c->stream, (const __half *)L.p, L.pitch, L.w, L.h,
                           (float *)tmp, dst->stride);
    else
        hipLaunchKernelGGL(to_plane<__hip_bfloat16>, grid, dim3(256), 0, c->stream, (const __hip_bfloat16 *)L.p,
                           L.pitch, L.w, L.h, (float *)tmp, dst->stride);
    VTG_HIP(c, hipGetLastError());
    float *base = dst->p - (size_t)dst->stride * VT_PAD_Y - VT_PAD_X;
    VTG_HIP(c, hipMemcpyAsync(base, tmp, bytes, hipMemcpyDeviceToHost, c->stream));
    VTG_HIP(c, hipStreamSynchronize(c->stream));
    return VTG_OK;
}

// ---- reference-shaped calls over host planes

int vtg_img_u8_to_q6(vtg_ctx *c, const struct vt_plane *dst, const uint8_t *src, int src_stride, int y0, int y1)
{
    if (!c || !dst || !src || y0 < 0 || y1 > dst->h || y0 > y1 || src_stride < dst->w)
        return fail(c, VTG_EINVAL, "vtg_img_u8_to_q6: bad argument");
    if (y0 == y1)
        return VTG_OK;
    vtg_pyr *p = nullptr;
    int rc = vtg_pyr_create(c, dst->w, y1 - y0, 1, VTG_STORE_F32, &p);
    if (rc)
        return rc;
    rc = vtg_pyr_build(c, p, src + (size_t)y0 * src_stride, src_stride, 0);
    if (!rc) {
        // rows y0..y1 of dst as a plane of their own: the padding of each row is the same either way
        struct vt_plane part = *dst;
        part.p = plane_row(dst, y0);
        part.h = y1 - y0;
        const Level &L = p->lv[0];
        void *tmp = nullptr;
        const size_t bytes = (size_t)dst->stride * part.h * sizeof(float);
        rc = scratch(c, plane_bytes(dst->w, part.h), &tmp);
        if (!rc) {
            const dim3 grid((dst->stride + 255) / 256, part.h + 2 * VT_PAD_Y);
            hipLaunchKernelGGL(to_plane<float>, grid, dim3(256), 0, c->stream, (const float *)L.p, L.pitch, L.w, L.h,
                               (float *)tmp, dst->stride);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess)
                e = hipMemcpyAsync(part.p - VT_PAD_X, (float *)tmp + (size_t)VT_PAD_Y * dst->stride, bytes,
                                   hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess)
                e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess)
                rc = hip_fail(c, e, "vtg_img_u8_to_q6");
        }
    }
    vtg_pyr_destroy(c, p);
    return rc;
}

int vtg_img_reduce(vtg_ctx *c, const struct vt_plane *dst, const struct vt_plane *src, int y0, int y1)
{
    if (!c || !dst || !src || dst->w != (src->w + 1) / 2 || dst->h != (src->h + 1) / 2 || y0 < 0 || y1 > dst->h ||
        y0 > y1)
        return fail(c, VTG_EINVAL, "vtg_img_reduce: bad argument");
    if (y0 == y1)
        return VTG_OK;
    vtg_pyr *p = nullptr;
    int rc = vtg_pyr_create(c, src->w, src->h, 2, VTG_STORE_F32, &p);
    if (rc)
        return rc;
    void *tmp = nullptr;
    const size_t sbytes = plane_bytes(src->w, src->h), dbytes = plane_bytes(dst->w, dst->h);
    rc = scratch(c, sbytes > dbytes ? sbytes : dbytes, &tmp);
    const Level *L = p->lv;
    hipError_t e = hipSuccess;
    if (!rc) {
        const float *base = src->p - (size_t)src->stride * VT_PAD_Y - VT_PAD_X;
        e = hipMemcpyAsync(tmp, base, sbytes, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(from_plane, dim3((src->w + 255) / 256, src->h), dim3(256), 0, c->stream,
                               (const float *)tmp, src->stride, src->w, src->h, (float *)L[0].p, L[0].pitch);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
        {
            std::vector<RJob> jobs{RJob{L[0].p, (void *)L[1].p, nullptr, L[0].w, L[0].h, L[0].pitch, L[1].w, L[1].h,
                                        L[1].pitch, 0, 0, 0}};
            rc = launch_jobs<float, float, float, false, true>(c, jobs, std::vector<int>{16});
        }
        if (e == hipSuccess && !rc) {
            const dim3 grid((dst->stride + 255) / 256, dst->h + 2 * VT_PAD_Y);
            hipLaunchKernelGGL(to_plane<float>, grid, dim3(256), 0, c->stream, (const float *)L[1].p, L[1].pitch,
                               L[1].w, L[1].h, (float *)tmp, dst->stride);
            e = hipGetLastError();
            if (e == hipSuccess)
                e = hipMemcpyAsync(plane_row(dst, y0) - VT_PAD_X,
                                   (float *)tmp + (size_t)(VT_PAD_Y + y0) * dst->stride,
                                   (size_t)(y1 - y0) * dst->stride * sizeof(float), hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess)
                e = hipStreamSynchronize(c->stream);
        }
        if (e != hipSuccess)
            rc = hip_fail(c, e, "vtg_img_reduce");
    }
    vtg_pyr_destroy(c, p);
    return rc;
}

}  // extern "C"
