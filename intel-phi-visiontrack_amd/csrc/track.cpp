// track.cpp: the tracker's job-op layer (include/vtg_track.h) over the library's own calls: a ring of pyramid sets
// (tk_build: OP_U8 + OP_REDUCE as one vtg_pyr_build per frame, track.c:584-629), LK of the feature list from the
// previous set with one guess for all (OP_LK, track.c:275-292, the guess from the head's rate, track.c:2113-2114),
// failed features dropped (track.c:1305), and an armed appearance searched (OP_SEARCH + search_slot,
// track.c:301-329, 1661-1726). Host bookkeeping lives in memory from the caller's allocator and the per-feature host
// work runs on the caller's executor, as the reference's tracker takes both (track.h:12-22).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "vtg_track.h"

struct vtg_tracker {
    vtg_tk_config cfg;
    vtg_ctx *ctx = nullptr;
    int nsets = 0;
    vtg_pyr *sets[4] = {nullptr, nullptr, nullptr, nullptr};
    int built[4] = {0, 0, 0, 0};       // the set holds a frame
    int last_set = -1;                 // the set of the last tracked frame
    tk_alloc_fn alloc = nullptr;
    void *alloc_ctx = nullptr;
    tk_run_fn run_pyr = nullptr, run_track = nullptr;
    void *pyr_ctx = nullptr, *track_ctx = nullptr;
    int threads_pyr = 1, threads_track = 1;
    // the feature list (host, from alloc)
    float *fx = nullptr, *fy = nullptr;
    uint32_t *fnew = nullptr;
    int nf = 0;
    // LK staging: xy, guess and records (host from alloc; device from one allocation)
    float *hxy = nullptr, *hguess = nullptr;
    lk_out *hout = nullptr;
    void *dev = nullptr;
    float *dxy = nullptr, *dguess = nullptr;
    lk_out *dout = nullptr;
    uint8_t *dframe = nullptr;
    // the head's rate and the clock of the last tracked frame
    double rate_u = 0.0, rate_v = 0.0;
    uint64_t last_ns = 0;
    bool have_ns = false;
    // the armed appearance
    vtg_tmpl *app = nullptr;
    int app_level = 0, app_tw = 0, app_th = 0, app_cx = 0, app_cy = 0, app_r = -1;
    // tk_level's plane
    vt_plane level_plane{nullptr, 0, 0, 0};
    float *level_mem = nullptr;
    size_t level_bytes = 0;
    char err[512] = {0};
};

namespace {

int tfail(vtg_tracker *t, int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t->err, sizeof t->err, fmt, ap);
    va_end(ap);
    return code;
}

int tctx(vtg_tracker *t, int rc, const char *what)
{
    if (rc)
        snprintf(t->err, sizeof t->err, "%s: %s", what, vtg_ctx_error(t->ctx));
    return rc;
}

void *take(vtg_tracker *t, size_t bytes)
{
    void *p = t->alloc(t->alloc_ctx, (bytes + 63) & ~(size_t)63);
    if (p)
        memset(p, 0, bytes);
    return p;
}

// The build's host side on the pyramid executor: tid 0 hands the frame to the device and builds the set (the
// reference's OP_U8 / OP_REDUCE rows run on the team; here the rows are the device's).
struct BuildJob {
    vtg_tracker *t;
    int set;
    const uint8_t *left;
    int rc;
};

void build_job(void *arg, int tid, int n)
{
    (void)n;
    BuildJob *j = (BuildJob *)arg;
    if (tid != 0)
        return;
    vtg_tracker *t = j->t;
    const size_t bytes = (size_t)t->cfg.width * t->cfg.height;
    j->rc = tctx(t, vtg_upload(t->ctx, t->dframe, j->left, bytes), "tk_build: upload");
    if (!j->rc)
        j->rc = tctx(t, vtg_pyr_build(t->ctx, t->sets[j->set], t->dframe, (int)t->cfg.width, 1), "tk_build");
}

// The per-feature host side of tracking on the track executor: features [i0, i1) of each thread's share take their
// records (position, displacement, sigma; keep or drop).
struct TrackJob {
    vtg_tracker *t;
    std::vector<uint8_t> *keep;
    vtg_tk_result *out;
};

void track_job(void *arg, int tid, int n)
{
    TrackJob *j = (TrackJob *)arg;
    vtg_tracker *t = j->t;
    const int i0 = (int)((long)t->nf * tid / n), i1 = (int)((long)t->nf * (tid + 1) / n);
    for (int i = i0; i < i1; i++) {
        const lk_out &r = t->hout[i];
        (*j->keep)[i] = r.status == 0;
    }
}

}  // namespace

extern "C" {

vtg_tracker *vtg_tk_create(const vtg_tk_config *cfg, int threads_pyr, int threads_track, tk_alloc_fn alloc,
                           void *alloc_ctx, tk_run_fn run_pyr, void *pyr_ctx, tk_run_fn run_track, void *track_ctx)
{
    if (!cfg || !alloc || !run_pyr || !run_track || threads_pyr < 1 || threads_track < 1) {
        fprintf(stderr, "vtg_tk_create: a config, an allocator and two executors are required\n");
        return nullptr;
    }
    if (cfg->width < 16 || cfg->height < 16 || cfg->levels < 1 || cfg->levels > VTG_MAX_LEVELS || cfg->sets < 2 ||
        cfg->sets > 4 || cfg->max_features < 1 || cfg->max_features > VTG_TK_MAX_FEATURES) {
        fprintf(stderr, "vtg_tk_create: %ux%u, %u levels, %u sets, %u features: out of range\n", cfg->width,
                cfg->height, cfg->levels, cfg->sets, cfg->max_features);
        return nullptr;
    }
    vtg_tracker *t = new vtg_tracker;
    t->cfg = *cfg;
    t->alloc = alloc;
    t->alloc_ctx = alloc_ctx;
    t->run_pyr = run_pyr;
    t->pyr_ctx = pyr_ctx;
    t->run_track = run_track;
    t->track_ctx = track_ctx;
    t->threads_pyr = threads_pyr;
    t->threads_track = threads_track;
    t->nsets = (int)cfg->sets;
    const int nf = (int)cfg->max_features;
    t->fx = (float *)take(t, sizeof(float) * nf);
    t->fy = (float *)take(t, sizeof(float) * nf);
    t->fnew = (uint32_t *)take(t, sizeof(uint32_t) * nf);
    t->hxy = (float *)take(t, 2 * sizeof(float) * nf);
    t->hguess = (float *)take(t, 2 * sizeof(float) * nf);
    t->hout = (lk_out *)take(t, sizeof(lk_out) * nf);
    if (!t->fx || !t->fy || !t->fnew || !t->hxy || !t->hguess || !t->hout) {
        fprintf(stderr, "vtg_tk_create: the allocator returned NULL\n");
        delete t;
        return nullptr;
    }
    int rc = vtg_ctx_create(cfg->device, &t->ctx);
    for (int s = 0; !rc && s < t->nsets; s++)
        rc = vtg_pyr_create(t->ctx, (int)cfg->width, (int)cfg->height, (int)cfg->levels, cfg->storage, &t->sets[s]);
    const size_t frame = ((size_t)cfg->width * cfg->height + 255) & ~(size_t)255;
    const size_t lists = 2 * (2 * sizeof(float) * nf) + sizeof(lk_out) * nf;
    if (!rc)
        rc = vtg_alloc(t->ctx, frame + lists, &t->dev);
    if (rc) {
        fprintf(stderr, "vtg_tk_create: %s\n", t->ctx ? vtg_ctx_error(t->ctx) : "no device");
        vtg_tk_destroy(t);
        return nullptr;
    }
    t->dframe = (uint8_t *)t->dev;
    t->dxy = (float *)((char *)t->dev + frame);
    t->dguess = t->dxy + 2 * nf;
    t->dout = (lk_out *)(t->dguess + 2 * nf);
    return t;
}

void vtg_tk_destroy(vtg_tracker *t)
{
    if (!t)
        return;
    if (t->ctx) {
        (void)vtg_ctx_sync(t->ctx);
        if (t->app)
            (void)vtg_ncc_template_destroy(t->ctx, t->app);
        for (int s = 0; s < 4; s++)
            if (t->sets[s])
                (void)vtg_pyr_destroy(t->ctx, t->sets[s]);
        if (t->dev)
            (void)vtg_free(t->ctx, t->dev);
        (void)vtg_ctx_destroy(t->ctx);
    }
    // host memory came from the caller's allocator, which the tracker never frees (track.h:15)
    delete t;
}

const char *vtg_tk_error(const vtg_tracker *t) { return t ? t->err : "no tracker"; }

int vtg_tk_sets(const vtg_tracker *t) { return t ? t->nsets : 0; }

vtg_ctx *vtg_tk_ctx(vtg_tracker *t) { return t ? t->ctx : nullptr; }

int vtg_tk_build(vtg_tracker *t, int set, const uint8_t *left, const vtg_msg *m)
{
    (void)m; // the build keys (pre_gauss, contrast) are not acted on by this layer
    if (!t)
        return VTG_EINVAL;
    if (set < 0 || set >= t->nsets || !left)
        return tfail(t, VTG_EINVAL, "vtg_tk_build: set %d of %d, frame %p", set, t->nsets, (const void *)left);
    BuildJob j{t, set, left, 0};
    t->run_pyr(t->pyr_ctx, build_job, &j);
    if (!j.rc)
        t->built[set] = 1;
    return j.rc;
}

int vtg_tk_set_features(vtg_tracker *t, int n, const float *x, const float *y)
{
    if (!t)
        return VTG_EINVAL;
    t->nf = 0;
    return vtg_tk_add_features(t, n, x, y);
}

int vtg_tk_add_features(vtg_tracker *t, int n, const float *x, const float *y)
{
    if (!t)
        return VTG_EINVAL;
    if (n < 0 || (n > 0 && (!x || !y)))
        return tfail(t, VTG_EINVAL, "vtg_tk_add_features: %d features", n);
    if (t->nf + n > (int)t->cfg.max_features)
        return tfail(t, VTG_ERANGE, "vtg_tk_add_features: %d + %d features, at most %u", t->nf, n,
                     t->cfg.max_features);
    for (int i = 0; i < n; i++) {
        t->fx[t->nf + i] = x[i];
        t->fy[t->nf + i] = y[i];
        t->fnew[t->nf + i] = 1;
    }
    t->nf += n;
    return VTG_OK;
}

int vtg_tk_set_rate(vtg_tracker *t, double rate_u, double rate_v)
{
    if (!t)
        return VTG_EINVAL;
    t->rate_u = rate_u;
    t->rate_v = rate_v;
    return VTG_OK;
}

int vtg_tk_arm_search(vtg_tracker *t, int set, int level, int x0, int y0, int tw, int th, int cx, int cy, int r)
{
    if (!t)
        return VTG_EINVAL;
    if (r < 0) {
        t->app_r = -1;
        return VTG_OK;
    }
    if (set < 0 || set >= t->nsets || !t->built[set] || level < 0 || level >= (int)t->cfg.levels)
        return tfail(t, VTG_EINVAL, "vtg_tk_arm_search: set %d (built %d), level %d", set,
                     set >= 0 && set < t->nsets ? t->built[set] : 0, level);
    if (t->app && (t->app_tw != tw || t->app_th != th || t->app_level != level)) {
        (void)vtg_ctx_sync(t->ctx);
        (void)vtg_ncc_template_destroy(t->ctx, t->app);
        t->app = nullptr;
    }
    int rc = t->app ? vtg_ncc_template_update(t->ctx, t->app, t->sets[set], level, x0, y0)
                    : vtg_ncc_template(t->ctx, t->sets[set], level, x0, y0, tw, th, &t->app);
    if (rc)
        return tctx(t, rc, "vtg_tk_arm_search");
    t->app_level = level;
    t->app_tw = tw;
    t->app_th = th;
    t->app_cx = cx;
    t->app_cy = cy;
    t->app_r = r;
    return VTG_OK;
}

int vtg_tk_track(vtg_tracker *t, int set, const vtg_msg *m, vtg_tk_result *out)
{
    if (!t)
        return VTG_EINVAL;
    if (set < 0 || set >= t->nsets || !t->built[set] || !m || !out)
        return tfail(t, VTG_EINVAL, "vtg_tk_track: set %d (built %d), message %p, result %p", set,
                     set >= 0 && set < t->nsets ? t->built[set] : 0, (const void *)m, (void *)out);
    memset(out, 0, offsetof(vtg_tk_result, feat));
    out->instance = m->instance;
    out->search.ncc = out->search.second = -2.0;
    // the time since the last tracked frame and every feature's start: the head's rate times it
    const double dt = t->have_ns ? (double)(int64_t)(m->acq_ns - t->last_ns) * 1e-9 : 0.0;
    t->last_ns = m->acq_ns;
    t->have_ns = true;
    const float gu = (float)(t->rate_u * dt), gv = (float)(t->rate_v * dt);
    out->guess_u = gu;
    out->guess_v = gv;
    const int prev = t->last_set;
    const bool flow = (m->dest & VTG_PIPE_FLOW) && prev >= 0 && prev != set && t->built[prev] && t->nf > 0;
    int rc = VTG_OK;
    std::vector<uint8_t> keep((size_t)t->nf, 1);
    std::vector<float> du, dv;
    if (flow) {
        for (int i = 0; i < t->nf; i++) {
            t->hxy[2 * i] = t->fx[i];
            t->hxy[2 * i + 1] = t->fy[i];
            t->hguess[2 * i] = gu;
            t->hguess[2 * i + 1] = gv;
        }
        const size_t pb = 2 * sizeof(float) * t->nf;
        lk_cfg lk{(int)t->cfg.patch_w, (int)t->cfg.patch_h, (int)t->cfg.max_iter, t->cfg.eps, t->cfg.min_eig,
                  t->cfg.max_residual, 0};
        rc = tctx(t, vtg_upload(t->ctx, t->dxy, t->hxy, pb), "tk_track: upload");
        if (!rc)
            rc = tctx(t, vtg_upload(t->ctx, t->dguess, t->hguess, pb), "tk_track: upload");
        if (!rc)
            rc = tctx(t, vtg_lk_track(t->ctx, t->sets[prev], t->sets[set], (int)t->cfg.levels - 1, &lk, t->nf, t->dxy,
                                      t->dguess, t->cfg.fb_max, t->dout), "tk_track: OP_LK");
        if (!rc)
            rc = tctx(t, vtg_download(t->ctx, t->hout, t->dout, sizeof(lk_out) * t->nf), "tk_track: download");
        if (rc)
            return rc;
        TrackJob j{t, &keep, out};
        t->run_track(t->track_ctx, track_job, &j);
    }
    // the result: kept features at their new places, then the new ones; the list keeps the same
    int n = 0;
    for (int i = 0; i < t->nf; i++) {
        vtg_feat_out &f = out->feat[n];
        memset(&f, 0, sizeof f);
        if (flow) {
            if (!keep[i]) {
                out->dropped++;
                continue;
            }
            const lk_out &r = t->hout[i];
            f.x = r.x;
            f.y = r.y;
            f.u = (float)((double)r.x - (double)t->fx[i]);
            f.v = (float)((double)r.y - (double)t->fy[i]);
            f.sigma = std::sqrt(std::max(std::max(r.cov_xx, r.cov_yy), 0.0f));
            du.push_back(f.u);
            dv.push_back(f.v);
            out->tracked++;
        } else {
            f.x = t->fx[i];
            f.y = t->fy[i];
            f.flags = VTG_FF_NEW;
        }
        t->fx[n] = f.x;
        t->fy[n] = f.y;
        t->fnew[n] = 0;
        n++;
    }
    t->nf = n;
    out->nfeat = (uint32_t)n;
    // the head's rate: the median displacement over the time
    if (!du.empty() && dt > 0.0) {
        std::nth_element(du.begin(), du.begin() + du.size() / 2, du.end());
        std::nth_element(dv.begin(), dv.begin() + dv.size() / 2, dv.end());
        t->rate_u = du[du.size() / 2] / dt;
        t->rate_v = dv[dv.size() / 2] / dt;
    }
    out->rate_u = t->rate_u;
    out->rate_v = t->rate_v;
    // the armed appearance, searched in this frame's set
    if (t->app && t->app_r >= 0) {
        vtg_found f;
        rc = tctx(t, vtg_ncc_search(t->ctx, t->sets[set], t->app, t->app_cx - t->app_r, t->app_cx + t->app_r,
                                    t->app_cy - t->app_r, t->app_cy + t->app_r, 1, nullptr, &f),
                  "tk_track: OP_SEARCH");
        if (rc)
            return rc;
        out->search.armed = 1;
        out->search.hx = f.hx;
        out->search.hy = f.hy;
        out->search.ncc = f.ncc;
        out->search.second = f.second;
        // the match's centre: the corner plus half the template, in level-0 pixels
        const double sc = (double)(1 << t->app_level);
        out->search.x = (float)((0.5 * f.hx + 0.5 * (t->app_tw - 1)) * sc);
        out->search.y = (float)((0.5 * f.hy + 0.5 * (t->app_th - 1)) * sc);
    }
    t->last_set = set;
    return VTG_OK;
}

const vt_plane *vtg_tk_level(vtg_tracker *t, int set, int level)
{
    if (!t || set < 0 || set >= t->nsets || !t->built[set] || level < 0 || level >= (int)t->cfg.levels) {
        if (t)
            tfail(t, VTG_EINVAL, "vtg_tk_level: set %d, level %d", set, level);
        return nullptr;
    }
    int w = (int)t->cfg.width, h = (int)t->cfg.height;
    for (int l = 0; l < level; l++)
        w = (w + 1) / 2, h = (h + 1) / 2;
    const int stride = ((w + 15) & ~15) + 64; // plane_stride (plane.h:20-24): 32 floats of pad each side
    const size_t bytes = sizeof(float) * (size_t)stride * (size_t)(h + 8);
    if (bytes > t->level_bytes) {
        // the allocator's memory is never freed, so a larger plane is a new allocation
        t->level_mem = (float *)take(t, bytes);
        if (!t->level_mem) {
            tfail(t, VTG_ENOMEM, "vtg_tk_level: the allocator returned NULL");
            return nullptr;
        }
        t->level_bytes = bytes;
    }
    t->level_plane.p = t->level_mem + (size_t)stride * 4 + 32;
    t->level_plane.w = w;
    t->level_plane.h = h;
    t->level_plane.stride = stride;
    if (tctx(t, vtg_pyr_read_level(t->ctx, t->sets[set], level, &t->level_plane), "vtg_tk_level"))
        return nullptr;
    return &t->level_plane;
}

}  // extern "C"
