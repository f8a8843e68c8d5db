// tune.cpp: the autotuner. For a frame pair of a size class it times every candidate schedule of the two launches that
// carry the batch -- the pyramid reduce (output tiles of 8, 16 or 32 rows) and the LK kernel (3 or 4 points per
// wavefront, where the window allows 3; 1, 2, 4 or 8 wavefronts per workgroup) -- on the caller's own frames and
// points, and keeps the fastest for that class in the context. Later builds and LK launches of pyramids of the class use the picks. A batch of several
// classes stays one LK launch: it takes the pick of the pair holding the most points. Every schedule computes the same
// bytes (the tests hold them to it), so tuning changes time only.
#include <algorithm>
#include <cfloat>

#include "vtg_internal.hpp"

namespace vtg {

uint64_t pyr_class(const vtg_pyr *p)
{
    return ((uint64_t)p->w << 40) ^ ((uint64_t)p->h << 16) ^ ((uint64_t)p->levels << 8) ^ (uint64_t)p->storage;
}

int tuned_pyr_rows(const vtg_ctx *c, const vtg_pyr *p)
{
    const auto it = c->pyr_rows.find(pyr_class(p));
    return it == c->pyr_rows.end() ? 16 : it->second;
}

namespace {

// The median of `reps` timed runs of f on the context's stream (microseconds), after one untimed run.
template <typename F>
int median_us(vtg_ctx *c, int reps, F &&f, float *us)
{
    hipEvent_t a = nullptr, b = nullptr;
    VTG_HIP(c, hipEventCreate(&a));
    hipError_t e = hipEventCreate(&b);
    if (e != hipSuccess) {
        (void)hipEventDestroy(a);
        return hip_fail(c, e, "hipEventCreate");
    }
    int rc = f();
    std::vector<float> t;
    for (int r = 0; r < reps && !rc; r++) {
        e = hipEventRecord(a, c->stream);
        if (e == hipSuccess && !(rc = f()))
            e = hipEventRecord(b, c->stream);
        if (e == hipSuccess && !rc)
            e = hipEventSynchronize(b);
        float ms = 0.0f;
        if (e == hipSuccess && !rc)
            e = hipEventElapsedTime(&ms, a, b);
        if (e != hipSuccess && !rc)
            rc = hip_fail(c, e, "autotune timing");
        t.push_back(1e3f * ms);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc)
        return rc;
    std::sort(t.begin(), t.end());
    *us = t[t.size() / 2];
    return VTG_OK;
}

}  // namespace

}  // namespace vtg

using namespace vtg;

extern "C" {

int vtg_tune_pair(vtg_ctx *c, vtg_pyr *I, vtg_pyr *J, const uint8_t *f0, const uint8_t *f1, int pitch, int top,
                  const struct lk_cfg *cfg, int n, const float *xy, int solve_mode, struct vtg_tune *res)
{
    if (!c || !I || !J || !f0 || !f1 || !cfg || !xy || !res || n <= 0 || pitch < I->w)
        return fail(c, VTG_EINVAL, "vtg_tune_pair: bad argument");
    if (I->w != J->w || I->h != J->h || I->levels != J->levels || I->storage != J->storage || top < 0 ||
        top >= I->levels)
        return fail(c, VTG_EINVAL, "vtg_tune_pair: the pyramids differ or top %d is out of range", top);
    *res = vtg_tune{};
    vtg_pyr *ps[2] = {I, J};
    const uint8_t *fr[2] = {f0, f1};
    const int pt[2] = {pitch, pitch};
    // the pyramid reduce: tile heights
    float best = FLT_MAX;
    for (int rows : {8, 16, 32}) {
        const int rr[2] = {rows, rows};
        float us = 0.0f;
        int rc = median_us(c, 5, [&] { return launch_pyr_build_many(c, ps, 2, fr, pt, rr); }, &us);
        if (rc)
            return rc;
        res->candidates++;
        if (rows == 16)
            res->pyr_us_default = us;
        if (us < best)
            best = us, res->pyr_rows = rows, res->pyr_us = us;
    }
    c->pyr_rows[pyr_class(I)] = res->pyr_rows;
    // the LK kernel: schedules on the caller's points (the pyramids hold the pair after the runs above)
    lk_out *out = nullptr;
    VTG_HIP(c, hipMalloc(&out, (size_t)n * sizeof(lk_out)));
    const int off[2] = {0, n};
    const vtg_pyr *Ic = I, *Jc = J;
    best = FLT_MAX;
    int rc = VTG_OK;
    const LkSched dflt = LkSched{c->lk_waves, c->lk_ppw ? c->lk_ppw : lk_default_ppw(cfg)};
    for (int ppw : {3, 4}) {
        if (ppw == 3 && lk_default_ppw(cfg) != 3)
            continue; // the window has no three-point form
        for (int waves : {1, 2, 4, 8}) {
            if (rc)
                break;
            const LkSched s{waves, ppw};
            float us = 0.0f;
            rc = median_us(c, 3, [&] {
                return launch_lk_many(c, 1, &Ic, &Jc, &top, off, cfg, xy, nullptr, 0.0f, solve_mode, out, &s);
            }, &us);
            if (rc)
                break;
            res->candidates++;
            if (waves == dflt.waves && ppw == dflt.ppw)
                res->lk_us_default = us;
            if (us < best)
                best = us, res->lk_waves = waves, res->lk_ppw = ppw, res->lk_us = us;
        }
    }
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(out);
    if (rc)
        return rc;
    c->lk_sched[lk_class(I, top, cfg)] = LkSched{res->lk_waves, res->lk_ppw};
    return VTG_OK;
}

int vtg_tune_get(vtg_ctx *c, const vtg_pyr *I, int top, const struct lk_cfg *cfg, struct vtg_tune *res)
{
    if (!c || !I || !cfg || !res)
        return fail(c, VTG_EINVAL, "vtg_tune_get: bad argument");
    *res = vtg_tune{};
    res->pyr_rows = tuned_pyr_rows(c, I);
    const LkSched s = tuned_lk(c, I, top, cfg);
    res->lk_waves = s.waves;
    res->lk_ppw = s.ppw ? s.ppw : lk_default_ppw(cfg);
    return VTG_OK;
}

int vtg_tune_reset(vtg_ctx *c)
{
    if (!c)
        return VTG_EINVAL;
    c->pyr_rows.clear();
    c->lk_sched.clear();
    return VTG_OK;
}

}  // extern "C"
