// serve.cpp: the serving loop. Frame pairs arrive from several streams of different resolutions; each stream keeps
// its two most recent pyramids resident in device memory, so a new frame costs one pyramid build plus the LK of its
// points from the previous frame (and, when asked, an NCC search of the new frame for a template cut from the
// previous one).
//
// Requests are queued per stream by vtg_srv_submit (a host copy of the frame and points; the call returns a ticket at
// once). A scheduler thread forms batches continuously: whenever one of the server's lanes (a HIP stream each) is
// free, it takes the oldest waiting frame of every stream that has none in flight, as many as fit the device budget
// and the batch limit, and launches them together (the frames' pyramids in one launch per level, all their points in
// one LK launch per schedule, the searches in one launch per stage). Frames that arrive while batches run are
// admitted into the next free lane, so batches overlap on the device. Admission is by memory: every byte the server
// places on the device -- each stream's resident pyramids and template, each batch's frames, points and results --
// comes from one arena of the configured budget; a batch that does not fit waits for earlier ones to release theirs.
// A stream has at most one frame in flight (its next frame reads the pyramid the previous one builds), which also
// keeps a stream's results in submission order.
//
// Each request's arithmetic is that of the direct calls on the same pyramids and points: which other requests share
// its batch changes nothing in its results.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "vtg_internal.hpp"

using namespace vtg;

namespace {

using Clock = std::chrono::steady_clock;

size_t r256(size_t n) { return (n + 255) & ~(size_t)255; }

// First-fit sub-allocation of the server's device arena in 256-byte granules; adjacent free blocks merge.
struct Arena {
    char *base = nullptr;
    size_t size = 0, used = 0, peak = 0;
    std::map<size_t, size_t> free_; // offset -> bytes

    void init(void *b, size_t n)
    {
        base = (char *)b;
        size = n;
        free_.clear();
        free_[0] = n;
        used = peak = 0;
    }
    size_t largest() const
    {
        size_t m = 0;
        for (const auto &f : free_)
            m = std::max(m, f.second);
        return m;
    }
    void *alloc(size_t n)
    {
        n = r256(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n)
                continue;
            const size_t off = it->first, sz = it->second;
            free_.erase(it);
            if (sz > n)
                free_[off + n] = sz - n;
            used += n;
            peak = std::max(peak, used);
            return base + off;
        }
        return nullptr;
    }
    void release(void *p, size_t n)
    {
        if (!p)
            return;
        n = r256(n);
        used -= n;
        auto it = free_.emplace((size_t)((char *)p - base), n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Req {
    uint64_t ticket = 0;
    int stream = 0, n = 0, ncc_x0 = -1, ncc_y0 = -1, search = -1;
    std::vector<uint8_t> frame;
    std::vector<float> xy, guess;
    Clock::time_point t_submit, t_done;
    bool done = false, has_lk = false, has_found = false;
    int rc = 0;
    std::vector<lk_out> out;
    vtg_found found{0, 0, -2.0, -2.0, 0.0, 0.0, 0, 0, 0, -2.0f};
};

// Pinned host staging of one stream (at most one frame in flight): frame | xy | guess | out | found.
struct Stream {
    bool open = false;
    int w = 0, h = 0, levels = 0, max_points = 0;
    vtg_pyr *pyr[2] = {nullptr, nullptr};
    vtg_tmpl *tmpl = nullptr;
    void *resident = nullptr;
    size_t resident_bytes = 0;
    char *pin = nullptr;
    long frames = 0; // frames launched; frame k builds into pyr[k % 2]
    bool busy = false;
    std::deque<Req *> queue;

    size_t o_xy() const { return r256((size_t)w * h); }
    size_t o_guess() const { return o_xy() + (size_t)8 * max_points; }
    size_t o_out() const { return o_guess() + (size_t)8 * max_points; }
    size_t o_found() const { return o_out() + (size_t)32 * max_points; }
    size_t pin_bytes() const { return o_found() + 256; }
};

struct Lane {
    vtg_ctx *ctx = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    std::vector<Req *> reqs;
    void *block = nullptr;
    size_t block_bytes = 0;
};

}  // namespace

struct vtg_srv {
    vtg_srv_cfg cfg{};
    std::mutex mu;
    std::condition_variable work, fin;
    std::thread sched;
    bool stop = false;
    vtg_ctx *main = nullptr;
    std::vector<Stream> streams;
    std::vector<Lane> lanes;
    void *arena_mem = nullptr;
    Arena arena;
    std::unordered_map<uint64_t, Req *> tickets;
    uint64_t next_ticket = 1;
    size_t stream_rr = 0;
    struct vtg_srv_stats st{};
    char err[256] = {0};
};

namespace {

int srv_fail(vtg_srv *s, int code, const char *msg)
{
    if (s)
        snprintf(s->err, sizeof s->err, "%s", msg);
    return code;
}

// The device bytes of a batch: frames, then points, guesses (when any request has them), records and picks.
struct Layout {
    std::vector<size_t> frame;
    size_t xy = 0, guess = 0, out = 0, found = 0, total = 0;
    bool any_guess = false;
    int points = 0;
};

Layout layout(const vtg_srv *s, const std::vector<Req *> &rs)
{
    Layout L;
    size_t off = 0;
    for (const Req *r : rs) {
        const Stream &st = s->streams[r->stream];
        L.frame.push_back(off);
        off += r256((size_t)st.w * st.h);
        if (r->has_lk) {
            L.points += r->n;
            L.any_guess |= !r->guess.empty();
        }
    }
    L.xy = off;
    off += r256((size_t)8 * L.points);
    L.guess = off;
    if (L.any_guess)
        off += r256((size_t)8 * L.points);
    L.out = off;
    off += r256((size_t)32 * L.points);
    L.found = off;
    off += (size_t)64 * rs.size();
    L.total = r256(off);
    return L;
}

// The most a batch of one frame of a w x h stream with n points can take (points with guesses).
size_t one_frame_bytes(int w, int h, int n)
{
    return r256(r256((size_t)w * h) + 2 * r256((size_t)8 * n) + r256((size_t)32 * n) + 64);
}

// Stage and launch a batch on lane ln (the server lock held; every call here is asynchronous on the lane's stream).
// The batch's points are those of its requests with a previous frame, in request order.
int launch(vtg_srv *s, Lane &ln, const std::vector<Req *> &rs, const Layout &L)
{
    vtg_ctx *c = ln.ctx;
    char *d = (char *)ln.block;
    std::vector<vtg_pyr *> ps;
    std::vector<const uint8_t *> fr;
    std::vector<int> pitch;
    std::vector<const vtg_pyr *> I, J, ncc_p;
    std::vector<const vtg_tmpl *> ncc_t;
    std::vector<int> top, offs{0}, region;
    int pre = 0, nsearch = 0;
    for (size_t i = 0; i < rs.size(); i++) {
        Req *r = rs[i];
        Stream &st = s->streams[r->stream];
        const int cur = (int)(st.frames % 2), prev = 1 - cur;
        std::memcpy(st.pin, r->frame.data(), (size_t)st.w * st.h);
        VTG_HIP(c, hipMemcpyAsync(d + L.frame[i], st.pin, (size_t)st.w * st.h, hipMemcpyHostToDevice, c->stream));
        ps.push_back(st.pyr[cur]);
        fr.push_back((const uint8_t *)(d + L.frame[i]));
        pitch.push_back(st.w);
        if (r->has_lk) {
            float *pxy = (float *)(st.pin + st.o_xy()), *pg = (float *)(st.pin + st.o_guess());
            std::memcpy(pxy, r->xy.data(), (size_t)8 * r->n);
            VTG_HIP(c, hipMemcpyAsync(d + L.xy + (size_t)8 * pre, pxy, (size_t)8 * r->n, hipMemcpyHostToDevice,
                                      c->stream));
            if (L.any_guess) {
                if (r->guess.empty())
                    std::memset(pg, 0, (size_t)8 * r->n);
                else
                    std::memcpy(pg, r->guess.data(), (size_t)8 * r->n);
                VTG_HIP(c, hipMemcpyAsync(d + L.guess + (size_t)8 * pre, pg, (size_t)8 * r->n,
                                          hipMemcpyHostToDevice, c->stream));
            }
            I.push_back(st.pyr[prev]);
            J.push_back(st.pyr[cur]);
            top.push_back(st.levels - 1);
            pre += r->n;
            offs.push_back(pre);
        }
        if (r->has_found) {
            int rc = vtg_ncc_template_update(c, st.tmpl, st.pyr[prev], 0, r->ncc_x0, r->ncc_y0);
            if (rc)
                return rc;
            ncc_p.push_back(st.pyr[cur]);
            ncc_t.push_back(st.tmpl);
            region.insert(region.end(), {1, st.w, 1, st.h});
            r->search = nsearch++;
        }
    }
    int rc = launch_pyr_build_many(c, ps.data(), (int)ps.size(), fr.data(), pitch.data(), nullptr);
    if (!rc && !I.empty())
        rc = launch_lk_many(c, (int)I.size(), I.data(), J.data(), top.data(), offs.data(), &s->cfg.lk,
                            (const float *)(d + L.xy), L.any_guess ? (const float *)(d + L.guess) : nullptr,
                            s->cfg.fb_max, s->cfg.solve_mode, (lk_out *)(d + L.out), nullptr);
    vtg_found *fd = (vtg_found *)(d + L.found);
    if (!rc && nsearch)
        rc = vtg_ncc_search_many(c, nsearch, ncc_p.data(), ncc_t.data(), region.data(), 1, nullptr, fd);
    if (rc)
        return rc;
    // each stream's frame now occupies its pyramid set: its next frame tracks from it (only once the launches went)
    for (Req *r : rs)
        s->streams[r->stream].frames++;
    // results to each stream's pinned staging
    pre = 0;
    for (Req *r : rs) {
        Stream &st = s->streams[r->stream];
        if (r->has_lk) {
            VTG_HIP(c, hipMemcpyAsync(st.pin + st.o_out(), d + L.out + (size_t)32 * pre, (size_t)32 * r->n,
                                      hipMemcpyDeviceToHost, c->stream));
            pre += r->n;
        }
        if (r->has_found)
            VTG_HIP(c, hipMemcpyAsync(st.pin + st.o_found(), fd + r->search, sizeof(vtg_found),
                                      hipMemcpyDeviceToHost, c->stream));
    }
    VTG_HIP(c, hipEventRecord(ln.done, c->stream));
    return VTG_OK;
}

void finish(vtg_srv *s, Lane &ln, int rc)
{
    const auto now = Clock::now();
    for (Req *r : ln.reqs) {
        Stream &st = s->streams[r->stream];
        if (!rc && r->has_lk) {
            r->out.resize(r->n);
            std::memcpy(r->out.data(), st.pin + st.o_out(), (size_t)32 * r->n);
        }
        if (!rc && r->has_found)
            std::memcpy(&r->found, st.pin + st.o_found(), sizeof(vtg_found));
        r->rc = rc;
        r->t_done = now;
        r->done = true;
        std::vector<uint8_t>().swap(r->frame);
        std::vector<float>().swap(r->xy);
        std::vector<float>().swap(r->guess);
        st.busy = false;
        s->st.completed++;
    }
    s->arena.release(ln.block, ln.block_bytes);
    ln.block = nullptr;
    ln.block_bytes = 0;
    ln.reqs.clear();
    ln.busy = false;
    s->fin.notify_all();
}

void loop(vtg_srv *s)
{
    (void)hipSetDevice(s->cfg.device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->stop) {
        bool progress = false;
        for (Lane &ln : s->lanes) {
            if (!ln.busy)
                continue;
            const hipError_t q = hipEventQuery(ln.done);
            if (q == hipSuccess) {
                finish(s, ln, VTG_OK);
                progress = true;
            } else if (q != hipErrorNotReady) {
                snprintf(s->err, sizeof s->err, "serving lane: %s", hipGetErrorString(q));
                finish(s, ln, VTG_EIO);
                progress = true;
            }
        }
        for (Lane &ln : s->lanes) {
            if (ln.busy)
                continue;
            // the oldest waiting frame of each stream with nothing in flight, round robin, while the batch fits
            std::vector<Req *> rs;
            const size_t ns = s->streams.size(), room = s->arena.largest();
            for (size_t q = 0; q < ns && (int)rs.size() < s->cfg.max_batch; q++) {
                Stream &st = s->streams[(s->stream_rr + q) % ns];
                if (!st.open || st.busy || st.queue.empty())
                    continue;
                Req *r = st.queue.front();
                r->has_lk = st.frames > 0 && r->n > 0;
                r->has_found = st.frames > 0 && r->ncc_x0 >= 0;
                rs.push_back(r);
                if (layout(s, rs).total > room) {
                    rs.pop_back();
                    continue;
                }
                st.queue.pop_front();
                st.busy = true;
            }
            if (rs.empty()) {
                // nothing fits: while no lane holds arena memory, a waiting frame that does not fit the whole free
                // arena never will (other streams' pyramids took the room); finish it with VTG_ENOMEM
                bool idle = true;
                for (const Lane &o : s->lanes)
                    idle &= !o.busy;
                if (idle)
                    for (Stream &st : s->streams) {
                        if (!st.open || st.busy || st.queue.empty())
                            continue;
                        Req *r = st.queue.front();
                        const std::vector<Req *> one{r};
                        if (layout(s, one).total <= room)
                            continue;
                        st.queue.pop_front();
                        r->rc = VTG_ENOMEM;
                        r->t_done = Clock::now();
                        r->done = true;
                        s->st.completed++;
                        snprintf(s->err, sizeof s->err, "serving: a %dx%d frame does not fit the free budget", st.w,
                                 st.h);
                        progress = true;
                    }
                if (progress)
                    s->fin.notify_all();
                break;
            }
            s->stream_rr = (s->stream_rr + 1) % ns;
            const Layout L = layout(s, rs);
            ln.block = s->arena.alloc(L.total);
            ln.block_bytes = L.total;
            ln.reqs = rs;
            ln.busy = true;
            int busy = 0;
            for (const Lane &o : s->lanes)
                busy += o.busy;
            s->st.batches++;
            s->st.max_batch = std::max(s->st.max_batch, (int)rs.size());
            s->st.max_inflight = std::max(s->st.max_inflight, busy);
            for (Req *r : rs)
                s->st.lk_points += r->has_lk ? r->n : 0;
            const int rc = launch(s, ln, rs, L);
            if (rc) {
                snprintf(s->err, sizeof s->err, "serving launch: %s", ln.ctx->err);
                (void)hipStreamSynchronize(ln.ctx->stream);
                finish(s, ln, rc);
            }
            progress = true;
        }
        s->st.hbm_peak = std::max(s->st.hbm_peak, s->arena.peak);
        if (!progress)
            s->work.wait_for(lk, std::chrono::microseconds(20));
    }
}

}  // namespace

extern "C" {

int vtg_srv_create(const struct vtg_srv_cfg *cfg, vtg_srv **out)
{
    if (!cfg || !out || cfg->lanes < 1 || cfg->lanes > 8 || cfg->max_batch < 1 || cfg->max_batch > 64 ||
        cfg->hbm_budget < ((size_t)1 << 20) || cfg->tmpl_w < 16 || cfg->tmpl_w > 64 || (cfg->tmpl_w & 15) ||
        cfg->tmpl_h < 8 || cfg->tmpl_h > 64)
        return VTG_EINVAL;
    *out = nullptr;
    if (cfg->lk.w < 5 || cfg->lk.w > LK_MAX_W || cfg->lk.h < 5 || cfg->lk.h > LK_MAX_H ||
        (cfg->solve_mode != VTG_SOLVE_F64 && cfg->solve_mode != VTG_SOLVE_MIXED))
        return VTG_EINVAL;
    vtg_srv *s = new vtg_srv();
    s->cfg = *cfg;
    int rc = vtg_ctx_create(cfg->device, &s->main);
    if (!rc && hipMalloc(&s->arena_mem, cfg->hbm_budget) != hipSuccess)
        rc = VTG_ENOMEM;
    if (!rc)
        s->arena.init(s->arena_mem, cfg->hbm_budget);
    s->lanes.resize(cfg->lanes);
    for (Lane &ln : s->lanes) {
        if (rc)
            break;
        rc = vtg_ctx_create(cfg->device, &ln.ctx);
        if (!rc && hipEventCreateWithFlags(&ln.done, hipEventDisableTiming) != hipSuccess)
            rc = VTG_EIO;
    }
    if (rc) {
        for (Lane &ln : s->lanes) {
            if (ln.done)
                (void)hipEventDestroy(ln.done);
            if (ln.ctx)
                vtg_ctx_destroy(ln.ctx);
        }
        if (s->arena_mem)
            (void)hipFree(s->arena_mem);
        if (s->main)
            vtg_ctx_destroy(s->main);
        delete s;
        return rc;
    }
    s->st.hbm_budget = cfg->hbm_budget;
    s->sched = std::thread(loop, s);
    *out = s;
    return VTG_OK;
}

int vtg_srv_destroy(vtg_srv *s)
{
    if (!s)
        return VTG_EINVAL;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->work.notify_all();
    s->sched.join();
    for (Lane &ln : s->lanes) {
        (void)hipStreamSynchronize(ln.ctx->stream);
        (void)hipEventDestroy(ln.done);
        vtg_ctx_destroy(ln.ctx);
    }
    for (Stream &st : s->streams) {
        for (Req *r : st.queue)
            delete r;
        if (!st.open)
            continue;
        delete st.pyr[0];
        delete st.pyr[1];
        delete st.tmpl;
        (void)hipHostFree(st.pin);
    }
    for (auto &t : s->tickets)
        if (t.second->done)
            delete t.second;
    (void)hipFree(s->arena_mem);
    vtg_ctx_destroy(s->main);
    delete s;
    return VTG_OK;
}

const char *vtg_srv_error(vtg_srv *s) { return s ? s->err : "null server"; }

int vtg_srv_open(vtg_srv *s, int w, int h, int levels, int max_points, int *stream)
{
    if (!s || !stream || max_points < 0 || w < s->cfg.tmpl_w + 4 || h < s->cfg.tmpl_h + 4)
        return srv_fail(s, VTG_EINVAL, "vtg_srv_open: bad argument");
    std::lock_guard<std::mutex> g(s->mu);
    Stream st;
    st.w = w;
    st.h = h;
    st.levels = levels;
    st.max_points = max_points;
    const size_t pb = r256(pyr_bytes(w, h, levels, VTG_STORE_Q6)), tb = r256(tmpl_bytes(s->cfg.tmpl_w, s->cfg.tmpl_h,
                                                                                           true));
    st.resident_bytes = 2 * pb + tb;
    st.resident = s->arena.alloc(st.resident_bytes);
    if (!st.resident)
        return srv_fail(s, VTG_ENOMEM, "vtg_srv_open: the stream's resident pyramids exceed the free budget");
    if (one_frame_bytes(w, h, max_points) > s->arena.largest()) {
        // no frame of this stream could ever be admitted: refuse it now rather than queue frames that never run
        s->arena.release(st.resident, st.resident_bytes);
        return srv_fail(s, VTG_ENOMEM, "vtg_srv_open: the budget left after the resident pyramids holds no frame of "
                                       "this stream");
    }
    char *m = (char *)st.resident;
    int rc = pyr_new(s->main, w, h, levels, VTG_STORE_Q6, m, &st.pyr[0]);
    if (!rc)
        rc = pyr_new(s->main, w, h, levels, VTG_STORE_Q6, m + pb, &st.pyr[1]);
    if (!rc)
        rc = tmpl_new(s->main, s->cfg.tmpl_w, s->cfg.tmpl_h, 0, true, m + 2 * pb, &st.tmpl);
    if (!rc && hipHostMalloc((void **)&st.pin, st.pin_bytes(), hipHostMallocDefault) != hipSuccess)
        rc = VTG_ENOMEM;
    // each lane's search scratch sized for a full batch of this stream's frames, now rather than in the loop
    for (Lane &ln : s->lanes) {
        if (rc)
            break;
        const size_t cells = (size_t)((w - 1) / 16 + 1) * ((h - 1) / 16 + 1) * sizeof(vtg_cell);
        void *p = nullptr;
        rc = scratch(ln.ctx, (size_t)s->cfg.max_batch * (r256(cells) + 8192), &p);
    }
    if (rc) {
        delete st.pyr[0];
        delete st.pyr[1];
        delete st.tmpl;
        if (st.pin)
            (void)hipHostFree(st.pin);
        s->arena.release(st.resident, st.resident_bytes);
        return srv_fail(s, rc, "vtg_srv_open: allocation failed");
    }
    st.open = true;
    s->st.hbm_resident += st.resident_bytes;
    s->streams.push_back(std::move(st));
    *stream = (int)s->streams.size() - 1;
    return VTG_OK;
}

int vtg_srv_close(vtg_srv *s, int stream)
{
    if (!s)
        return VTG_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    if (stream < 0 || stream >= (int)s->streams.size() || !s->streams[stream].open)
        return srv_fail(s, VTG_EINVAL, "vtg_srv_close: no such stream");
    s->fin.wait(lk, [&] { return !s->streams[stream].busy && s->streams[stream].queue.empty(); });
    Stream &st = s->streams[stream];
    delete st.pyr[0];
    delete st.pyr[1];
    delete st.tmpl;
    st.pyr[0] = st.pyr[1] = nullptr;
    st.tmpl = nullptr;
    (void)hipHostFree(st.pin);
    st.pin = nullptr;
    s->arena.release(st.resident, st.resident_bytes);
    s->st.hbm_resident -= st.resident_bytes;
    st.open = false;
    return VTG_OK;
}

int vtg_srv_submit(vtg_srv *s, int stream, const uint8_t *frame, int pitch, int n, const float *xy,
                   const float *guess, int ncc_x0, int ncc_y0, uint64_t *ticket)
{
    if (!s || !frame || !ticket || n < 0 || (n > 0 && !xy))
        return srv_fail(s, VTG_EINVAL, "vtg_srv_submit: bad argument");
    Req *r = new Req();
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (stream < 0 || stream >= (int)s->streams.size() || !s->streams[stream].open) {
            delete r;
            return srv_fail(s, VTG_EINVAL, "vtg_srv_submit: no such stream");
        }
        const Stream &st = s->streams[stream];
        if (pitch < st.w || n > st.max_points) {
            delete r;
            return srv_fail(s, VTG_EINVAL, "vtg_srv_submit: pitch below the width or more points than the stream's "
                                           "maximum");
        }
        if (ncc_x0 >= 0 && (ncc_x0 < 1 || ncc_y0 < 1 || ncc_x0 + s->cfg.tmpl_w >= st.w - 1 ||
                            ncc_y0 + s->cfg.tmpl_h >= st.h - 1)) {
            delete r;
            return srv_fail(s, VTG_EINVAL, "vtg_srv_submit: the template window does not fit the frame");
        }
        r->stream = stream;
        r->n = n;
        r->ncc_x0 = ncc_x0;
        r->ncc_y0 = ncc_y0;
        r->frame.resize((size_t)st.w * st.h);
        for (int y = 0; y < st.h; y++)
            std::memcpy(r->frame.data() + (size_t)y * st.w, frame + (size_t)y * pitch, st.w);
        if (n > 0) {
            r->xy.assign(xy, xy + 2 * (size_t)n);
            if (guess)
                r->guess.assign(guess, guess + 2 * (size_t)n);
        }
        r->t_submit = Clock::now();
        r->ticket = s->next_ticket++;
        s->tickets[r->ticket] = r;
        s->streams[stream].queue.push_back(r);
        s->st.submitted++;
        *ticket = r->ticket;
    }
    s->work.notify_one();
    return VTG_OK;
}

static int deliver(vtg_srv *s, Req *r, struct lk_out *out, int *nrec, struct vtg_found *found, double *latency_ms)
{
    if (out && r->has_lk)
        std::memcpy(out, r->out.data(), (size_t)32 * r->n);
    if (nrec)
        *nrec = r->has_lk ? r->n : 0;
    if (found)
        *found = r->found;
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    const int rc = r->rc;
    s->tickets.erase(r->ticket);
    delete r;
    return rc;
}

int vtg_srv_poll(vtg_srv *s, uint64_t ticket, struct lk_out *out, int *nrec, struct vtg_found *found,
                 double *latency_ms)
{
    if (!s)
        return VTG_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end())
        return srv_fail(s, VTG_EINVAL, "vtg_srv_poll: unknown ticket");
    if (!it->second->done)
        return VTG_EAGAIN;
    return deliver(s, it->second, out, nrec, found, latency_ms);
}

int vtg_srv_wait(vtg_srv *s, uint64_t ticket, struct lk_out *out, int *nrec, struct vtg_found *found,
                 double *latency_ms)
{
    if (!s)
        return VTG_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end())
        return srv_fail(s, VTG_EINVAL, "vtg_srv_wait: unknown ticket");
    Req *r = it->second;
    s->fin.wait(lk, [&] { return r->done; });
    return deliver(s, r, out, nrec, found, latency_ms);
}

int vtg_srv_get_stats(vtg_srv *s, struct vtg_srv_stats *out)
{
    if (!s || !out)
        return VTG_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->st;
    out->hbm_in_use = s->arena.used;
    out->hbm_peak = std::max(s->st.hbm_peak, s->arena.peak);
    return VTG_OK;
}

}  // extern "C"
