// door.cpp: the front door of the serving loop: frames arrive as the reference's wire format describes them
// (src/common/vt_proto.h:143-197), a ring of 88-byte descriptors published by a count, each naming a host slot that
// holds its frame, and are taken as the reference's ingest takes them (src/card/session_track.c:76-139, 188-222):
//   - every descriptor is checked before anything acts on it, in the reference's order: magic and check word, then
//     type / destination / camera, then slot, then length (vt_slot_bytes: w x h, twice that for a stereo camera 0,
//     rounded up to 64 bytes, vt_proto.h:283-287), then the camera's instance order; a bad one is counted as an
//     exception (the last one kept), its slot is free at once, and the session goes on;
//   - a good frame lands: its image is copied out of the host slot (the reference's DMA into a device slot), so the
//     slot is free once the descriptor is taken;
//   - backpressure: at most dev_slots frames are taken and not yet answered; past that the door stops taking
//     descriptors, so the host's slots fill and its own policy acts;
//   - landed frames go to the serving loop while fewer than max_inflight of the door's frames run there (the
//     reference builds a frame's pyramid when a set is free); when the newest landed frame asks for LIVE priority,
//     every landed frame before it is skipped (no result), whichever call took it (build_one, session_track.c:197-206);
//   - answers come back in the order the frames went to the server. A submit that fails leaves its frame landed, to
//     go at the next call.
// Host code over the serving loop; the GPU work is the server's.
#include <cstring>
#include <deque>
#include <vector>

#include "vtg_internal.hpp"

namespace {

uint64_t msg_check(const vtg_msg *m)
{
    uint64_t w[sizeof(vtg_msg) / 8];
    std::memcpy(w, m, sizeof w);
    uint64_t h = 0x9E3779B97F4A7C15ull;
    for (unsigned i = 0; i < sizeof(vtg_msg) / 8 - 1; i++) {
        h ^= w[i];
        h *= 0xC2B2AE3D27D4EB4Full;
        h ^= h >> 29;
    }
    return h;
}

// The bytes a frame of camera k takes in its slot (vt_slot_bytes, src/common/vt_proto.h:283-287).
uint64_t slot_bytes(const vtg_door_cfg &c, int k)
{
    const uint64_t one = (uint64_t)c.w[k] * (uint64_t)c.h[k];
    return ((k == 0 && c.stereo ? 2 : 1) * one + 63) & ~63ull;
}

}  // namespace

struct vtg_door {
    vtg_srv *srv = nullptr;
    vtg_door_cfg cfg{};
    int stream[3] = {-1, -1, -1};
    std::vector<float> xy[3];
    uint64_t taken = 0;                  // descriptors consumed (the reference's frames_taken)
    uint64_t last_instance[3] = {0, 0, 0};
    bool have_instance[3] = {false, false, false};
    uint64_t exceptions = 0, skipped = 0, submitted = 0, answered = 0;
    uint32_t last_exception = VTG_EXC_NONE;
    struct Landed {
        vtg_msg msg;
        std::vector<uint8_t> image;      // the frame (camera 0's left image when stereo)
    };
    std::deque<Landed> landed;           // taken, not yet with the server
    struct Pending {
        vtg_msg msg;
        uint64_t ticket;
    };
    std::deque<Pending> inflight;        // with the server, not yet answered
};

namespace {

// Hand landed frames to the server while fewer than max_inflight run there, the newest LIVE one first.
int pump(vtg_door *d)
{
    const vtg_door_cfg &c = d->cfg;
    const size_t cap = c.max_inflight > 0 ? (size_t)c.max_inflight : 1;
    while (!d->landed.empty() && d->inflight.size() < cap) {
        if (d->landed.size() > 1 && d->landed.back().msg.priority == VTG_PRIO_LIVE) {
            d->skipped += d->landed.size() - 1;
            d->landed.erase(d->landed.begin(), d->landed.end() - 1);
        }
        vtg_door::Landed &f = d->landed.front();
        const int k = (int)f.msg.camera;
        uint64_t ticket = 0;
        const int rc = vtg_srv_submit(d->srv, d->stream[k], f.image.data(), c.w[k], c.npoints[k],
                                      d->xy[k].empty() ? nullptr : d->xy[k].data(), nullptr, -1, -1, &ticket);
        if (rc)
            return rc; // the frame stays landed
        d->inflight.push_back(vtg_door::Pending{f.msg, ticket});
        d->submitted++;
        d->landed.pop_front();
    }
    return VTG_OK;
}

void report(const vtg_door *d, vtg_door_state *st)
{
    if (!st)
        return;
    st->frames_taken = d->taken;
    st->exceptions = d->exceptions;
    st->last_exception = d->last_exception;
    st->skipped = d->skipped;
    st->submitted = d->submitted;
    st->in_flight = d->inflight.size() + d->landed.size();
}

}  // namespace

extern "C" {

uint64_t vtg_msg_check(const struct vtg_msg *m) { return m ? msg_check(m) : 0; }

int vtg_door_create(vtg_srv *srv, const struct vtg_door_cfg *cfg, const float *const *xy, vtg_door **out)
{
    if (!srv || !cfg || !xy || !out)
        return VTG_EINVAL;
    *out = nullptr;
    if (cfg->cameras < 1 || cfg->cameras > 3 || cfg->slots < 1 || cfg->dev_slots < 1 || cfg->max_points < 0 ||
        cfg->max_inflight < 0)
        return VTG_EINVAL;
    for (int k = 0; k < cfg->cameras; k++)
        if (cfg->w[k] < 8 || cfg->h[k] < 8 || cfg->npoints[k] < 0 || cfg->npoints[k] > cfg->max_points ||
            (cfg->npoints[k] > 0 && !xy[k]) || slot_bytes(*cfg, k) > cfg->slot_bytes)
            return VTG_EINVAL;
    vtg_door *d = new vtg_door();
    d->srv = srv;
    d->cfg = *cfg;
    for (int k = 0; k < cfg->cameras; k++) {
        d->xy[k].assign(xy[k], xy[k] + 2 * (size_t)cfg->npoints[k]);
        int rc = vtg_srv_open(srv, cfg->w[k], cfg->h[k], cfg->levels[k], cfg->max_points, &d->stream[k]);
        if (rc) {
            for (int j = 0; j < k; j++)
                vtg_srv_close(srv, d->stream[j]);
            delete d;
            return rc;
        }
    }
    *out = d;
    return VTG_OK;
}

int vtg_door_destroy(vtg_door *d)
{
    if (!d)
        return VTG_EINVAL;
    for (int k = 0; k < d->cfg.cameras; k++)
        vtg_srv_close(d->srv, d->stream[k]);
    delete d;
    return VTG_OK;
}

int vtg_door_ingest(vtg_door *d, const struct vtg_msg *desc, uint64_t frames_pub, const uint8_t *slots,
                    struct vtg_door_state *st)
{
    if (!d || (frames_pub > d->taken && (!desc || !slots)))
        return VTG_EINVAL;
    const vtg_door_cfg &c = d->cfg;
    while (d->taken < frames_pub && d->inflight.size() + d->landed.size() < (size_t)c.dev_slots) {
        vtg_msg m;
        std::memcpy(&m, &desc[d->taken % VTG_DESC_RING], sizeof m);
        uint32_t ex = VTG_EXC_NONE;
        if (m.magic != VTG_MSG_MAGIC || m.check != msg_check(&m))
            ex = VTG_EXC_BAD_CHECK;
        else if (m.type != VTG_MSG_FRAME || !(m.dest & VTG_PIPE_PYR) || m.camera >= (uint32_t)c.cameras)
            ex = VTG_EXC_BAD_TYPE;
        else if (m.slot >= c.slots)
            ex = VTG_EXC_BAD_SLOT;
        else if ((uint64_t)m.length != slot_bytes(c, (int)m.camera))
            ex = VTG_EXC_BAD_LENGTH;
        else if (d->have_instance[m.camera] && m.instance <= d->last_instance[m.camera])
            ex = VTG_EXC_INSTANCE_ORDER;
        d->taken++;
        if (ex != VTG_EXC_NONE) {
            d->exceptions++;
            d->last_exception = ex;
            continue;
        }
        d->last_instance[m.camera] = m.instance;
        d->have_instance[m.camera] = true;
        const int k = (int)m.camera;
        const uint8_t *src = slots + (uint64_t)m.slot * c.slot_bytes;
        d->landed.push_back(vtg_door::Landed{m, std::vector<uint8_t>(src, src + (size_t)c.w[k] * c.h[k])});
    }
    const int rc = pump(d);
    report(d, st);
    return rc;
}

int vtg_door_result(vtg_door *d, int wait, struct vtg_door_result *r, struct lk_out *out, int max_out)
{
    if (!d || !r)
        return VTG_EINVAL;
    int rc = pump(d);
    if (rc)
        return rc;
    if (d->inflight.empty())
        return VTG_EAGAIN;
    const vtg_door::Pending &p = d->inflight.front();
    const int k = (int)p.msg.camera;
    if ((!out && d->cfg.npoints[k] > 0) || max_out < d->cfg.npoints[k])
        return VTG_EINVAL;
    int nrec = 0;
    vtg_found found;
    double ms = 0.0;
    rc = wait ? vtg_srv_wait(d->srv, p.ticket, out, &nrec, &found, &ms)
              : vtg_srv_poll(d->srv, p.ticket, out, &nrec, &found, &ms);
    if (rc)
        return rc;
    r->instance = p.msg.instance;
    r->camera = p.msg.camera;
    r->slot = p.msg.slot;
    r->acq_ns = p.msg.acq_ns;
    r->nrec = nrec;
    r->latency_ms = ms;
    d->inflight.pop_front();
    d->answered++;
    return pump(d); // the next landed frame takes the freed place
}

}  // extern "C"
