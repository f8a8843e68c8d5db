// ctx.cpp: contexts, device memory, errors and the per-kernel event timer of the C-ABI.
#include <cstring>

#include "vtg_internal.hpp"

namespace vtg {

int fail(vtg_ctx *c, int code, const char *fmt, ...)
{
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof c->err, fmt, ap);
        va_end(ap);
    }
    return code;
}

int hip_fail(vtg_ctx *c, hipError_t e, const char *what)
{
    return fail(c, e == hipErrorOutOfMemory ? VTG_ENOMEM : VTG_EIO, "%s: %s", what, hipGetErrorString(e));
}

static hipEvent_t take_event(vtg_ctx *c)
{
    if (!c->free_events.empty()) {
        hipEvent_t e = c->free_events.back();
        c->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess)
        return nullptr;
    return e;
}

void timer_begin(vtg_ctx *c, const char *name, hipEvent_t *a)
{
    (void)name;
    *a = nullptr;
    if (!c->timing)
        return;
    *a = take_event(c);
    if (*a)
        (void)hipEventRecord(*a, c->stream);
}

void timer_end(vtg_ctx *c, const char *name, hipEvent_t a)
{
    if (!c->timing || !a)
        return;
    hipEvent_t b = take_event(c);
    if (!b)
        return;
    (void)hipEventRecord(b, c->stream);
    c->pending.push_back(Timer{name, a, b});
}

static void collect(vtg_ctx *c)
{
    for (auto &t : c->pending) {
        (void)hipEventSynchronize(t.b);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, t.a, t.b);
        bool found = false;
        for (auto &s : c->stats)
            if (s.name == t.name) {
                s.ms += ms;
                s.launches++;
                found = true;
                break;
            }
        if (!found)
            c->stats.push_back({t.name, (double)ms, 1});
        c->free_events.push_back(t.a);
        c->free_events.push_back(t.b);
    }
    c->pending.clear();
}

int scratch(vtg_ctx *c, size_t bytes, void **p)
{
    if (bytes > c->scratch_bytes) {
        if (c->scratch) {
            VTG_HIP(c, hipStreamSynchronize(c->stream));
            VTG_HIP(c, hipFree(c->scratch));
            c->scratch = nullptr;
            c->scratch_bytes = 0;
        }
        const size_t want = (bytes + ((size_t)1 << 20) - 1) & ~(((size_t)1 << 20) - 1);
        VTG_HIP(c, hipMalloc(&c->scratch, want));
        c->scratch_bytes = want;
    }
    *p = c->scratch;
    return VTG_OK;
}

int lk_scratch(vtg_ctx *c, size_t bytes, void **p)
{
    if (bytes > c->lk_buf_bytes) {
        if (c->lk_buf) {
            VTG_HIP(c, hipStreamSynchronize(c->stream));
            VTG_HIP(c, hipFree(c->lk_buf));
            c->lk_buf = nullptr;
            c->lk_buf_bytes = 0;
        }
        const size_t want = (bytes + ((size_t)1 << 20) - 1) & ~(((size_t)1 << 20) - 1);
        VTG_HIP(c, hipMalloc(&c->lk_buf, want));
        c->lk_buf_bytes = want;
    }
    *p = c->lk_buf;
    return VTG_OK;
}

int pinned(vtg_ctx *c, size_t bytes, void **p)
{
    if (bytes > c->pinned_bytes) {
        if (c->pinned) {
            VTG_HIP(c, hipStreamSynchronize(c->stream));
            VTG_HIP(c, hipHostFree(c->pinned));
            c->pinned = nullptr;
            c->pinned_bytes = 0;
        }
        const size_t want = (bytes + 4095) & ~(size_t)4095;
        VTG_HIP(c, hipHostMalloc(&c->pinned, want, hipHostMallocDefault));
        c->pinned_bytes = want;
    }
    *p = c->pinned;
    return VTG_OK;
}

}  // namespace vtg

using namespace vtg;

extern "C" {

int vtg_ctx_create(int device, vtg_ctx **out)
{
    if (!out)
        return VTG_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return VTG_EIO;
    if (device < 0 || device >= n)
        return VTG_EINVAL;
    vtg_ctx *c = new vtg_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return VTG_EIO;
    }
    if (fixed_scratch_init(c)) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return VTG_EIO;
    }
    *out = c;
    return VTG_OK;
}

int vtg_ctx_destroy(vtg_ctx *c)
{
    if (!c)
        return VTG_EINVAL;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &t : c->pending) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (auto e : c->free_events)
        (void)hipEventDestroy(e);
    for (auto e : c->marks)
        if (e)
            (void)hipEventDestroy(e);
    if (c->scratch)
        (void)hipFree(c->scratch);
    if (c->filt)
        (void)hipFree(c->filt);
    if (c->fused)
        (void)hipFree(c->fused);
    if (c->lk_buf)
        (void)hipFree(c->lk_buf);
    if (c->pinned)
        (void)hipHostFree(c->pinned);
    if (c->join)
        (void)hipEventDestroy(c->join);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return VTG_OK;
}

int vtg_ctx_sync(vtg_ctx *c)
{
    if (!c)
        return VTG_EINVAL;
    VTG_HIP(c, hipStreamSynchronize(c->stream));
    return VTG_OK;
}

const char *vtg_ctx_error(const vtg_ctx *c) { return c ? c->err : "no context"; }

int vtg_ctx_wait(vtg_ctx *c, vtg_ctx *other)
{
    if (!c || !other)
        return VTG_EINVAL;
    if (c->device != other->device)
        return fail(c, VTG_EINVAL, "vtg_ctx_wait: contexts on different devices");
    if (!other->join)
        VTG_HIP(c, hipEventCreateWithFlags(&other->join, hipEventDisableTiming));
    VTG_HIP(c, hipEventRecord(other->join, other->stream));
    VTG_HIP(c, hipStreamWaitEvent(c->stream, other->join, 0));
    return VTG_OK;
}

int vtg_ctx_mark(vtg_ctx *c, uint64_t *mark)
{
    if (!c || !mark)
        return VTG_EINVAL;
    const uint64_t m = c->mark_next;
    hipEvent_t &e = c->marks[m % vtg_ctx::kMarks];
    if (!e)
        VTG_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    VTG_HIP(c, hipEventRecord(e, c->stream));
    c->mark_next = m + 1;
    *mark = m;
    return VTG_OK;
}

int vtg_ctx_mark_wait(vtg_ctx *c, uint64_t mark)
{
    if (!c)
        return VTG_EINVAL;
    if (mark == 0 || mark >= c->mark_next)
        return fail(c, VTG_EINVAL, "vtg_ctx_mark_wait: mark %llu was not recorded", (unsigned long long)mark);
    // the slot holds this mark or a later one recorded over it; either way it completes no earlier than this mark
    VTG_HIP(c, hipEventSynchronize(c->marks[mark % vtg_ctx::kMarks]));
    return VTG_OK;
}

void *vtg_ctx_stream(vtg_ctx *c) { return c ? (void *)c->stream : nullptr; }

int vtg_alloc(vtg_ctx *c, size_t bytes, void **dptr)
{
    if (!c || !dptr || bytes == 0)
        return fail(c, VTG_EINVAL, "vtg_alloc: bad argument");
    VTG_HIP(c, hipMalloc(dptr, bytes));
    return VTG_OK;
}

int vtg_free(vtg_ctx *c, void *dptr)
{
    if (!c)
        return VTG_EINVAL;
    if (dptr) {
        VTG_HIP(c, hipStreamSynchronize(c->stream));
        VTG_HIP(c, hipFree(dptr));
    }
    return VTG_OK;
}

int vtg_upload(vtg_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (!c || (!dst && bytes) || (!src && bytes))
        return fail(c, VTG_EINVAL, "vtg_upload: bad argument");
    VTG_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    VTG_HIP(c, hipStreamSynchronize(c->stream));
    return VTG_OK;
}

int vtg_download(vtg_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (!c || (!dst && bytes) || (!src && bytes))
        return fail(c, VTG_EINVAL, "vtg_download: bad argument");
    VTG_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    VTG_HIP(c, hipStreamSynchronize(c->stream));
    return VTG_OK;
}

int vtg_download_async(vtg_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (!c || (!dst && bytes) || (!src && bytes))
        return fail(c, VTG_EINVAL, "vtg_download_async: bad argument");
    VTG_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    return VTG_OK;
}

int vtg_upload_async(vtg_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (!c || (!dst && bytes) || (!src && bytes))
        return fail(c, VTG_EINVAL, "vtg_upload_async: bad argument");
    VTG_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    return VTG_OK;
}

int vtg_host_alloc(vtg_ctx *c, size_t bytes, void **host)
{
    if (!c || !host)
        return fail(c, VTG_EINVAL, "vtg_host_alloc: bad argument");
    *host = nullptr;
    const hipError_t e = hipHostMalloc(host, bytes ? bytes : 1, hipHostMallocDefault);
    if (e != hipSuccess)
        return fail(c, VTG_ENOMEM, "vtg_host_alloc: %zu bytes: %s", bytes, hipGetErrorString(e));
    return VTG_OK;
}

int vtg_host_free(vtg_ctx *c, void *host)
{
    if (!c)
        return VTG_EINVAL;
    if (host)
        VTG_HIP(c, hipHostFree(host));
    return VTG_OK;
}

int vtg_copy_d2d(vtg_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (!c || (!dst && bytes) || (!src && bytes))
        return fail(c, VTG_EINVAL, "vtg_copy_d2d: bad argument");
    VTG_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return VTG_OK;
}

int vtg_graph_begin(vtg_ctx *c)
{
    if (!c)
        return VTG_EINVAL;
    if (c->timing)
        return fail(c, VTG_EINVAL, "vtg_graph_begin: per-kernel timing is on (events cannot be captured)");
    VTG_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    return VTG_OK;
}

int vtg_graph_end(vtg_ctx *c, void **graph_exec)
{
    if (!c || !graph_exec)
        return VTG_EINVAL;
    hipGraph_t g = nullptr;
    VTG_HIP(c, hipStreamEndCapture(c->stream, &g));
    hipGraphExec_t e = nullptr;
    hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (err != hipSuccess)
        return hip_fail(c, err, "hipGraphInstantiate");
    *graph_exec = (void *)e;
    return VTG_OK;
}

int vtg_graph_launch(vtg_ctx *c, void *graph_exec)
{
    if (!c || !graph_exec)
        return VTG_EINVAL;
    VTG_HIP(c, hipGraphLaunch((hipGraphExec_t)graph_exec, c->stream));
    return VTG_OK;
}

int vtg_graph_destroy(vtg_ctx *c, void *graph_exec)
{
    if (!c || !graph_exec)
        return VTG_EINVAL;
    VTG_HIP(c, hipStreamSynchronize(c->stream));
    VTG_HIP(c, hipGraphExecDestroy((hipGraphExec_t)graph_exec));
    return VTG_OK;
}

int vtg_timing_enable(vtg_ctx *c, int on)
{
    if (!c)
        return VTG_EINVAL;
    c->timing = on != 0;
    return VTG_OK;
}

int vtg_timing_read(vtg_ctx *c, const char *kernel, double *total_ms, long *launches)
{
    if (!c || !kernel)
        return VTG_EINVAL;
    collect(c);
    double ms = 0;
    long n = 0;
    for (auto &s : c->stats)
        if (s.name == kernel) {
            ms = s.ms;
            n = s.launches;
        }
    if (total_ms)
        *total_ms = ms;
    if (launches)
        *launches = n;
    return VTG_OK;
}

int vtg_timing_reset(vtg_ctx *c)
{
    if (!c)
        return VTG_EINVAL;
    collect(c);
    c->stats.clear();
    return VTG_OK;
}

}  // extern "C"
