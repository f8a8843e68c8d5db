// vtg_internal.hpp: what the library's translation units share: the context, device pyramid and template records,
// error plumbing and the launch timer. Host-side C++ plus the device-side level descriptors.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "vtgpu.h"

namespace vtg {

constexpr int kMaxLevels = VTG_MAX_LEVELS;

// One level as the kernels see it: element pointer, size, pitch in elements.
struct Level {
    const void *p;
    int w, h, pitch;
};

// A pyramid as the kernels see it, passed by value in kernel arguments.
struct PyrView {
    Level l[kMaxLevels];
    int levels;
    int storage;
};

struct Timer {
    std::string name;
    hipEvent_t a, b;
};

// An LK launch schedule: wavefronts (of 4 points each) per workgroup, 1 .. 8.
struct LkSched {
    int waves;
    int ppw; // points per wavefront: 0 (3 where the window's shape allows, else 4), 3 or 4
};

}  // namespace vtg

struct vtg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    char err[512] = {0};
    bool timing = false;
    std::vector<vtg::Timer> pending;         // launches bracketed by events, not yet read
    std::vector<hipEvent_t> free_events;     // recycled events
    struct Stat { std::string name; double ms; long launches; };
    std::vector<Stat> stats;
    // scratch owned by the context (grown on demand, freed on destroy)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    void *filt = nullptr;                    // the contrast stretch's tables and per-frame tables (built on first use)
    void *fused = nullptr;                   // the fused launch's work counter, error word and per-level done counts
    bool fused_init = false;
    void *lk_buf = nullptr;                  // the LK launches' point order (grown on demand, stream-ordered reuse)
    size_t lk_buf_bytes = 0;
    int lk_order = 1;                        // track a launch's points in spatial order (0: in the caller's order)
    int lk_grad = 0;                         // VTG_GRAD_SOBEL or VTG_GRAD_SCHARR
    hipEvent_t join = nullptr;               // for vtg_ctx_wait
    static constexpr int kMarks = 64;
    hipEvent_t marks[kMarks] = {};           // vtg_ctx_mark: mark m recorded in slot m % kMarks
    uint64_t mark_next = 1;
    void *pinned = nullptr;                  // small pinned host buffer for results
    size_t pinned_bytes = 0;
    int lk_ppw = 0;                          // LK points per wavefront where no pick applies (0: the shape's best)
    int lk_waves = 1;                        // LK schedule: wavefronts per workgroup where no pick applies
    // the autotuner's picks per size class (keys from pyr_class / lk_class)
    std::map<uint64_t, int> pyr_rows;
    std::map<uint64_t, vtg::LkSched> lk_sched;
};

struct vtg_pyr {
    int w = 0, h = 0, levels = 0, storage = 0;
    void *block = nullptr;
    size_t bytes = 0;
    vtg::Level lv[vtg::kMaxLevels];
    size_t offset[vtg::kMaxLevels];
    int elem[vtg::kMaxLevels];
    vtg::PyrView *dview = nullptr;           // the levels' descriptor in device memory (read by batched kernels)
    bool owned = true;                       // block from hipMalloc (else placed in a caller's memory)
    bool wrapped = false;                    // level 0 is a caller's device frame (vtg_pyr_wrap): never built
    vtg::Level byte0{nullptr, 0, 0, 0};      // VTG_STORE_F32_U8: the u8 frame beside level 0
    size_t byte0_off = 0;
};

namespace vtg {
// The level the NCC search reads: level 0 of a VTG_STORE_F32_U8 pyramid is the u8 frame kept beside it.
inline const Level &ncc_level(const vtg_pyr *p, int l) { return l == 0 && p->byte0.p ? p->byte0 : p->lv[l]; }
inline int ncc_elem(const vtg_pyr *p, int l) { return l == 0 && p->byte0.p ? 1 : p->elem[l]; }
}  // namespace vtg

struct vtg_tmpl {
    int tw = 0, th = 0, level = 0;
    float norm = 0.0f, mean = 0.0f;
    int64_t sum_u = 0;          // sum of the template's pixels in the level's own units (u8 at level 0)
    void *block = nullptr;      // device: t (fp32 zero-mean, tw*th), raw (tw*th fp32 Q6), shifted u8 words
    float *t = nullptr;
    float *raw = nullptr;
    uint32_t *words = nullptr;  // 4 shifts x th rows x (tw/4 + 1) words, level 0 only
    int words_per_row = 0;
    bool owned = true;          // block from hipMalloc (else placed in a caller's memory)
};

namespace vtg {

int fail(vtg_ctx *c, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
int hip_fail(vtg_ctx *c, hipError_t e, const char *what);

#define VTG_HIP(ctx, call)                                  \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess)                               \
            return ::vtg::hip_fail((ctx), e_, #call);       \
    } while (0)

// Launch bracketing for the per-kernel timer: begin() before a launch, end() after it (no-ops when timing is off).
void timer_begin(vtg_ctx *c, const char *name, hipEvent_t *a);
void timer_end(vtg_ctx *c, const char *name, hipEvent_t a);

// Context scratch of at least `bytes` (device), stream-ordered reuse.
int scratch(vtg_ctx *c, size_t bytes, void **p);
// The LK launches' own scratch (apart from the searches', which another stream may still read).
int lk_scratch(vtg_ctx *c, size_t bytes, void **p);
int pinned(vtg_ctx *c, size_t bytes, void **p);

PyrView view(const vtg_pyr *p);
// A pyramid laid out in `mem` (pyr_bytes of device memory), or in its own allocation when mem is null.
size_t pyr_bytes(int w, int h, int levels, int storage);
int pyr_new(vtg_ctx *c, int w, int h, int levels, int storage, void *mem, vtg_pyr **out);
// A template record of tw x th for a level of element size 1 (u8) or more, laid out in `mem` (tmpl_bytes), or in its
// own allocation when mem is null; its contents come from vtg_ncc_template_update.
size_t tmpl_bytes(int tw, int th, bool u8);
int tmpl_new(vtg_ctx *c, int tw, int th, int level, bool u8, void *mem, vtg_tmpl **out);
// An F32 pyramid holding copies of host planes levels[0..n-1] (their visible pixels).
int pyr_from_planes(vtg_ctx *c, const vt_plane *levels, int n, vtg_pyr **out);

// kernels (one per translation unit)
int launch_pyr_build(vtg_ctx *c, vtg_pyr *p, const uint8_t *frame_dev, int pitch);
int launch_pyr_build_many(vtg_ctx *c, vtg_pyr *const *ps, int n, const uint8_t *const *frames, const int *pitch,
                          const int *rows, int flags = 0);
// Level 0 of pyramids of storage Q6W or F32 from device u8 frames, filtered as `flags` asks (filters.hip).
int launch_level0_many(vtg_ctx *c, vtg_pyr *const *ps, int n, const uint8_t *const *frames, const int *pitch,
                       int flags);
// sched: one schedule for every pair, or null for each pair's size-class pick (the context default otherwise);
// fb_top: the flowback's top level (-1: each pair's own top)
int launch_lk_many(vtg_ctx *c, int np, const vtg_pyr *const *I, const vtg_pyr *const *J, const int *top,
                   const int *off, const lk_cfg *cfg, const float *xy, const float *guess, float fb_max,
                   int solve_mode, lk_out *out, const LkSched *sched, int fb_top = -1);

// Size classes of the autotuner and the picks kept for them.
uint64_t pyr_class(const vtg_pyr *p);
uint64_t lk_class(const vtg_pyr *I, int top, const lk_cfg *cfg);
int tuned_pyr_rows(const vtg_ctx *c, const vtg_pyr *p);
LkSched tuned_lk(const vtg_ctx *c, const vtg_pyr *I, int top, const lk_cfg *cfg);
// Points per wavefront of the LK layout a window runs in by default: 3 where that layout is compiled, else 4.
int lk_default_ppw(const lk_cfg *cfg);
int launch_ncc_template(vtg_ctx *c, const vtg_pyr *p, int level, int x0, int y0, vtg_tmpl *t);
// Look up the fixed search array's address on the context's device (current), once per device (ncc.hip).
int fixed_scratch_init(vtg_ctx *c);
int launch_ncc_score(vtg_ctx *c, const vtg_pyr *p, const vtg_tmpl *t, int n, const int *xy, int half, double *out);

}  // namespace vtg
