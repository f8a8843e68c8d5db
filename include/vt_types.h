/* vt_types.h: the data types that cross the drop-in boundary, laid out exactly as the reference lays them out so a
 * host program built against the reference's headers passes the same bytes.
 *
 *   struct vt_plane  src/kern/plane.h:16-35   padded fp32 Q6 plane (pixel x 64, integer valued), 32 lanes / 4 rows
 *                                             of replicated border, stride a multiple of 16 lanes
 *   struct lk_cfg    src/kern/lk.h:24-31      LK window, iteration limit, thresholds
 *   struct lk_out    src/kern/lk.h:33-40      32 B per point: estimate, min eigenvalue, residual, covariance,
 *                                             iterations summed over levels, status bits
 *
 * Each block is guarded by the reference header's own include guard, so a translation unit that already included
 * plane.h or lk.h from the reference sees one definition. */
#ifndef VTG_TYPES_H
#define VTG_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifndef VT_PLANE_H
#define VT_PLANE_H
#define VT_PAD_X 32
#define VT_PAD_Y 4
#define VT_Q6 64

struct vt_plane {
    float *p; /* pixel (0, 0), inside the padded block */
    int w, h;
    int stride; /* floats per row, a multiple of 16 */
};

static inline int plane_stride(int w) { return ((w + 15) & ~15) + 2 * VT_PAD_X; }
static inline size_t plane_bytes(int w, int h) { return (size_t)plane_stride(w) * (size_t)(h + 2 * VT_PAD_Y) * sizeof(float); }
static inline struct vt_plane plane_at(void *mem, int w, int h)
{
    const int s = plane_stride(w);
    struct vt_plane pl = {(float *)mem + (size_t)s * VT_PAD_Y + VT_PAD_X, w, h, s};
    return pl;
}
static inline float *plane_row(const struct vt_plane *pl, int y) { return pl->p + (ptrdiff_t)y * pl->stride; }
#endif

#ifndef VT_LK_H
#define VT_LK_H
#define LK_MAX_W 48
#define LK_MAX_H 31

enum { LK_OK = 0, LK_OUT = 1, LK_LOWEIG = 2, LK_RESIDUAL = 4, LK_FLOWBACK = 8 };

struct lk_cfg {
    int w, h;           /* window, pixels: 5 .. LK_MAX_W by 5 .. LK_MAX_H */
    int max_iter;       /* iterations per level */
    float eps;          /* stop when a step is below this, pixels of the level */
    float min_eig;      /* smaller eigenvalue of G per window pixel, 8-bit grey per pixel gradients */
    float max_residual; /* mean |T - J| in 8-bit grey */
    int prefetch;       /* CPU cache hint in the reference; no effect on results, ignored here */
};

struct lk_out {
    float x, y;                   /* estimate at level 0 */
    float min_eig;                /* at level 0 */
    float residual;               /* mean |T - J|, 8-bit grey */
    float cov_xx, cov_xy, cov_yy; /* s^2 G^-1, pixels^2 */
    uint16_t iters;               /* summed over the levels */
    uint16_t status;              /* LK_* bits */
};
#endif

/* Still-cell vectors (src/kern/still.h:19-25, src/common/vt_proto.h:385-416). Named vtg_* so a translation unit may
 * also include the reference's vt_proto.h; the layouts are the reference's, so pointers convert both ways. */
#ifndef VTG_STILL_TYPES
#define VTG_STILL_TYPES
#define VTG_CELL_W 16                /* a cell: 16 x 8 pixels of the level read (VT_CELL_W, VT_CELL_H) */
#define VTG_CELL_H 8
#define VTG_CELL_LK 0xffu            /* state: the low byte is LK's status, 0 = a good vector */
#define VTG_CELL_CHANGED 0x100u      /* this cell's own evidence said it moved */
#define VTG_CELL_MOVING 0x200u       /* it, or a neighbour, moved: its mean starts again */
#define VTG_CELL_COUNT_SHIFT 16      /* the high half: frames in its mean */

struct vtg_still_cfg { /* = struct still_cfg */
    int w, h;          /* of the level read */
    int avg;           /* frames at most in a cell's mean */
    float motion_px;   /* a good vector longer than this: moved */
    float motion_grey; /* mean |frame - mean| over a cell above this, 8-bit grey: moved */
    int threads;       /* the reference's team size; unused here */
};

struct vtg_still_cell { /* = struct vt_cell, 24 B */
    float u, v;       /* pixels of the level; meaningful when LK's status is 0 */
    float diff;       /* mean |frame - mean| over the cell, 8-bit grey */
    uint32_t state;   /* VTG_CELL_* */
    float sigma;      /* sqrt(cov_xx + cov_yy) of LK's own covariance (0 unless LK's status is 0) */
    float pad;
};

struct vtg_motion { /* = the leading fields of struct vt_motion, through frames */
    uint32_t cols, rows;
    uint32_t cell_at;   /* not filled (the reference's transport offset) */
    uint32_t tracked;   /* cells with a good vector */
    uint32_t moving;    /* cells that moved at the newest frame */
    float count_mean;   /* frames in a cell's mean */
    float drift_u, drift_v, drift_rms; /* mean vector (and scatter) over cells that did not move */
    float diff_median;  /* the camera's noise: the median cell difference, 8-bit grey */
    float diff_limit;   /* what a cell's difference had to exceed */
    float pad;
    uint64_t cycles_vectors, cycles_mean; /* not filled */
    uint64_t frames;
};

#define VTG_FF_STEREO 1u    /* a feature's flags (VT_FF_*, src/common/vt_proto.h:324-327): xr, yr, disparity hold a match */
#define VTG_FF_3D 2u        /* its 3D point passed every check */
#define VTG_FF_AMBIGUOUS 4u /* the scan's second peak was too close to its first */
#define VTG_FF_NEW 8u       /* detected at this frame (kept through the match) */

struct vtg_stereo_geom { long double fx, fy, cx, cy, baseline; }; /* = struct stereo_geom (stereo.h:22) */

struct vtg_stereo_cfg { /* the tracker's stereo settings (vt_proto.h:97-99, track.c:488-490) */
    int stereo_level;          /* the scan's level: the tracker takes the lowest with max_disp / 2^level <= 64 */
    float min_disp, max_disp;  /* pixels of level 0 */
    float min_ncc, ambiguity;  /* the scan's peak must reach min_ncc; its second peak must not exceed ambiguity x peak */
    float fb_max;              /* the flowback limit, pixels (0: no flowback) */
    float max_reproj;          /* pixels */
    struct lk_cfg lk;
    struct vtg_stereo_geom geom;
};

struct vtg_stereo_feature { /* one feature of the tracker's stereo step (track.c:1859-1914) */
    float x, y;         /* in: the feature in the left image, pixels of level 0 */
    float disparity;    /* in: a prior (> 0) or none (<= 0); out: x - xr, or -1 when there is no match */
    uint32_t flags;     /* in: VTG_FF_NEW is kept; out: VTG_FF_STEREO, VTG_FF_3D, VTG_FF_AMBIGUOUS */
    float xr, yr;       /* out: the match in the right image (0 when none) */
    float z;            /* out: the 3D point's depth (0 unless VTG_FF_3D) */
    float pad;
    long double p[3];   /* out: the triangulated point, written when triangulation succeeds */
};

struct vtg_stereo_scan { /* = struct stereo_scan (src/kern/stereo.h:12-16) */
    float disparity;   /* of the best peak, in pixels of the level; negative when no candidate fits */
    float ncc;         /* the best peak's correlation */
    float second;      /* the best correlation at least 3 candidates away from it */
};

/* The front door's message (src/common/vt_proto.h:143-197): the host's descriptor of one frame in one of its slots,
 * read from a ring of VTG_DESC_RING; same layout as struct vt_msg (88 bytes). */
#define VTG_MSG_MAGIC 0x4d54u /* "TM" */
#define VTG_DESC_RING 256
enum { VTG_MSG_NONE, VTG_MSG_FRAME, VTG_MSG_RESULT };
#define VTG_PIPE_PYR 1u
#define VTG_PIPE_FLOW 2u
#define VTG_PIPE_STEREO 4u
#define VTG_PIPE_EST 8u
enum { VTG_PRIO_ORDERED, VTG_PRIO_LIVE /* frames waiting ahead of this one are skipped */ };
enum { VTG_EXC_NONE, VTG_EXC_BAD_CHECK, VTG_EXC_BAD_TYPE, VTG_EXC_BAD_SLOT, VTG_EXC_BAD_LENGTH,
       VTG_EXC_INSTANCE_ORDER };  /* = enum vt_exception's ingest cases */

struct vtg_msg {
    uint16_t magic, type;       /* VTG_MSG_MAGIC, VTG_MSG_FRAME */
    uint16_t dest, priority;    /* VTG_PIPE_ mask (must hold VTG_PIPE_PYR), VTG_PRIO_ */
    uint32_t algorithm, length; /* the bytes of the frame in its slot */
    uint64_t instance;          /* the source's frame number: strictly increasing per camera */
    uint64_t acq_ns, pub_ns;    /* acquisition and publication times, host clock */
    uint32_t slot, cmd;         /* the slot holding the frame; a command (not acted on here) */
    float roi[4];
    uint32_t sc_key;
    float sc_value;
    uint32_t camera, generated; /* which camera: 0 .. cameras - 1 */
    uint64_t check;             /* vtg_msg_check of everything before it */
};
#endif

#endif
