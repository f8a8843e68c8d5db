// ncc.hip: zero-mean NCC template search on the GPU (search_at / tk_correlate, src/kern/track.c:184-208; the cell map
// OP_SEARCH, track.c:301-329; search_slot / search_half / peak_offset, track.c:1615-1726; the template,
// store_appearance, track.c:935-974).
//
// The reference scores a window as st / sqrt(norm * var) with window values taken relative to its first pixel
// v = p - p0, st = sum v t, var = sum v^2 - (sum v)^2 / n, in fp32 lane sums. At level 0 every pixel is 64 x a byte,
// so with u = bytes and t = fl(64 u_t - mean) the same sums are exact integers up to one known term:
//     st  = 4096 (sum u u_t - u0 sum u_t) - 64 mean (sum u - n u0)       [+ 64 sum (u - u0) (t - 64 u_t + mean)]
//     sum v = 64 (sum u - n u0),   sum v^2 = 4096 (sum u^2 - 2 u0 sum u + n u0^2)
// The bracketed term is the fp32 rounding of the stored template, e = t - 64 u_t + mean. It is 0 for every pixel
// when tw th is a power of two (then 64 u_t - mean fits in 24 bits). Otherwise the fast maps leave it out only where
// the template kernel's bound on what that moves a score, 4 * 64 * 255 sum |e| / norm, is at most 1e-6; a template
// above it (near-black ones reach 1.8e-5) is flagged and its search's map runs in the reference's arithmetic.
// The search of a 64-wide template computes sum u u_t on the int8 matrix cores (ncc_xcorr_cells: the cross term of
// every corner as a banded matrix product, exact i32), then, in the same workgroup, the box sums sum u, sum u^2 of
// every corner's window by sliding sums over the frame tile in LDS (down the columns, then along the rows by DPP
// prefix scans) and the score in fp64; the cross terms never leave the chip.
// Narrower templates (16, 32) take the dot4 kernel below, which computes sum u u_t with v_dot4_u32_u8 (4 byte
// products per lane per instruction, exact u32) beside the same box sums.
//
// Fast kernel layout: a workgroup owns 256 x 16 template corners = 16 cells of the map. Its frame tile (corners plus
// the template's extent) sits in LDS as 32-bit words. Wave s (0..3) takes corners x = X0 + 4 q + s for lane q, so all
// lanes of a wave share one byte alignment: the template is pre-shifted by that alignment (4 copies, tw/4 + 1 words per
// row) and read as uniform scalar loads, and each lane reads tw/4 + 1 aligned words per frame row. A lane keeps 8 frame
// rows in registers and slides them down, so each row read from LDS feeds 8 corners; two passes of 8 rows make 16.
// The products of one template row run word-outer over the 8 corners' accumulators (independent chains), and the next
// row's template words are requested before them.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "vtg_internal.hpp"

namespace vtg {

namespace {

constexpr int NX = 256, NY = 16, YB = 8;

template <typename T> struct Q6v;
template <> struct Q6v<uint8_t> {
    static __device__ __forceinline__ float get(uint8_t v) { return (float)v * 64.0f; }
};
template <> struct Q6v<uint16_t> {
    static __device__ __forceinline__ float get(uint16_t v) { return (float)v; }
};
template <> struct Q6v<float> {
    static __device__ __forceinline__ float get(float v) { return v; }
};
template <> struct Q6v<__half> {
    static __device__ __forceinline__ float get(__half v) { return __half2float(v) * 64.0f; }
};
template <> struct Q6v<__hip_bfloat16> {
    static __device__ __forceinline__ float get(__hip_bfloat16 v) { return __bfloat162float(v) * 64.0f; }
};

__device__ __forceinline__ double wave_sum_d(double v)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    return v;
}

// ---------------------------------------------------------------- template

template <typename T> struct Exact { static constexpr bool value = false; };
template <> struct Exact<uint8_t> { static constexpr bool value = true; };
template <> struct Exact<uint16_t> { static constexpr bool value = true; };
template <> struct Exact<float> { static constexpr bool value = true; };

// The reference's 16-lane horizontal sum: pairs i^1, i^2, i^4, i^8, every lane updated each step (vx.h:213-221).
__device__ __forceinline__ float tree16(const float *p)
{
    float a[16], b[16];
    for (int i = 0; i < 16; i++)
        a[i] = p[i];
    for (int st = 1; st < 16; st <<= 1) {
        for (int i = 0; i < 16; i++)
            b[i] = a[i] + a[i ^ st];
        for (int i = 0; i < 16; i++)
            a[i] = b[i];
    }
    return a[0];
}

// Pixels as Q6 floats; mean and norm by the reference's 16-lane order: lane j sums columns j, j+16, ... of every
// row in row order, then the tree (store_appearance, track.c:948-966). For integer-valued levels the lane sums are
// exact integers (at most 256 x 16,320 < 2^24), so they are formed in any order; the sum of squares of the zero-mean
// values is not, and runs as 16 sequential chains.
// One template of a batch: the level and corner it is cut from and where its records go.
struct TJob {
    Level lv;
    int x0, y0;
    float *t, *raw, *nm;
    long long *su;
    uint32_t *words; // level 0 (u8) only
    int *slow;       // the template's rounding term may move a score by more than 1e-6 (u8 levels)
};

// The template's rounding term. The matrix-core and dot4 maps score a corner from the identity in the header, which
// takes the stored template t_k as 64 u_k - mean exactly. It is fl(64 u_k - mean), off by e_k (|e_k| <= half an ulp
// of t_k); the score st / sqrt(norm var) then misses 64 sum (u - u0) e_k / sqrt(norm var), and since a corner is
// scored only where var > norm / 16 and |u - u0| <= 255, that is below 4 * 64 * 255 sum |e_k| / norm. When tw th is a
// power of two every e_k is 0 (the mean of multiples of 64 over a power of two has at most 6 + log2(tw th) fractional
// bits, so 64 u_k - mean fits in 24); otherwise the template kernel computes the bound and flags the template when it
// exceeds 1e-6, and the fast maps leave that search to ncc_cells_generic, the reference's arithmetic.
constexpr double ROUND_BOUND = 4.0 * 64.0 * 255.0;
constexpr double ROUND_LIMIT = 1e-6;

constexpr int TMAX = 8; // templates per launch

struct TBatch {
    TJob j[TMAX];
    int n, tw, th, wpr;
};

template <typename T>
__device__ __forceinline__ void template_one(Level lv, int x0, int y0, int tw, int th, float *t, float *raw,
                                             float *out_norm_mean, long long *out_sum_u, uint32_t *words, int wpr,
                                             int *out_slow)
{
    __shared__ float px[64 * 64];
    __shared__ int ipart[256];
    __shared__ float part[16];
    __shared__ float mean_s;
    __shared__ double epart[4];
    const int tid = threadIdx.x, n = tw * th;
    const T *base = (const T *)lv.p;
    {
        float v[16]; // n <= 64 x 64: every load issued before the first store
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const int i = tid + 256 * k, ic = i < n ? i : n - 1;
            const int r = ic / tw, kk = ic - r * tw;
            v[k] = Q6v<T>::get(base[(size_t)(y0 + r) * lv.pitch + x0 + kk]);
        }
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const int i = tid + 256 * k;
            if (i < n) {
                px[i] = v[k];
                raw[i] = v[k];
            }
        }
    }
    __syncthreads();
    if (words) {
        // byte b of word m of shift s, row r: template byte at column 4 m + b - s, 0 outside (level 0: Q6 = 64 u)
        for (int i = tid; i < 4 * th * wpr; i += 256) {
            const int sft = i / (th * wpr), rem = i - sft * th * wpr, r = rem / wpr, m = rem - r * wpr;
            uint32_t w = 0;
            for (int b = 0; b < 4; b++) {
                const int col = 4 * m + b - sft;
                if (col >= 0 && col < tw)
                    w |= (uint32_t)((int)px[r * tw + col] >> 6) << (8 * b);
            }
            words[i] = w;
        }
    }
    if constexpr (Exact<T>::value) {
        const int j = tid & 15, g = tid >> 4;
        int s = 0;
        for (int r = g; r < th; r += 16)
            for (int k = 0; k < tw; k += 16)
                s += (int)px[r * tw + k + j];
        ipart[tid] = s;
        __syncthreads();
        if (tid < 16) {
            int tot = 0;
            for (int q = 0; q < 16; q++)
                tot += ipart[q * 16 + tid];
            part[tid] = (float)tot;
        }
    } else {
        if (tid < 16) {
            float s = 0.0f;
            const int per = tw / 16;
            for (int r = 0; r < th; r += 4) {
                float z[16];
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const int rr = r + (q >> 2), kk = q & 3;
                    z[q] = (rr < th && kk < per) ? px[rr * tw + 16 * kk + tid] : 0.0f;
                }
#pragma unroll
                for (int q = 0; q < 16; q++)
                    if (r + (q >> 2) < th && (q & 3) < per)
                        s = s + z[q];
            }
            part[tid] = s;
        }
    }
    __syncthreads();
    if (tid == 0) {
        const float sum = tree16(part);
        mean_s = (float)((double)sum / (double)n);
        if (words) {
            long long tot = 0; // the lane sums are exact integers; their float tree need not be
            for (int q = 0; q < 16; q++)
                tot += (long long)part[q];
            *out_sum_u = tot >> 6;
        }
    }
    __syncthreads();
    const float mean = mean_s;
    double es = 0.0; // sum |e_k| of this thread's pixels
    for (int i = tid; i < n; i += 256) {
        const float z = px[i] - mean;
        es += fabs((double)z - ((double)px[i] - (double)mean));
        px[i] = z;
        t[i] = z;
    }
    es = wave_sum_d(es);
    if ((tid & 63) == 0)
        epart[tid >> 6] = es;
    __syncthreads();
    if (tid < 16) {
        // lane tid: columns tid, tid + 16, ... of each row in row order, a sequential fp32 chain; 4 rows of loads
        // are issued before their fused multiply-adds
        float s = 0.0f;
        const int per = tw / 16;
        for (int r = 0; r < th; r += 4) {
            float z[16];
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int rr = r + (q >> 2), kk = q & 3;
                z[q] = (rr < th && kk < per) ? px[rr * tw + 16 * kk + tid] : 0.0f;
            }
#pragma unroll
            for (int q = 0; q < 16; q++)
                if (r + (q >> 2) < th && (q & 3) < per)
                    s = fmaf(z[q], z[q], s);
        }
        part[tid] = s;
    }
    __syncthreads();
    if (tid == 0) {
        const float norm = tree16(part);
        out_norm_mean[0] = norm;
        out_norm_mean[1] = mean;
        const double e_sum = epart[0] + epart[1] + epart[2] + epart[3];
        *out_slow = words && ROUND_BOUND * e_sum > ROUND_LIMIT * (double)norm;
    }
}

// Grid: one workgroup per template of the batch (store_appearance for several slots or streams in one round).
template <typename T>
__global__ __launch_bounds__(256) void ncc_template_kernel(TBatch b)
{
    const TJob &j = b.j[blockIdx.x];
    template_one<T>(j.lv, j.x0, j.y0, b.tw, b.th, j.t, j.raw, j.nm, j.su, j.words, b.wpr, j.slow);
}

// ---------------------------------------------------------------- scores at given corners, the reference's arithmetic

// The score of one corner by 16 lanes that stand for the reference's 16 vector lanes: lane j sums columns j, j+16, ...
// of each row in row order with fp32 fused multiply-adds, the group sums with the same pair tree, and the finish is
// fp64 (the reference's long double) (search_at, track.c:184-206; search_half, track.c:1615-1647 with half set and
// (hx, hy) in half pixels). Four groups of a wave score four corners; every lane of the wave must call it.
template <typename T>
__device__ double score16(const Level &lv, const float *t, int tw, int th, double norm, int hx, int hy, bool half,
                          int lane)
{
    const int j = lane & 15;
    const int x = half ? (hx >> 1) : hx, y = half ? (hy >> 1) : hy;
    const int ox = half ? (hx & 1) : 0, oy = half ? (hy & 1) : 0;
    const int ext = half ? 1 : 0;
    const bool fits = !(x < 1 || y < 1 || x + tw + ext >= lv.w - 1 || y + th + ext >= lv.h - 1);
    float st = 0.0f, sj = 0.0f, sjj = 0.0f;
    if (fits) {
        const T *base = (const T *)lv.p;
        const size_t pitch = (size_t)lv.pitch;
        const float first = Q6v<T>::get(base[(size_t)y * pitch + x]);
        for (int r = 0; r < th; r++) {
            const T *row = base + (size_t)(y + r) * pitch + x + j;
            for (int k = 0; k < tw; k += 16) {
                float v = Q6v<T>::get(row[k]);
                if (ox)
                    v = (v + Q6v<T>::get(row[k + 1])) * 0.5f;
                if (oy) {
                    float w = Q6v<T>::get(row[k + pitch]);
                    if (ox)
                        w = (w + Q6v<T>::get(row[k + pitch + 1])) * 0.5f;
                    v = (v + w) * 0.5f;
                }
                v = v - first;
                st = fmaf(v, t[r * tw + k + j], st);
                sj = sj + v;
                sjj = fmaf(v, v, sjj);
            }
        }
    }
    for (int step = 1; step < 16; step <<= 1) {
        st = st + __shfl_xor(st, step);
        sj = sj + __shfl_xor(sj, step);
        sjj = sjj + __shfl_xor(sjj, step);
    }
    if (!fits)
        return -1.0;
    const double n = (double)(tw * th), a = st, b = sj, c = sjj;
    const double var = c - b * b / n;
    return (var <= norm * 0.0625 || norm <= 0.0) ? 0.0 : a / sqrt(norm * var);
}

template <typename T>
__global__ __launch_bounds__(256) void ncc_score_kernel(Level lv, const float *t, int tw, int th, const float *nm,
                                                        int n, const int *xy, int half, double *out)
{
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int first = (blockIdx.x * 4 + wave) * 4;
    if (first >= n)
        return;
    const int i = first + (lane >> 4);
    const bool valid = i < n;
    const double v = score16<T>(lv, t, tw, th, (double)nm[0], valid ? xy[2 * i] : -64, valid ? xy[2 * i + 1] : -64,
                                half != 0, lane);
    if (valid && (lane & 15) == 0)
        out[i] = v;
}

// ---------------------------------------------------------------- a batch of searches

struct Cand { float v; int idx; };

constexpr int RING_MAX = 9; // half-pixel grid of one candidate: 2 (around + 1) + 1, around <= 3

// What the finish kernels pass on: the top cells, then every half-pixel score around each.
struct Aux {
    int ntop, tops[8];
    int done; // refinement workgroups of the search finished (cleared by ncc_top8)
    int pad[6];
    double hs[8][RING_MAX * RING_MAX];
};

// One search of a batch: the level, the template, the clipped corner window and where its map and records go.
struct SJob {
    Level lv;
    const float *t;          // fp32 zero-mean template
    const float *nm;         // norm, mean
    const long long *sum_t;  // sum of the template's bytes (u8 levels)
    const uint32_t *words;   // the 4 pre-shifted byte copies (u8 levels)
    int x0, x1, y0, y1;      // clipped corner window
    int cols, rows;          // cells of the map; rows == 0: nothing to search
    vtg_cell *cells;
    vtg_found *found;
    Aux *aux;
    const int *slow;         // the template is flagged slow (its rounding term): the map is ncc_cells_generic's
};

constexpr int SMAX = 8; // searches per launch

// Searches launched together: one template size, one level element type, one step.
struct SBatch {
    SJob j[SMAX];
    int n, tw, th, step;
    int slow_only; // ncc_cells_generic: only the searches whose template is flagged slow
};

// ---------------------------------------------------------------- cell map, generic: one wave per cell

// Any level type, any template: the cell's corners in row-major order, four at a time, the first best kept by a
// strict > (OP_SEARCH, track.c:301-329), scores in the reference's arithmetic. Grid: (cells / 4, searches).
template <typename T>
__global__ __launch_bounds__(256) void ncc_cells_generic(SBatch b)
{
    const SJob &a = b.j[blockIdx.y];
    if (b.slow_only && !*a.slow)
        return; // the whole workgroup: this search's map is the fast kernel's
    const int ncell = a.cols * a.rows;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int cell = blockIdx.x * 4 + wave;
    if (cell >= ncell)
        return;
    const int step = b.step;
    const int cy0 = a.y0 + (cell / a.cols) * 16, cx0 = a.x0 + (cell % a.cols) * 16;
    const int cy1 = cy0 + 16 <= a.y1 + 1 ? cy0 + 16 : a.y1 + 1, cx1 = cx0 + 16 <= a.x1 + 1 ? cx0 + 16 : a.x1 + 1;
    const int nx = (cx1 - cx0 + step - 1) / step, ny = (cy1 - cy0 + step - 1) / step, np = nx * ny;
    const double norm = (double)a.nm[0];
    float top = -2.0f;
    int bx = cx0, by = cy0;
    for (int q0 = 0; q0 < np; q0 += 4) {
        const int p = q0 + (lane >> 4);
        const int x = p < np ? cx0 + (p % nx) * step : -64, y = p < np ? cy0 + (p / nx) * step : -64;
        const float v = (float)score16<T>(a.lv, a.t, b.tw, b.th, norm, x, y, false, lane);
        for (int q = 0; q < 4 && q0 + q < np; q++) {
            const float vq = __shfl(v, 16 * q);
            if (vq > top)
                top = vq, bx = cx0 + ((q0 + q) % nx) * step, by = cy0 + ((q0 + q) / nx) * step;
        }
    }
    if (lane == 0)
        a.cells[cell] = vtg_cell{top, bx, by};
}

// ---------------------------------------------------------------- cell map, fast: u8 level, dot4

__device__ __forceinline__ bool better(float v, int y, int x, float v2, int y2, int x2)
{
    return v > v2 || (v == v2 && (y < y2 || (y == y2 && x < x2)));
}

// The frame tile of a workgroup sits in LDS as 32-bit words: LR = NY + th - 1 rows of LW = NX / 4 + WPR + 1 words,
// word 0 of a row holding the 4 bytes from the aligned column A = X0 & ~3. After it: the vertical window sums V1, V2
// (sum of u and u^2 over th rows) of the pass's YB corner rows for each of the VB = NX + tw - 1 byte columns, then the
// box sums B1, B2 of the pass's YB x NX corners.
__host__ __device__ constexpr size_t fast_lds_words(int th, int tw, int wpr)
{
    return (size_t)(NY + th - 1) * (NX / 4 + wpr + 1) + (size_t)2 * YB * (NX + tw - 1) + (size_t)2 * YB * NX;
}

// Grid: (column blocks of 16 cells, cell rows, searches); blocks past a search's own map leave at once.
template <int WPR>
__global__ __launch_bounds__(256) void ncc_cells_fast(SBatch b)
{
    extern __shared__ uint32_t lds[];
    __shared__ float cv[4][16];
    __shared__ int cx[4][16], cyy[4][16];
    const SJob &a = b.j[blockIdx.z];
    if ((int)blockIdx.x * 16 >= a.cols || (int)blockIdx.y >= a.rows || *a.slow)
        return; // the whole workgroup (a slow template's map is ncc_cells_generic's): no barrier is pending
    const uint8_t *img = (const uint8_t *)a.lv.p;
    const int ih = a.lv.h, pitch = a.lv.pitch, tw = b.tw, th = b.th, stp = b.step;
    constexpr int LW = NX / 4 + WPR + 1;
    const int LR = NY + th - 1, VB = NX + tw - 1;
    uint32_t *V1 = lds + LR * LW, *V2 = V1 + YB * VB, *B1 = V2 + YB * VB, *B2 = B1 + YB * NX;
    const int X0 = a.x0 + blockIdx.x * NX, Y0 = a.y0 + blockIdx.y * NY;
    const int A = X0 & ~3, off = X0 - A;
    const int tid = threadIdx.x;
    for (int i = tid; i < LR * LW; i += 256) {
        const int r = i / LW, c = i - r * LW;
        const int gy = Y0 + r < ih ? Y0 + r : ih - 1;
        const int gx = A + 4 * c;
        uint32_t v = 0;
        if (gx + 3 < pitch)
            v = *(const uint32_t *)(img + (size_t)gy * pitch + gx);
        lds[i] = v;
    }
    __syncthreads();
    const uint8_t *tb = (const uint8_t *)lds; // the tile as bytes: row r, column b of the corners at r * 4 LW + off + b
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int sh = (off + wave) & 3, wb = (off + wave) >> 2;
    const uint32_t *T = a.words + (size_t)sh * th * WPR;
    const int x = X0 + 4 * lane + wave, c = 4 * lane + wave;
    const bool x_ok = x <= a.x1 && ((x - a.x0) % stp) == 0;
    const double n = (double)(tw * th), norm = (double)a.nm[0], mean = (double)a.nm[1];
    const double sum_t = (double)a.sum_t[0];
    float best_v = -3.0f;
    int best_x = 0, best_y = 0;
    const uint32_t *lbase = lds + lane + wb;
    // running vertical sums of byte columns tid and tid + 256, carried from pass to pass
    uint32_t vs1[2] = {0u, 0u}, vs2[2] = {0u, 0u};
    for (int sp = 0; sp < NY / YB; sp++) {
        // vertical sums of the pass's corner rows d = sp YB .. sp YB + YB - 1, sliding down one row at a time
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int bc = tid + 256 * q;
            if (bc < VB) {
                const uint8_t *col = tb + off + bc;
                uint32_t s1 = vs1[q], s2 = vs2[q];
                for (int d = 0; d < YB; d++) {
                    const int rr = sp * YB + d;
                    if (rr == 0) {
                        for (int r = 0; r < th; r++) {
                            const uint32_t u = col[r * 4 * LW];
                            s1 += u;
                            s2 += u * u;
                        }
                    } else {
                        const uint32_t u_in = col[(rr + th - 1) * 4 * LW], u_out = col[(rr - 1) * 4 * LW];
                        s1 += u_in - u_out;
                        s2 += u_in * u_in - u_out * u_out;
                    }
                    V1[d * VB + bc] = s1;
                    V2[d * VB + bc] = s2;
                }
                vs1[q] = s1;
                vs2[q] = s2;
            }
        }
        __syncthreads();
        {
            // box sums: task tid = (quantity, corner row, 16-corner segment), sliding right
            const int row = tid >> 4, seg = tid & 15, d = row & (YB - 1);
            const uint32_t *src = (row < YB ? V1 : V2) + d * VB;
            uint32_t *dst = (row < YB ? B1 : B2) + d * NX;
            const int c0 = seg * 16;
            uint32_t s = 0;
            for (int k = 0; k < tw; k++)
                s += src[c0 + k];
            dst[c0] = s;
            for (int k = 1; k < 16; k++) {
                s += src[c0 + k + tw - 1] - src[c0 + k - 1];
                dst[c0 + k] = s;
            }
        }
        __syncthreads();
        uint32_t F[YB][WPR];
        uint32_t P[YB];
#pragma unroll
        for (int d = 0; d < YB; d++)
            P[d] = 0;
        // rows 0 .. YB - 2 of the pass into the ring
#pragma unroll
        for (int j = 0; j < YB - 1; j++) {
            const uint32_t *src = lbase + (sp * YB + j) * LW;
#pragma unroll
            for (int m = 0; m < WPR; m++)
                F[j][m] = src[m];
        }
        uint32_t tr[WPR];
#pragma unroll
        for (int m = 0; m < WPR; m++)
            tr[m] = T[m];
        // template row r: the ring holds frame rows r .. r + YB - 2 and takes row r + YB - 1; the next template row's
        // words are requested before this row's products, so their latency hides behind them
        auto step = [&](const int r, const int j) __attribute__((always_inline)) {
            const int slot = (j + YB - 1) % YB;
            const uint32_t *src = lbase + (sp * YB + r + YB - 1) * LW;
#pragma unroll
            for (int m = 0; m < WPR; m++)
                F[slot][m] = src[m];
            const int rn = r + 1 < th ? r + 1 : r;
            uint32_t tn[WPR];
#pragma unroll
            for (int m = 0; m < WPR; m++)
                tn[m] = T[rn * WPR + m];
#pragma unroll
            for (int m = 0; m < WPR; m++) {
#pragma unroll
                for (int d = 0; d < YB; d++)
                    P[d] = __builtin_amdgcn_udot4(F[(j + d) % YB][m], tr[m], P[d], false);
            }
#pragma unroll
            for (int m = 0; m < WPR; m++)
                tr[m] = tn[m];
        };
        for (int r = 0; r < th; r += YB) {
#pragma unroll
            for (int j = 0; j < YB; j++)
                step(r + j, j);
        }
        // scores of the pass's YB corners of this lane
#pragma unroll
        for (int d = 0; d < YB; d++) {
            const int y = Y0 + sp * YB + d;
            const int u0 = tb[(sp * YB + d) * 4 * LW + off + c];
            const double su = (double)B1[d * NX + c], sq = (double)B2[d * NX + c];
            const double pc = (double)P[d] - (double)u0 * sum_t;
            const double sv = su - n * (double)u0;
            const double st = 4096.0 * pc - 64.0 * mean * sv;
            const double sj = 64.0 * sv;
            const double sjj = 4096.0 * (sq - 2.0 * (double)u0 * su + n * (double)u0 * (double)u0);
            const double var = sjj - sj * sj / n;
            const double sc = (var <= norm * 0.0625 || norm <= 0.0) ? 0.0 : st / sqrt(norm * var);
            const float v = (float)sc;
            const bool ok = x_ok && y <= a.y1 && ((y - a.y0) % stp) == 0;
            if (ok && better(v, y, x, best_v, best_y, best_x))
                best_v = v, best_x = x, best_y = y;
        }
    }
    // the cell of this lane: lanes 4c .. 4c+3 of every wave; reduce the quad, then the 4 waves
#pragma unroll
    for (int o = 1; o < 4; o <<= 1) {
        const float v2 = __shfl_xor(best_v, o);
        const int x2 = __shfl_xor(best_x, o), y2 = __shfl_xor(best_y, o);
        if (better(v2, y2, x2, best_v, best_y, best_x))
            best_v = v2, best_x = x2, best_y = y2;
    }
    if ((lane & 3) == 0) {
        cv[wave][lane >> 2] = best_v;
        cx[wave][lane >> 2] = best_x;
        cyy[wave][lane >> 2] = best_y;
    }
    __syncthreads();
    if (tid < 16) {
        const int cc = tid;
        float v = cv[0][cc];
        int bx = cx[0][cc], by = cyy[0][cc];
        for (int w = 1; w < 4; w++)
            if (better(cv[w][cc], cyy[w][cc], cx[w][cc], v, by, bx))
                v = cv[w][cc], bx = cx[w][cc], by = cyy[w][cc];
        const int col = blockIdx.x * 16 + cc;
        if (col < a.cols)
            a.cells[(size_t)blockIdx.y * a.cols + col] = vtg_cell{v, bx, by};
    }
}

// ---------------------------------------------------------------- cell map, MFMA: u8 level, 64-wide template

// The cross term of a 64-wide template as a banded matrix product on the int8 matrix cores. For a 16 x 16 tile of
// corners (x across, y down) and image row R = y_t + d of the tile's reach (d = 0 .. th + 14):
//     P[x][y] += sum_k a[R][x + k] * b[d - (y - y_t)][k]          (k = 0 .. 63; b rows outside 0 .. th-1 are zero)
// i.e. one v_mfma_i32_16x16x64_i8 with A = 16 x 64 bytes of image row R sliding in x and B = 64 x 16 template bytes,
// column y holding template row d - y. Both operands are biased to signed bytes (a = u - 128, b = u_t - 128), which
// the scoring undoes exactly: sum u u_t = sum a b + 128 (sum u + sum u_t) - 16384 n. 64 of the th + 15 bands
// of a tile are full, so th / (th + 15) of the matrix work is useful (81 % at th = 64).
// A lane's A operand (row m = lane % 16, bytes 16 (lane / 16) .. + 15 of the window) is 16 unaligned bytes of an LDS
// row: 5 aligned words funnel-shifted by the lane's byte offset. Its B operand (column n = lane % 16, the same 16 k)
// is one aligned 16-byte read of a template row. The result tile lands as lane n + 16 q holding corners x = 4 q .. 4 q
// + 3 of row y = n, stored as one 16-byte write per tile. Each wave owns 4 x 4 tiles (64 x 64 corners), so every
// A operand feeds up to 4 products and every B operand 4.
typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int XT = 4, YT = 4;           // tiles per wave, across and down
constexpr int XB = 2 * 16 * XT;         // corners of a workgroup (2 x 2 waves) across
constexpr int YBK = 2 * 16 * YT;        // and down
constexpr int XS = XB + 80;             // LDS image row stride in bytes: the block, the template's reach, alignment
constexpr int TS = 80;                  // LDS template row stride in bytes (64 + 16 spreads a wave's rows over banks)
constexpr int XC_LOADS = ((YBK + 63) * (XS / 4) + 255) / 256; // tile words per thread at th = 64

// Templates shorter than 33 rows run as 33 rows whose extra rows are zero, so the band of every tile is at least 48
// rows and the phases of the product loop come in one order.
__host__ __device__ constexpr int xcorr_th(int th) { return th < 33 ? 33 : th; }

// Rows r0 .. r1 - 1 of a phase in which y-tiles TLO .. THI are active: each row's A operands (XT x-tiles) and, per
// active tile, its B operand and XT products.
template <int TLO, int THI>
__device__ __forceinline__ void xcorr_phase(int r0, int r1, const uint32_t *arow, const char *brow, int sh,
                                            v4i (&acc)[YT][XT])
{
    constexpr int XW = XS / 4;
#pragma unroll 2
    for (int R = r0; R < r1; R++) {
        v4i af[XT];
#pragma unroll
        for (int mx = 0; mx < XT; mx++) {
            const uint32_t *p = arow + R * XW + 4 * mx;
            const uint32_t w0 = p[0], w1 = p[1], w2 = p[2], w3 = p[3], w4 = p[4];
            af[mx] = v4i{(int)__builtin_amdgcn_alignbyte(w1, w0, sh), (int)__builtin_amdgcn_alignbyte(w2, w1, sh),
                         (int)__builtin_amdgcn_alignbyte(w3, w2, sh), (int)__builtin_amdgcn_alignbyte(w4, w3, sh)};
        }
#pragma unroll
        for (int t = TLO; t <= THI; t++) {
            const v4i bf = *(const v4i *)(brow + (R - 16 * t) * TS);
#pragma unroll
            for (int mx = 0; mx < XT; mx++)
                acc[t][mx] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[mx], bf, acc[t][mx], 0, 0, 0);
        }
    }
}

// The cross terms of the workgroup's 128 x 128 corners (block origin bx0, by0 in map corners) into acc: the image tile
// and the template into LDS, then the product loop. Returns the tile's byte offset of corner column 0 (X0 - A). The
// tile stays in LDS: rows Y0 .. Y0 + YBK + xcorr_th(th) - 2 from column A, XS bytes a row, biased to signed bytes.
__device__ __forceinline__ int xcorr_block(const SJob &a, int th, int bx0, int by0, uint32_t *lds, v4i (&acc)[YT][XT])
{
    const uint8_t *img = (const uint8_t *)a.lv.p;
    const int ih = a.lv.h, pitch = a.lv.pitch, the = xcorr_th(th);
    const int X0 = a.x0 + bx0, Y0 = a.y0 + by0;
    const int A = X0 & ~3, off = X0 - A;
    constexpr int XW = XS / 4, TW = TS / 4;
    const int TR = YBK + the - 1;
    uint32_t *tile = lds, *tpl = lds + TR * XW;
    const int tid = threadIdx.x;
    {
        // the image rows Y0 .. Y0 + TR - 1 from column A, biased to signed bytes; rows past the level repeat its
        // last row and words past the pitch read 0 (only corners outside the search window see either)
        constexpr int U = XC_LOADS; // every load of the tile in flight at once
        const int nw = TR * XW;
        uint32_t v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = tid + 256 * u;
            const int r = i / XW, c = i - r * XW;
            const int gy = Y0 + r < ih ? Y0 + r : ih - 1;
            const int gx = A + 4 * c;
            v[u] = (i < nw && gx + 3 < pitch) ? *(const uint32_t *)(img + (size_t)gy * pitch + gx) : 0x80808080u;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = tid + 256 * u;
            if (i < nw)
                tile[i] = v[u] ^ 0x80808080u;
        }
        // the template, rows -15 .. the + 14 (zero outside 0 .. th - 1), biased; from the unshifted byte copy
        const int ntw = (the + 30) * TW, wpr = 17;
        for (int i = tid; i < ntw; i += 256) {
            const int j = i / TW, m = i - j * TW, r = j - 15;
            tpl[i] = (r >= 0 && r < th && m < 16) ? a.words[r * wpr + m] ^ 0x80808080u : 0u;
        }
    }
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int wx = wave & 1, wy = wave >> 1;
    const int n16 = lane & 15, g = lane >> 4;
    const int o0 = off + 64 * wx + n16 + 16 * g; // byte of this lane's x-tile 0 operand in a tile row
    const int sh = o0 & 3;
    const uint32_t *arow = tile + (64 * wy) * XW + (o0 >> 2);
    const char *brow = (const char *)tpl + (15 - n16) * TS + 16 * g; // + d TS: template row d - n16 (+15)
#pragma unroll
    for (int t = 0; t < YT; t++)
#pragma unroll
        for (int mx = 0; mx < XT; mx++)
            acc[t][mx] = v4i{0, 0, 0, 0};
    // y-tile t takes rows R = 16 t .. 16 t + band - 1 (band = th_eff + 15 >= 48): seven phases, each a loop over a
    // fixed run of active tiles, in the same order for every template height
    const int band = the + 15;
    xcorr_phase<0, 0>(0, 16, arow, brow, sh, acc);
    xcorr_phase<0, 1>(16, 32, arow, brow, sh, acc);
    xcorr_phase<0, 2>(32, 48, arow, brow, sh, acc);
    xcorr_phase<0, 3>(48, band, arow, brow, sh, acc);
    xcorr_phase<1, 3>(band, band + 16, arow, brow, sh, acc);
    xcorr_phase<2, 3>(band + 16, band + 32, arow, brow, sh, acc);
    xcorr_phase<3, 3>(band + 32, band + 48, arow, brow, sh, acc);
    return off;
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_addu(uint32_t v)
{
    // lanes whose source lies outside the row, and rows the mask leaves out, add 0
    return v + (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, false);
}

// Inclusive prefix sum over the 64 lanes (exact integers), all in DPP: shifts by 1, 2, 4, 8 within each row of 16, then
// row 0's total into row 1 and row 2's into row 3 (row_bcast:15), then rows 0-1's into rows 2 and 3 (row_bcast:31).
// Every lane must be active.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    v = dpp_addu<0x111, 0xf>(v); // row_shr:1
    v = dpp_addu<0x112, 0xf>(v); // row_shr:2
    v = dpp_addu<0x114, 0xf>(v); // row_shr:4
    v = dpp_addu<0x118, 0xf>(v); // row_shr:8
    v = dpp_addu<0x142, 0xa>(v); // row_bcast:15
    v = dpp_addu<0x143, 0xc>(v); // row_bcast:31
    return v;
}

// The cell map of a 64-wide template in one pass: the cross terms of the workgroup's 128 x 128 corners on the matrix
// cores (xcorr_block), then every corner's box sums and score in the same workgroup, so the cross terms never leave
// the chip. After one barrier (the template's space becomes the stashes) each wave works alone on the 64 x 64 corners
// whose cross terms it holds: lane l owns corner column l of them and, for its box sums, byte columns l and l + 64 of
// the wave's 127. Its box (columns l .. l + 63) is the first half's total plus the sum over the columns k < l of
// (column l + 64's sum - column k's): the lane slides that difference of its two column sums (sum u, sum u^2 over th
// rows) down the frame tile in LDS, one DPP prefix scan per sum and corner row gives the sum over k < l, and the
// first half's total slides down the row sums of the wave's first 64 byte columns, which the workgroup computes once
// per tile row (v_dot4_u32_u8); no workgroup barrier. The wave moves its accumulators to its
// own stash 8 corner rows at a time, so lane l reads corner column l's cross term. With su, sq the box sums, u0 the
// window's first byte and n = tw th, the score is the header's identity on exact integers: sum u u_t = cross + 128 (su
// + sum u_t) - 16384 n, st = 4096 (sum u u_t - u0 sum u_t) - 64 mean (su - n u0), var = 4096 (sq - 2 u0 su + n u0^2)
// - (64 (su - n u0))^2 / n; st / 64 and var n / 4096 are exact in fp64, and the score st / sqrt(norm var), 0 where
// var <= norm / 16, takes one rounded quotient by n (exact when POW2: n a power of two) and a reciprocal square root
// refined by two Newton steps (relative error near 1e-16, far inside the float the map keeps). Each 16 corner rows
// close a row of the wave's 4 cells, the first
// best corner of each kept (OP_SEARCH, track.c:301-329). Grid: (column blocks of XB corners, row blocks of YBK
// corners, searches) over each search's padded map (16 cols x 16 rows corners); blocks past a search's own map leave
// at once.
constexpr int XC_SR = 8;             // corner rows a wave stashes at a time
constexpr int XC_SS = 64 + 4;        // stash row stride in words: 64 corners and a pad
constexpr int XC_SW = XC_SR * XC_SS; // stash words per wave

__host__ __device__ constexpr size_t xcells_rs_offset(int th)
{
    // the tile, then the larger of the template and the four waves' stashes
    return (size_t)(YBK + xcorr_th(th) - 1) * XS +
           ((size_t)(xcorr_th(th) + 30) * TS > (size_t)4 * XC_SW * 4 ? (size_t)(xcorr_th(th) + 30) * TS
                                                                      : (size_t)4 * XC_SW * 4);
}

__host__ __device__ constexpr size_t xcells_lds_bytes(int th)
{
    // then the row sums: per tile row and wave column, sum u and sum u^2 of the wave's first 64 byte columns
    return xcells_rs_offset(th) + (size_t)(YBK + xcorr_th(th) - 1) * 2 * 2 * 4;
}

// LDS written by some lanes of a wave, then read by others of the same wave: the wave's LDS operations complete in
// order, and the fences keep the compiler from moving them across.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool POW2>
__global__ __launch_bounds__(256) void ncc_xcorr_cells(SBatch b)
{
    extern __shared__ uint32_t lds[];
    const SJob &a = b.j[blockIdx.z];
    const int pcols = 16 * a.cols, prows = 16 * a.rows;
    const int bx0 = blockIdx.x * XB, by0 = blockIdx.y * YBK; // block origin in map corners
    if (bx0 >= pcols || by0 >= prows || *a.slow)
        return; // the whole workgroup (a slow template's map is ncc_cells_generic's): no barrier is pending
    const int th = b.th, stp = b.step;
    v4i acc[YT][XT];
    const int off = xcorr_block(a, th, bx0, by0, lds, acc);
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int wx = wave & 1, wy = wave >> 1, n16 = lane & 15, g = lane >> 4;
    const int TR = YBK + xcorr_th(th) - 1;
    int *st = (int *)(lds + TR * (XS / 4)) + wave * XC_SW; // this wave's stash
    // byte columns 64 wx + lane (first half) and 64 wx + 64 + lane (second half) of the block's corner space
    const uint8_t *ca = (const uint8_t *)lds + off + 64 * wx + lane, *cb = ca + 64;
    const int yw = 64 * wy; // the wave's first corner row in the block
    const int n = 64 * th, sum_t = (int)a.sum_t[0];
    const double norm = (double)a.nm[0], mean = (double)a.nm[1];
    const double nn = norm * (double)n;      // exact: 24 and 13 significant bits
    const double inv_n = 1.0 / (double)n;    // exact when POW2
    const int cx = bx0 + 64 * wx + lane, x = a.x0 + cx; // this lane's corner column in the map, and in the level
    const bool x_ok = cx < pcols && x <= a.x1 && ((x - a.x0) & (stp - 1)) == 0; // step 1 or 2
    const int ay0 = a.y0, ay1 = a.y1;
    __syncthreads(); // every wave is done with the template: its space takes the stashes
    // rs[2 (2 r + c) + q]: tile row r's sum u (q = 0) and sum u^2 (q = 1) over byte columns off + 64 c .. + 63
    uint32_t *rs = (uint32_t *)((char *)lds + xcells_rs_offset(th));
    for (int i = tid; i < 2 * TR; i += 256) {
        const int r = i >> 1, o = off + 64 * (i & 1), sh = 8 * (o & 3);
        const uint32_t *w = (const uint32_t *)((const uint8_t *)lds + r * XS) + (o >> 2);
        uint32_t q1 = 0, q2 = 0;
#pragma unroll
        for (int k = 0; k <= 16; k++) {
            // 17 words: the first without its bytes before o, the last with only its bytes before o + 64 (none when
            // o is word-aligned)
            const uint32_t m = k == 0 ? 0xffffffffu << sh : k == 16 ? (sh ? (1u << sh) - 1u : 0u) : 0xffffffffu;
            const uint32_t u = (w[k] ^ 0x80808080u) & m;
            q1 = __builtin_amdgcn_udot4(u, 0x01010101u, q1, false);
            q2 = __builtin_amdgcn_udot4(u, u, q2, false);
        }
        rs[2 * i] = q1;
        rs[2 * i + 1] = q2;
    }
    __syncthreads();
    const uint32_t *rw = rs + 2 * wx; // this wave's row sums: rw[4 r], rw[4 r + 1]
    // d1, d2: the lane's second column's sums minus its first column's; t1, t2: the first half's totals
    uint32_t d1 = 0, d2 = 0, t1 = 0, t2 = 0;
    for (int r = 0; r < th; r++) {
        const uint32_t ua = ca[(yw + r) * XS] ^ 0x80u, ub = cb[(yw + r) * XS] ^ 0x80u;
        d1 += ub - ua;
        d2 += ub * ub - ua * ua;
        t1 += rw[4 * (yw + r)];
        t2 += rw[4 * (yw + r) + 1];
    }
    float best_v = -3.0f;
    int best_x = 0, best_y = 0;
#pragma unroll
    for (int h = 0; h < 8; h++) {
        // the wave's corner rows 8 h .. 8 h + 7: y-tile h / 2, held by its lanes with n16 / 8 == h % 2
        wave_lds_sync(); // the last 8 rows' stash reads are done
        if ((n16 >> 3) == (h & 1))
#pragma unroll
            for (int mx = 0; mx < XT; mx++)
                *(v4i *)(st + (n16 & 7) * XC_SS + 16 * mx + 4 * g) = acc[h >> 1][mx];
        wave_lds_sync();
        for (int rr = 0; rr < XC_SR; rr++) {
            const int yb = yw + 8 * h + rr; // corner row in the block
            if (h > 0 || rr > 0) {
                const int ri = (yb + th - 1) * XS, ro = (yb - 1) * XS;
                const uint32_t ia = ca[ri] ^ 0x80u, oa = ca[ro] ^ 0x80u, ib = cb[ri] ^ 0x80u, ob = cb[ro] ^ 0x80u;
                d1 += (ib - ob) - (ia - oa);
                d2 += (ib * ib - ob * ob) - (ia * ia - oa * oa);
                t1 += rw[4 * (yb + th - 1)] - rw[4 * (yb - 1)];
                t2 += rw[4 * (yb + th - 1) + 1] - rw[4 * (yb - 1) + 1];
            }
            // box of corner column lane: first-half columns lane .. 63, second-half columns 0 .. lane - 1, i.e. the
            // first half's total plus the exclusive prefix of the differences (the arithmetic wraps consistently)
            const uint32_t box1 = t1 + (wave_incl_scan(d1) - d1);
            const uint32_t box2 = t2 + (wave_incl_scan(d2) - d2);
            const int u0 = ca[yb * XS] ^ 0x80;
            const int y = a.y0 + by0 + yb;
            // exact integers (every true value fits in 31 bits; the 32-bit arithmetic wraps consistently):
            // pc = sum (u - u0) u_t, s1 = sum (u - u0), s2 = sum (u - u0)^2
            const int pc = st[rr * XC_SS + lane] + 128 * ((int)box1 + sum_t) - 16384 * n - u0 * sum_t;
            const int s1 = (int)box1 - n * u0;
            const int s2 = (int)(box2 - 2u * (uint32_t)u0 * box1 + (uint32_t)(n * u0 * u0));
            // st = 64 b and var = 4096 an / n with b = 64 pc - mean s1 and an = n s2 - s1^2, both exact in fp64; the
            // score st / sqrt(norm var) = b / sqrt(norm an / n), scored where var > norm / 16, i.e. 65536 an > norm n
            const double ds1 = (double)s1;
            const double b = fma(-mean, ds1, 64.0 * (double)pc);
            const double an = fma((double)n, (double)s2, -(ds1 * ds1));
            const double w = POW2 ? norm * an * inv_n : norm * an / (double)n;
            double r = __builtin_amdgcn_rsq(w); // refined twice by Newton's step r (3 - w r^2) / 2
            r = fma(r, fma(-0.5 * w * r, r, 0.5), r);
            r = fma(r, fma(-0.5 * w * r, r, 0.5), r);
            const float v = (65536.0 * an <= nn || norm <= 0.0) ? 0.0f : (float)(b * r);
            // the lane's corners come in increasing y, so an earlier one wins a tie: better() reduces to a higher
            // score here (bitwise, so no branch; the window's bounds read once, above)
            const bool ok = x_ok & (y <= ay1) & (((y - ay0) & (stp - 1)) == 0);
            const bool take = ok & (v > best_v);
            best_v = take ? v : best_v;
            best_x = take ? x : best_x;
            best_y = take ? y : best_y;
        }
        if (h & 1) {
            // a row of the wave's 4 cells: lanes 16 c .. 16 c + 15 hold cell column c
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                const float v2 = __shfl_xor(best_v, o);
                const int x2 = __shfl_xor(best_x, o), y2 = __shfl_xor(best_y, o);
                if (better(v2, y2, x2, best_v, best_y, best_x))
                    best_v = v2, best_x = x2, best_y = y2;
            }
            const int ccol = (bx0 >> 4) + 4 * wx + (lane >> 4), crow = (by0 >> 4) + 4 * wy + (h >> 1);
            if ((lane & 15) == 0 && ccol < a.cols && crow < a.rows)
                a.cells[(size_t)crow * a.cols + ccol] = vtg_cell{best_v, best_x, best_y};
            best_v = -3.0f;
            best_x = best_y = 0;
        }
    }
}

// ---------------------------------------------------------------- finish: top 8, half-pixel refinement, peaks

__device__ __forceinline__ bool before(const Cand &a, const Cand &b)
{
    return a.v > b.v || (a.v == b.v && a.idx < b.idx);
}

// A cell's rank key: smaller is better. High word: the score, order-inverted (higher score, smaller key); low word:
// the cell index (earlier cell, smaller key): exactly the order search_slot's insertion keeps (track.c:1686-1694).
__device__ __forceinline__ unsigned long long cell_key(float v, int idx)
{
    const unsigned int b = __float_as_uint(v);
    const unsigned int ord = (b & 0x80000000u) ? ~b : (b | 0x80000000u); // increasing with v
    return ((unsigned long long)(~ord) << 32) | (unsigned int)idx;
}

__device__ __forceinline__ unsigned long long min64(unsigned long long a, unsigned long long b) { return a < b ? a : b; }

constexpr int TOP_REG = 32; // keys per thread held in registers: maps of up to 32,768 cells are read once

// One step of a wave-wide minimum of 64-bit keys: both halves moved by the same DPP control (all ones, the identity,
// where the source lies outside the row or the row mask leaves the row out).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned long long dpp_min64(unsigned long long m)
{
    const unsigned int lo = (unsigned int)__builtin_amdgcn_update_dpp(-1, (int)(unsigned int)m, CTRL, ROW_MASK, 0xf, false);
    const unsigned int hi =
        (unsigned int)__builtin_amdgcn_update_dpp(-1, (int)(unsigned int)(m >> 32), CTRL, ROW_MASK, 0xf, false);
    return min64(m, ((unsigned long long)hi << 32) | lo);
}

// The smallest key of the wave, the same in every lane: a butterfly within each row of 16 (after each step the lanes
// of a group hold one value, so the mirrors act as exchanges), then the rows joined into lane 63. Every lane must be
// active.
__device__ __forceinline__ unsigned long long wave_min64(unsigned long long m)
{
    m = dpp_min64<0xb1, 0xf>(m);  // quad_perm [1,0,3,2]
    m = dpp_min64<0x4e, 0xf>(m);  // quad_perm [2,3,0,1]
    m = dpp_min64<0x141, 0xf>(m); // row_half_mirror
    m = dpp_min64<0x140, 0xf>(m); // row_mirror
    m = dpp_min64<0x142, 0xa>(m); // row_bcast:15 into rows 1 and 3
    m = dpp_min64<0x143, 0xc>(m); // row_bcast:31 into rows 2 and 3: lane 63 holds the minimum
    const unsigned int lo = (unsigned int)__builtin_amdgcn_readlane((int)(unsigned int)m, 63);
    const unsigned int hi = (unsigned int)__builtin_amdgcn_readlane((int)(unsigned int)(m >> 32), 63);
    return ((unsigned long long)hi << 32) | lo;
}

// The 8 best cells, scores above -2 only: 8 rounds of a workgroup-wide minimum of the keys above the last one taken.
// Keys are unique (the cell index is their low word), so each round's minimum belongs to one thread; with the map in
// registers each thread keeps its smallest key not yet taken and only the owner of the key just taken looks again.
// Each round: a DPP minimum per wave, the 16 waves' minima through LDS (double-buffered by round, one barrier), every
// thread taking the minimum of the 16. Grid: one workgroup per search.
__global__ __launch_bounds__(1024) void ncc_top8(SBatch b)
{
    __shared__ unsigned long long wmin[2][16];
    const SJob &a = b.j[blockIdx.x];
    const vtg_cell *cells = a.cells;
    const int ncell = a.cols * a.rows;
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool inreg = ncell <= 1024 * TOP_REG;
    unsigned long long key[TOP_REG];
    unsigned long long cur = ~0ull; // with the map in registers: this thread's smallest key not yet taken
    if (inreg) {
        float v[TOP_REG];
#pragma unroll
        for (int q = 0; q < TOP_REG; q++) {
            const int i = tid + 1024 * q;
            v[q] = i < ncell ? cells[i].v : -3.0f;
        }
#pragma unroll
        for (int q = 0; q < TOP_REG; q++) {
            key[q] = v[q] > -2.0f ? cell_key(v[q], tid + 1024 * q) : ~0ull;
            cur = min64(cur, key[q]);
        }
    }
    unsigned long long prev = 0;
    int n = 0;
    for (int round = 0; round < 8; round++) {
        unsigned long long m = ~0ull;
        if (inreg) {
            m = cur;
        } else {
            for (int i = tid; i < ncell; i += 1024) {
                const float v = cells[i].v;
                const unsigned long long k = cell_key(v, i);
                if (v > -2.0f && (round == 0 || k > prev))
                    m = min64(m, k);
            }
        }
        m = wave_min64(m);
        if (lane == 0)
            wmin[round & 1][wave] = m;
        __syncthreads(); // the other buffer is rewritten only after the next round's barrier
        unsigned long long t = wmin[round & 1][0];
#pragma unroll
        for (int w = 1; w < 16; w++)
            t = min64(t, wmin[round & 1][w]);
        prev = t;
        if (prev == ~0ull)
            break;
        if (tid == 0)
            a.aux->tops[round] = (int)(unsigned int)prev;
        n++;
        if (inreg && cur == prev) {
            // this thread's key was taken: its next smallest
            unsigned long long c2 = ~0ull;
#pragma unroll
            for (int q = 0; q < TOP_REG; q++)
                c2 = key[q] > prev ? min64(c2, key[q]) : c2;
            cur = c2;
        }
    }
    if (tid == 0) {
        a.aux->ntop = n;
        a.aux->done = 0;
    }
}

// The score at half-pixel corner (hx, hy) from a staged region: pixel (x, y) of the level at reg[(y - ry) * rw + x - rx],
// template (16 TWQ wide) in LDS too; same arithmetic as score16. The half-pixel averages are selected rather than
// branched on, so the four corners of a wave do not diverge; every load of a row is issued before its sums.
template <int TWQ>
__device__ double score16_lds(const float *reg, int rx, int ry, int rw, const float *t, int th, double norm, int lw,
                              int lh, int hx, int hy, int lane)
{
    constexpr int tw = 16 * TWQ;
    const int j = lane & 15;
    const int x = hx >> 1, y = hy >> 1;
    const bool ox = hx & 1, oy = hy & 1;
    const bool fits = !(x < 1 || y < 1 || x + tw + 1 >= lw - 1 || y + th + 1 >= lh - 1);
    float st = 0.0f, sj = 0.0f, sjj = 0.0f;
    if (fits) {
        const float *p = reg + (y - ry) * rw + (x - rx);
        const float first = p[0];
        p += j;
        const float *tp = t + j;
#pragma unroll 2
        for (int r = 0; r < th; r++) {
            float a0[TWQ], a1[TWQ], b0[TWQ], b1[TWQ], tt[TWQ];
#pragma unroll
            for (int q = 0; q < TWQ; q++) {
                a0[q] = p[16 * q];
                a1[q] = p[16 * q + 1];
                b0[q] = p[16 * q + rw];
                b1[q] = p[16 * q + rw + 1];
                tt[q] = tp[16 * q];
            }
#pragma unroll
            for (int q = 0; q < TWQ; q++) {
                float v = ox ? (a0[q] + a1[q]) * 0.5f : a0[q];
                const float w = ox ? (b0[q] + b1[q]) * 0.5f : b0[q];
                v = oy ? (v + w) * 0.5f : v;
                v = v - first;
                st = fmaf(v, tt[q], st);
                sj = sj + v;
                sjj = fmaf(v, v, sjj);
            }
            p += rw;
            tp += tw;
        }
    }
    for (int s = 1; s < 16; s <<= 1) {
        st = st + __shfl_xor(st, s);
        sj = sj + __shfl_xor(sj, s);
        sjj = sjj + __shfl_xor(sjj, s);
    }
    if (!fits)
        return -1.0;
    const double n = (double)(tw * th), a = st, b = sj, c = sjj;
    const double var = c - b * b / n;
    return (var <= norm * 0.0625 || norm <= 0.0) ? 0.0 : a / sqrt(norm * var);
}

// Top cell c of a search, by one workgroup of 256: every half-pixel corner within around + 1 of twice its best corner
// (the inner 2 around + 1 square is search_slot's refinement, track.c:1697-1711; the ring gives take_back's
// neighbours, track.c:1737-1741), the window and the template staged in LDS (sm), 16 corners at a time.
template <typename T, int TWQ>
__device__ __forceinline__ void refine_one(const SBatch &b, const SJob &a, int c, float *sm)
{
    constexpr int tw = 16 * TWQ;
    const int th = b.th, step = b.step;
    const Level lv = a.lv;
    const int tid = threadIdx.x, lane = tid & 63, grp = tid >> 4;
    const int around = step > 1 ? 2 * step - 1 : 1, R = around + 1, side = 2 * R + 1;
    const vtg_cell cl = a.cells[a.aux->tops[c]];
    const int cx = 2 * cl.x, cy = 2 * cl.y;
    const int rx = (cx - R) >> 1, ry = (cy - R) >> 1;
    const int rw = ((cx + R) >> 1) - rx + tw + 2, rh = ((cy + R) >> 1) - ry + th + 2;
    float *tl = sm, *reg = sm + tw * th;
    const T *base = (const T *)lv.p;
    {
        // every load of the staging issued before the first store: the template (<= 16 per thread) and the region
        // (<= 24 per thread)
        float tv[16], rv[24];
        const int nt = tw * th, nr = rw * rh;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const int i = tid + 256 * k;
            tv[k] = a.t[i < nt ? i : nt - 1];
        }
#pragma unroll
        for (int k = 0; k < 24; k++) {
            const int i = tid + 256 * k, ic = i < nr ? i : nr - 1;
            const int r = ic / rw, kk = ic - r * rw;
            const int y = ry + r < 0 ? 0 : (ry + r >= lv.h ? lv.h - 1 : ry + r);
            const int x = rx + kk < 0 ? 0 : (rx + kk >= lv.w ? lv.w - 1 : rx + kk);
            rv[k] = Q6v<T>::get(base[(size_t)y * lv.pitch + x]);
        }
#pragma unroll
        for (int k = 0; k < 16; k++)
            if (tid + 256 * k < nt)
                tl[tid + 256 * k] = tv[k];
#pragma unroll
        for (int k = 0; k < 24; k++)
            if (tid + 256 * k < nr)
                reg[tid + 256 * k] = rv[k];
    }
    __syncthreads();
    const double norm = (double)a.nm[0];
    for (int q0 = 0; q0 < side * side; q0 += 16) {
        const int q = q0 + grp;
        const int hy = cy - R + q / side, hx = cx - R + q % side;
        const double v =
            score16_lds<TWQ>(reg, rx, ry, rw, tl, th, norm, lv.w, lv.h, q < side * side ? hx : -64, hy, lane);
        if (q < side * side && (lane & 15) == 0)
            a.aux->hs[c][q] = v;
    }
}

// The pick, the second peak clear of it by half a template (search_slot, track.c:1704-1725) and the parabola offsets
// from the pick's ring (take_back, track.c:1737-1741; peak_offset, 1650-1657). Selection only: no scoring here. By the
// NT threads of one workgroup, with LDS for the search's record (ax) and the candidates' picks. A search with nothing
// to search gets the empty record (score -2).
template <int NT>
__device__ void finish_search(const SBatch &b, const SJob &a, Aux &ax, vtg_cell *tc, double *scs, int *sxs, int *sys)
{
    const int ncell = a.cols * a.rows, tw = b.tw, th = b.th, step = b.step;
    {
        // the record into LDS, every load issued before the first store
        constexpr int NW = (int)(sizeof(Aux) / 4), PER = (NW + NT - 1) / NT;
        const unsigned int *src = (const unsigned int *)a.aux;
        unsigned int *dst = (unsigned int *)&ax;
        unsigned int v[PER];
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = threadIdx.x + NT * k;
            v[k] = i < NW ? src[i] : 0u;
        }
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = threadIdx.x + NT * k;
            if (i < NW)
                dst[i] = v[k];
        }
    }
    __syncthreads();
    const int ntop = a.rows == 0 ? 0 : ax.ntop;
    const int around = step > 1 ? 2 * step - 1 : 1, R = around + 1, side = 2 * R + 1;
    if ((int)threadIdx.x < ntop) {
        // lane c: candidate c's best half-pixel corner, first in row-major order
        const int c = threadIdx.x;
        const vtg_cell cl = a.cells[ax.tops[c]];
        tc[c] = cl;
        const int cx = 2 * cl.x, cy = 2 * cl.y;
        double s = -2.0;
        int bx = cx, by = cy;
        for (int hy = cy - around; hy <= cy + around; hy++)
            for (int hx = cx - around; hx <= cx + around; hx++) {
                const double v = ax.hs[c][(hy - cy + R) * side + hx - cx + R];
                if (v > s)
                    s = v, bx = hx, by = hy;
            }
        scs[c] = s;
        sxs[c] = bx;
        sys[c] = by;
    }
    __syncthreads();
    if (threadIdx.x != 0)
        return;
    double fn = -2.0, sec = -2.0;
    int fhx = 0, fhy = 0, fc = -1;
    for (int c = 0; c < ntop; c++)
        if (scs[c] > fn)
            fn = scs[c], fhx = sxs[c], fhy = sys[c], fc = c;
    for (int c = 0; c < ntop; c++) {
        const int dx = sxs[c] - fhx, dy = sys[c] - fhy;
        if ((dx * dx > tw * tw || dy * dy > th * th) && scs[c] > sec)
            sec = scs[c];
    }
    auto peak = [](double l, double m, double r) {
        const double den = l - 2.0 * m + r;
        if (den >= 0.0)
            return 0.0;
        const double o = 0.5 * (l - r) / den;
        return o < -0.5 ? -0.5 : o > 0.5 ? 0.5 : o;
    };
    vtg_found f;
    f.hx = fhx;
    f.hy = fhy;
    f.ncc = fn;
    f.second = sec;
    f.ox = f.oy = 0.0;
    f.cells = ncell;
    f.best_x = f.best_y = 0;
    f.best_v = -2.0f;
    if (fc >= 0) {
        const vtg_cell cl = tc[fc];
        const int ix = fhx - 2 * cl.x + R, iy = fhy - 2 * cl.y + R;
        const double *g = ax.hs[fc];
        f.ox = peak(g[iy * side + ix - 1], fn, g[iy * side + ix + 1]);
        f.oy = peak(g[(iy - 1) * side + ix], fn, g[(iy + 1) * side + ix]);
        const vtg_cell bc = tc[0];
        f.best_x = bc.x;
        f.best_y = bc.y;
        f.best_v = bc.v;
    }
    *a.found = f;
}

// Grid: (8, searches), 256 threads. Workgroup c refines top cell c of its search (refine_one) when the search has one;
// the last workgroup of a search to finish (the count in the search's record) then makes its record (finish_search)
// from every candidate's scores. Nothing waits: each workgroup counts itself once and leaves.
template <typename T, int TWQ>
__global__ __launch_bounds__(256) void ncc_refine_final(SBatch b)
{
    extern __shared__ float sm[];
    __shared__ Aux ax;
    __shared__ vtg_cell tc[8];
    __shared__ double scs[8];
    __shared__ int sxs[8], sys[8];
    __shared__ int last;
    const SJob &a = b.j[blockIdx.y];
    const int c = blockIdx.x;
    if (a.rows != 0 && c < a.aux->ntop) // the same for the whole workgroup
        refine_one<T, TWQ>(b, a, c, sm);
    // this workgroup's scores visible at agent scope before it is counted; the last one counted reads them all
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0)
        last = __hip_atomic_fetch_add(&a.aux->done, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
               (int)gridDim.x - 1;
    __syncthreads();
    if (!last)
        return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    finish_search<256>(b, a, ax, tc, scs, sxs, sys);
}

template <typename F>
int by_level_type(vtg_ctx *c, const vtg_pyr *p, int level, F &&f)
{
    const int eb = ncc_elem(p, level);
    switch (p->storage) {
    case VTG_STORE_F32: return f((float *)nullptr);
    case VTG_STORE_Q6: return eb == 1 ? f((uint8_t *)nullptr) : f((uint16_t *)nullptr);
    case VTG_STORE_F16: return eb == 1 ? f((uint8_t *)nullptr) : f((__half *)nullptr);
    case VTG_STORE_BF16: return eb == 1 ? f((uint8_t *)nullptr) : f((__hip_bfloat16 *)nullptr);
    case VTG_STORE_Q6W: return f((uint16_t *)nullptr);
    case VTG_STORE_Q6F:
    case VTG_STORE_F32_U8: return eb == 1 ? f((uint8_t *)nullptr) : f((float *)nullptr);
    }
    return fail(c, VTG_EINVAL, "storage %d", p->storage);
}

}  // namespace

// n templates, cut from pyramids p[k] at level[k], corner (x0[k], y0[k]), into t[k]: one launch per run of up to TMAX
// consecutive templates that share size and level element type. Arguments were checked by the caller.
int launch_ncc_templates(vtg_ctx *c, int n, vtg_tmpl *const *t, const vtg_pyr *const *p, const int *level,
                         const int *x0, const int *y0)
{
    for (int k0 = 0; k0 < n;) {
        const vtg_pyr *p0 = p[k0];
        const int l0 = level[k0];
        TBatch b{};
        b.tw = t[k0]->tw;
        b.th = t[k0]->th;
        b.wpr = t[k0]->words_per_row;
        int k = k0;
        for (; k < n && b.n < TMAX; k++) {
            if (t[k]->tw != b.tw || t[k]->th != b.th || p[k]->storage != p0->storage ||
                ncc_elem(p[k], level[k]) != ncc_elem(p0, l0))
                break;
            TJob &j = b.j[b.n++];
            j.lv = ncc_level(p[k], level[k]);
            j.x0 = x0[k];
            j.y0 = y0[k];
            j.t = t[k]->t;
            j.raw = t[k]->raw;
            j.nm = (float *)t[k]->block;
            j.su = (long long *)((char *)t[k]->block + 16);
            j.slow = (int *)((char *)t[k]->block + 24);
            j.words = ncc_elem(p[k], level[k]) == 1 ? t[k]->words : nullptr;
        }
        const int rc = by_level_type(c, p0, l0, [&](auto *tag) -> int {
            using T = std::remove_pointer_t<decltype(tag)>;
            hipEvent_t ev;
            timer_begin(c, "ncc_template", &ev);
            hipLaunchKernelGGL(ncc_template_kernel<T>, dim3(b.n), dim3(256), 0, c->stream, b);
            VTG_HIP(c, hipGetLastError());
            timer_end(c, "ncc_template", ev);
            return VTG_OK;
        });
        if (rc)
            return rc;
        k0 = k;
    }
    return VTG_OK;
}

int launch_ncc_template(vtg_ctx *c, const vtg_pyr *p, int level, int x0, int y0, vtg_tmpl *t)
{
    return launch_ncc_templates(c, 1, &t, &p, &level, &x0, &y0);
}

int launch_ncc_score(vtg_ctx *c, const vtg_pyr *p, const vtg_tmpl *t, int n, const int *xy, int half, double *out)
{
    if (n <= 0)
        return VTG_OK;
    const Level lv = ncc_level(p, t->level);
    const float *nm = (const float *)t->block;
    return by_level_type(c, p, t->level, [&](auto *tag) -> int {
        using T = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL(ncc_score_kernel<T>, dim3((n + 15) / 16), dim3(256), 0, c->stream, lv, t->t, t->tw, t->th,
                           nm, n, xy, half, out);
        VTG_HIP(c, hipGetLastError());
        return VTG_OK;
    });
}

static size_t aux_bytes() { return (sizeof(Aux) + 255) & ~(size_t)255; }

// One launch per stage for up to SMAX searches that share template size, level element type and step.
enum { MAP_GENERIC = 0, MAP_DOT4 = 1, MAP_MFMA = 2 };

// The cell maps of a group of searches on f's stream (errors reported on c).
static int launch_maps(vtg_ctx *c, vtg_ctx *f, const vtg_pyr *p0, int level, int mode, int wpr, const SBatch &b)
{
    int maxcols = 0, maxrows = 0, maxcell = 0;
    for (int k = 0; k < b.n; k++) {
        maxcols = std::max(maxcols, b.j[k].cols);
        maxrows = std::max(maxrows, b.j[k].rows);
        maxcell = std::max(maxcell, b.j[k].cols * b.j[k].rows);
    }
    hipEvent_t ev;
    if (maxcell > 0) {
        if (mode == MAP_MFMA) {
            const dim3 gx((16 * maxcols + XB - 1) / XB, (16 * maxrows + YBK - 1) / YBK, b.n);
            const int nt = b.tw * b.th;
            timer_begin(f, "ncc_xcorr_cells", &ev);
            if ((nt & (nt - 1)) == 0)
                hipLaunchKernelGGL(ncc_xcorr_cells<true>, gx, dim3(256), xcells_lds_bytes(b.th), f->stream, b);
            else
                hipLaunchKernelGGL(ncc_xcorr_cells<false>, gx, dim3(256), xcells_lds_bytes(b.th), f->stream, b);
            VTG_HIP(c, hipGetLastError());
            timer_end(f, "ncc_xcorr_cells", ev);
        } else if (mode == MAP_DOT4) {
            const dim3 grid((maxcols + 15) / 16, maxrows, b.n);
            const size_t lds = fast_lds_words(b.th, b.tw, wpr) * sizeof(uint32_t);
            timer_begin(f, "ncc_cells", &ev);
            if (wpr == 17)
                hipLaunchKernelGGL(ncc_cells_fast<17>, grid, dim3(256), lds, f->stream, b);
            else if (wpr == 9)
                hipLaunchKernelGGL(ncc_cells_fast<9>, grid, dim3(256), lds, f->stream, b);
            else
                hipLaunchKernelGGL(ncc_cells_fast<5>, grid, dim3(256), lds, f->stream, b);
            VTG_HIP(c, hipGetLastError());
            timer_end(f, "ncc_cells", ev);
        } else {
            int rc = by_level_type(c, p0, level, [&](auto *tag) -> int {
                using T = std::remove_pointer_t<decltype(tag)>;
                timer_begin(f, "ncc_cells_generic", &ev);
                hipLaunchKernelGGL(ncc_cells_generic<T>, dim3((maxcell + 3) / 4, b.n), dim3(256), 0, f->stream, b);
                VTG_HIP(c, hipGetLastError());
                timer_end(f, "ncc_cells_generic", ev);
                return VTG_OK;
            });
            if (rc)
                return rc;
        }
        const int nt = b.tw * b.th;
        if (mode != MAP_GENERIC && (nt & (nt - 1)) != 0) {
            // a template whose rounding term the fast map would leave out (flagged by the template kernel): its
            // search's map in the reference's arithmetic instead; every other search's workgroups leave at once
            SBatch s = b;
            s.slow_only = 1;
            int rc = by_level_type(c, p0, level, [&](auto *tag) -> int {
                using T = std::remove_pointer_t<decltype(tag)>;
                timer_begin(f, "ncc_cells_generic", &ev);
                hipLaunchKernelGGL(ncc_cells_generic<T>, dim3((maxcell + 3) / 4, b.n), dim3(256), 0, f->stream, s);
                VTG_HIP(c, hipGetLastError());
                timer_end(f, "ncc_cells_generic", ev);
                return VTG_OK;
            });
            if (rc)
                return rc;
        }
    }
    return VTG_OK;
}

// Top 8, refinement and pick of a group of searches on f's stream, after its maps (errors reported on c).
static int launch_finish(vtg_ctx *c, vtg_ctx *f, const vtg_pyr *p0, int level, const SBatch &b)
{
    hipEvent_t ev;
    timer_begin(f, "ncc_top8", &ev);
    hipLaunchKernelGGL(ncc_top8, dim3(b.n), dim3(1024), 0, f->stream, b);
    VTG_HIP(c, hipGetLastError());
    timer_end(f, "ncc_top8", ev);
    {
        // the refinement of every search's top cells and each search's record, one launch
        const int around = b.step > 1 ? 2 * b.step - 1 : 1, R = around + 1;
        const int rmax = R + 1 + 2; // region extent beyond the template, whole pixels, each way
        const size_t lds =
            ((size_t)b.tw * b.th + (size_t)(b.tw + rmax + 1) * (b.th + rmax + 1)) * sizeof(float);
        int rc = by_level_type(c, p0, level, [&](auto *tag) -> int {
            using T = std::remove_pointer_t<decltype(tag)>;
            const dim3 grid(8, b.n);
            timer_begin(f, "ncc_refine_final", &ev);
            switch (b.tw / 16) {
            case 1: hipLaunchKernelGGL((ncc_refine_final<T, 1>), grid, dim3(256), lds, f->stream, b); break;
            case 2: hipLaunchKernelGGL((ncc_refine_final<T, 2>), grid, dim3(256), lds, f->stream, b); break;
            case 3: hipLaunchKernelGGL((ncc_refine_final<T, 3>), grid, dim3(256), lds, f->stream, b); break;
            default: hipLaunchKernelGGL((ncc_refine_final<T, 4>), grid, dim3(256), lds, f->stream, b); break;
            }
            VTG_HIP(c, hipGetLastError());
            timer_end(f, "ncc_refine_final", ev);
            return VTG_OK;
        });
        if (rc)
            return rc;
    }
    return VTG_OK;
}

// n searches into device records: each window clipped to where the template fits (as search_slot clips), maps in
// cells[k] or f's scratch, picks in found_dev[k] or, when found_dev is null, in the scratch (returned in
// *found_scratch). Every stage runs on f's stream after the work submitted to c so far (f may be c).
// Stream-ordered; no allocation once the scratch has grown to the batch's size. fixed: working memory of fixed_bytes
// to use instead of the scratch (nothing is allocated then).
static int search_many(vtg_ctx *c, vtg_ctx *f, int n, const vtg_pyr *const *p, const vtg_tmpl *const *t,
                       const int *region, int step, vtg_cell *const *cells, vtg_found *found_dev,
                       vtg_found **found_scratch, void *fixed = nullptr, size_t fixed_bytes = 0)
{
    if (!c || !f)
        return VTG_EINVAL;
    if (f != c && f->device != c->device)
        return fail(c, VTG_EINVAL, "vtg_ncc_search: contexts on different devices");
    if (n < 1 || !p || !t || !region || (step != 1 && step != 2))
        return fail(c, VTG_EINVAL, "vtg_ncc_search: bad argument");
    const vtg_tmpl *t0 = t[0];
    if (!p[0] || !t0 || t0->level >= p[0]->levels)
        return fail(c, VTG_EINVAL, "vtg_ncc_search: bad pyramid or template");
    const int level0 = t0->level;
    for (int k = 0; k < n; k++) {
        if (!p[k] || !t[k] || t[k]->level >= p[k]->levels)
            return fail(c, VTG_EINVAL, "vtg_ncc_search: bad pyramid or template %d", k);
        if (t[k]->tw != t0->tw || t[k]->th != t0->th || (t[k]->words != nullptr) != (t0->words != nullptr) ||
            p[k]->storage != p[0]->storage || ncc_elem(p[k], t[k]->level) != ncc_elem(p[0], level0))
            return fail(c, VTG_EINVAL, "vtg_ncc_search: search %d differs from the first in template size or level "
                                       "type", k);
    }
    // u8 level: 64-wide templates go to the matrix cores, 16- and 32-wide ones to the dot4 kernel
    const bool u8 = ncc_elem(p[0], level0) == 1 && t0->words;
    const int mode = u8 && t0->tw == 64 ? MAP_MFMA
                     : u8 && (t0->th % YB) == 0 && t0->th >= YB &&
                               (t0->words_per_row == 17 || t0->words_per_row == 9 || t0->words_per_row == 5)
                         ? MAP_DOT4
                         : MAP_GENERIC;
    // the scratch: aux records, then found records (when the caller has none), then maps (for searches without one)
    std::vector<int> win((size_t)4 * n), ncell(n);
    const size_t auxb = aux_bytes();
    size_t bytes = (size_t)n * auxb + (found_dev ? 0 : (size_t)n * 256);
    for (int k = 0; k < n; k++) {
        const Level &lv = ncc_level(p[k], t[k]->level);
        const int hi_x = lv.w - t[k]->tw - 2, hi_y = lv.h - t[k]->th - 2;
        int x0 = region[4 * k], x1 = region[4 * k + 1], y0 = region[4 * k + 2], y1 = region[4 * k + 3];
        x0 = x0 < 1 ? 1 : x0;
        x1 = x1 > hi_x ? hi_x : x1;
        y0 = y0 < 1 ? 1 : y0;
        y1 = y1 > hi_y ? hi_y : y1;
        win[4 * k] = x0, win[4 * k + 1] = x1, win[4 * k + 2] = y0, win[4 * k + 3] = y1;
        ncell[k] = (x1 < x0 || y1 < y0) ? 0 : ((x1 - x0) / 16 + 1) * ((y1 - y0) / 16 + 1);
        if (!(cells && cells[k]))
            bytes += ((size_t)ncell[k] * sizeof(vtg_cell) + 255) & ~(size_t)255;
    }
    void *sc = nullptr;
    int rc;
    if (fixed) {
        if (bytes > fixed_bytes)
            return fail(c, VTG_ERANGE, "vtg_ncc_search_fixed: the search needs %zu bytes of working memory, the fixed "
                                       "array holds %zu", bytes, fixed_bytes);
        sc = fixed;
    } else if ((rc = scratch(f, bytes, &sc))) {
        // every stage runs on f's stream, so f's scratch is reused in stream order
        return rc;
    }
    if (f != c && (rc = vtg_ctx_wait(f, c)))
        return fail(c, rc, "vtg_ncc_search: %s", vtg_ctx_error(f));
    char *aux0 = (char *)sc, *fnd0 = aux0 + (size_t)n * auxb;
    char *cel = fnd0 + (found_dev ? 0 : (size_t)n * 256);
    if (found_scratch)
        *found_scratch = found_dev ? nullptr : (vtg_found *)fnd0;
    for (int k0 = 0; k0 < n; k0 += SMAX) {
        SBatch b{};
        b.n = std::min(SMAX, n - k0);
        b.tw = t0->tw;
        b.th = t0->th;
        b.step = step;
        for (int i = 0; i < b.n; i++) {
            const int k = k0 + i;
            SJob &j = b.j[i];
            const int *w = &win[4 * k];
            j.lv = ncc_level(p[k], t[k]->level);
            j.t = t[k]->t;
            j.nm = (const float *)t[k]->block;
            j.sum_t = (const long long *)((const char *)t[k]->block + 16);
            j.slow = (const int *)((const char *)t[k]->block + 24);
            j.words = t[k]->words;
            j.x0 = w[0], j.x1 = w[1], j.y0 = w[2], j.y1 = w[3];
            j.cols = ncell[k] ? (w[1] - w[0]) / 16 + 1 : 0;
            j.rows = ncell[k] ? (w[3] - w[2]) / 16 + 1 : 0;
            if (cells && cells[k]) {
                j.cells = cells[k];
            } else {
                j.cells = (vtg_cell *)cel;
                cel += ((size_t)ncell[k] * sizeof(vtg_cell) + 255) & ~(size_t)255;
            }
            j.found = found_dev ? found_dev + k : (vtg_found *)(fnd0 + (size_t)k * 256);
            j.aux = (Aux *)(aux0 + (size_t)k * auxb);
        }
        if ((rc = launch_maps(c, f, p[0], level0, mode, t0->words_per_row, b)) ||
            (rc = launch_finish(c, f, p[0], level0, b)))
            return rc;
    }
    return VTG_OK;
}

}  // namespace vtg

// The working memory of vtg_ncc_search_fixed: a fixed array, one copy per device, held until the process ends. Host
// code reaches it only through the address hipGetSymbolAddress gives for the current device (the name itself is a
// host-side stand-in, not a device address). Enough for one exhaustive 64 x 64 search of a 1920 x 1080 level (its
// record and 7,424-cell map take 94 KiB), not for one of a 3840 x 2160 level (367 KiB).
static constexpr size_t kFixedBytes = (size_t)256 << 10;
static __device__ __align__(256) unsigned char scratch[kFixedBytes];

namespace vtg {

static void *g_fixed[64]; // scratch's address on each device, once looked up
static std::mutex g_fixed_mu;

int fixed_scratch_init(vtg_ctx *c)
{
    if (c->device < 0 || c->device >= 64)
        return fail(c, VTG_ERANGE, "device %d: the fixed search array is kept for devices 0..63", c->device);
    std::lock_guard<std::mutex> lk(g_fixed_mu);
    if (g_fixed[c->device])
        return VTG_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(::scratch)); // the context's device is current
    if (e != hipSuccess)
        return hip_fail(c, e, "hipGetSymbolAddress(scratch)");
    g_fixed[c->device] = ptr;
    return VTG_OK;
}

size_t tmpl_bytes(int tw, int th, bool u8)
{
    const size_t fbytes = ((size_t)tw * th * sizeof(float) + 255) & ~(size_t)255;
    return 256 + 2 * fbytes + (u8 ? (size_t)4 * th * (tw / 4 + 1) * sizeof(uint32_t) : 0);
}

int tmpl_new(vtg_ctx *c, int tw, int th, int level, bool u8, void *mem, vtg_tmpl **out)
{
    *out = nullptr;
    vtg_tmpl *t = new vtg_tmpl();
    t->tw = tw;
    t->th = th;
    t->level = level;
    t->words_per_row = tw / 4 + 1;
    t->owned = mem == nullptr;
    if (mem) {
        t->block = mem;
    } else {
        hipError_t e = hipMalloc(&t->block, tmpl_bytes(tw, th, u8));
        if (e != hipSuccess) {
            delete t;
            return hip_fail(c, e, "vtg_ncc_template: hipMalloc");
        }
    }
    const size_t fbytes = ((size_t)tw * th * sizeof(float) + 255) & ~(size_t)255;
    t->t = (float *)((char *)t->block + 256);
    t->raw = (float *)((char *)t->t + fbytes);
    t->words = u8 ? (uint32_t *)((char *)t->raw + fbytes) : nullptr;
    *out = t;
    return VTG_OK;
}

}  // namespace vtg

using namespace vtg;

extern "C" {

int vtg_ncc_template(vtg_ctx *c, const vtg_pyr *p, int level, int x0, int y0, int tw, int th, vtg_tmpl **out)
{
    if (!c || !p || !out || level < 0 || level >= p->levels)
        return fail(c, VTG_EINVAL, "vtg_ncc_template: bad argument");
    *out = nullptr;
    if (tw < 16 || tw > 64 || (tw & 15) || th < 1 || th > 64)
        return fail(c, VTG_ERANGE, "vtg_ncc_template: %dx%d (width 16..64 in 16s, height 1..64)", tw, th);
    const Level &lv = ncc_level(p, level);
    if (x0 < 1 || y0 < 1 || x0 + tw >= lv.w - 1 || y0 + th >= lv.h - 1)
        return fail(c, VTG_EINVAL, "vtg_ncc_template: window at (%d,%d) does not fit level %d (%dx%d)", x0, y0, level,
                    lv.w, lv.h);
    vtg_tmpl *t = nullptr;
    int rc = tmpl_new(c, tw, th, level, ncc_elem(p, level) == 1, nullptr, &t);
    if (rc)
        return rc;
    rc = launch_ncc_template(c, p, level, x0, y0, t);
    if (rc) {
        (void)hipFree(t->block);
        delete t;
        return rc;
    }
    *out = t;
    return VTG_OK;
}

int vtg_ncc_template_update(vtg_ctx *c, vtg_tmpl *t, const vtg_pyr *p, int level, int x0, int y0)
{
    if (!c || !t || !p || level < 0 || level >= p->levels)
        return fail(c, VTG_EINVAL, "vtg_ncc_template_update: bad argument");
    const Level &lv = ncc_level(p, level);
    if (x0 < 1 || y0 < 1 || x0 + t->tw >= lv.w - 1 || y0 + t->th >= lv.h - 1)
        return fail(c, VTG_EINVAL, "vtg_ncc_template_update: window at (%d,%d) does not fit level %d", x0, y0, level);
    if ((ncc_elem(p, level) == 1) != (t->words != nullptr))
        return fail(c, VTG_EINVAL, "vtg_ncc_template_update: level element size differs from the template's");
    t->level = level;
    return launch_ncc_template(c, p, level, x0, y0, t);
}

int vtg_ncc_template_update_many(vtg_ctx *c, int n, vtg_tmpl *const *t, const vtg_pyr *const *p, const int *level,
                                 const int *x0, const int *y0)
{
    if (!c)
        return VTG_EINVAL;
    if (n < 0 || (n > 0 && (!t || !p || !level || !x0 || !y0)))
        return fail(c, VTG_EINVAL, "vtg_ncc_template_update_many: bad argument");
    for (int k = 0; k < n; k++) {
        if (!t[k] || !p[k] || level[k] < 0 || level[k] >= p[k]->levels)
            return fail(c, VTG_EINVAL, "vtg_ncc_template_update_many: bad template, pyramid or level %d", k);
        const Level &lv = ncc_level(p[k], level[k]);
        if (x0[k] < 1 || y0[k] < 1 || x0[k] + t[k]->tw >= lv.w - 1 || y0[k] + t[k]->th >= lv.h - 1)
            return fail(c, VTG_EINVAL, "vtg_ncc_template_update_many: window %d at (%d,%d) does not fit level %d", k,
                        x0[k], y0[k], level[k]);
        if ((ncc_elem(p[k], level[k]) == 1) != (t[k]->words != nullptr))
            return fail(c, VTG_EINVAL, "vtg_ncc_template_update_many: level element size of %d differs from its "
                                       "template's", k);
    }
    for (int k = 0; k < n; k++)
        t[k]->level = level[k];
    return launch_ncc_templates(c, n, t, p, level, x0, y0);
}

int vtg_ncc_template_destroy(vtg_ctx *c, vtg_tmpl *t)
{
    if (!c || !t)
        return VTG_EINVAL;
    (void)hipStreamSynchronize(c->stream);
    if (t->owned)
        (void)hipFree(t->block);
    delete t;
    return VTG_OK;
}

int vtg_ncc_template_read(vtg_ctx *c, const vtg_tmpl *t, float *app, float *norm)
{
    if (!c || !t)
        return VTG_EINVAL;
    float nm[2];
    VTG_HIP(c, hipMemcpyAsync(nm, t->block, sizeof nm, hipMemcpyDeviceToHost, c->stream));
    if (app)
        VTG_HIP(c, hipMemcpyAsync(app, t->t, (size_t)t->tw * t->th * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    VTG_HIP(c, hipStreamSynchronize(c->stream));
    if (norm)
        *norm = nm[0];
    return VTG_OK;
}

int vtg_ncc_search_many(vtg_ctx *c, int n, const vtg_pyr *const *p, const vtg_tmpl *const *t, const int *region,
                        int step, struct vtg_cell *const *cells, struct vtg_found *found_dev)
{
    if (!found_dev)
        return fail(c, VTG_EINVAL, "vtg_ncc_search_many: null found records");
    return search_many(c, c, n, p, t, region, step, cells, found_dev, nullptr);
}

int vtg_ncc_search_many_split(vtg_ctx *map_ctx, vtg_ctx *finish_ctx, int n, const vtg_pyr *const *p,
                              const vtg_tmpl *const *t, const int *region, int step, struct vtg_cell *const *cells,
                              struct vtg_found *found_dev)
{
    if (!map_ctx || !finish_ctx)
        return VTG_EINVAL;
    if (!found_dev)
        return fail(map_ctx, VTG_EINVAL, "vtg_ncc_search_many_split: null found records");
    return search_many(map_ctx, finish_ctx, n, p, t, region, step, cells, found_dev, nullptr);
}

int vtg_ncc_search_async(vtg_ctx *c, const vtg_pyr *p, const vtg_tmpl *t, int x0, int x1, int y0, int y1, int step,
                         struct vtg_cell *cells, struct vtg_found *found_dev)
{
    if (!found_dev)
        return fail(c, VTG_EINVAL, "vtg_ncc_search_async: null found record");
    const int region[4] = {x0, x1, y0, y1};
    vtg_cell *cl[1] = {cells};
    return search_many(c, c, 1, &p, &t, region, step, cl, found_dev, nullptr);
}

int vtg_ncc_search_fixed(vtg_ctx *c, const vtg_pyr *p, const vtg_tmpl *t, int x0, int x1, int y0, int y1, int step,
                         struct vtg_cell *cells, struct vtg_found *found_dev)
{
    if (!c)
        return VTG_EINVAL;
    if (!found_dev)
        return fail(c, VTG_EINVAL, "vtg_ncc_search_fixed: null found record");
    void *fixed = c->device >= 0 && c->device < 64 ? g_fixed[c->device] : nullptr;
    if (!fixed)
        return fail(c, VTG_EIO, "vtg_ncc_search_fixed: the fixed array of device %d was not looked up", c->device);
    const int region[4] = {x0, x1, y0, y1};
    vtg_cell *cl[1] = {cells};
    return search_many(c, c, 1, &p, &t, region, step, cl, found_dev, nullptr, fixed, kFixedBytes);
}

int vtg_ncc_search(vtg_ctx *c, const vtg_pyr *p, const vtg_tmpl *t, int x0, int x1, int y0, int y1, int step,
                   struct vtg_cell *cells, struct vtg_found *found)
{
    if (!found)
        return fail(c, VTG_EINVAL, "vtg_ncc_search: null found record");
    const int region[4] = {x0, x1, y0, y1};
    vtg_cell *cl[1] = {cells};
    vtg_found *fd = nullptr;
    int rc = search_many(c, c, 1, &p, &t, region, step, cl, nullptr, &fd);
    if (rc)
        return rc;
    void *host = nullptr;
    if ((rc = pinned(c, sizeof(vtg_found), &host)))
        return rc;
    VTG_HIP(c, hipMemcpyAsync(host, fd, sizeof(vtg_found), hipMemcpyDeviceToHost, c->stream));
    VTG_HIP(c, hipStreamSynchronize(c->stream));
    *found = *(vtg_found *)host;
    return VTG_OK;
}
int vtg_ncc_score(vtg_ctx *c, const vtg_pyr *p, const vtg_tmpl *t, int n, const int *xy, int half, double *out)
{
    if (!c || !p || !t || n < 0 || (n > 0 && (!xy || !out)) || t->level >= p->levels)
        return fail(c, VTG_EINVAL, "vtg_ncc_score: bad argument");
    return launch_ncc_score(c, p, t, n, xy, half, out);
}

}  // extern "C"

extern "C" int vtg_correlate_planes(vtg_ctx *c, const struct vt_plane *pl, const float *app, int tw, int th,
                                    double norm, int n, const int *x, const int *y, double *out)
{
    if (!c || !pl || !app || n < 0 || (n > 0 && (!x || !y || !out)) || tw < 1 || th < 1 || tw * th > 64 * 64)
        return fail(c, VTG_EINVAL, "vtg_correlate_planes: bad argument");
    if (n == 0)
        return VTG_OK;
    vtg_pyr *p = nullptr;
    int rc = pyr_from_planes(c, pl, 1, &p);
    if (rc)
        return rc;
    // a template record around the caller's values: norm, mean (unused) and the fp32 pixels
    vtg_tmpl t;
    t.tw = tw;
    t.th = th;
    t.level = 0;
    const size_t fb = (size_t)tw * th * sizeof(float), xyb = (size_t)2 * n * sizeof(int);
    std::vector<int> xy((size_t)2 * n);
    for (int i = 0; i < n; i++)
        xy[2 * i] = x[i], xy[2 * i + 1] = y[i];
    void *dev = nullptr;
    hipError_t e = hipMalloc(&dev, 256 + fb + xyb + (size_t)n * sizeof(double));
    if (e != hipSuccess) {
        vtg_pyr_destroy(c, p);
        return hip_fail(c, e, "vtg_correlate_planes: hipMalloc");
    }
    t.block = dev;
    t.t = (float *)((char *)dev + 256);
    int *dxy = (int *)((char *)dev + 256 + fb);
    double *dout = (double *)((char *)dxy + xyb);
    const float nm[2] = {(float)norm, 0.0f};
    e = hipMemcpyAsync(dev, nm, sizeof nm, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(t.t, app, fb, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dxy, xy.data(), xyb, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        // norm travels as a float: the reference passes the template's float norm widened (track.c:312)
        rc = launch_ncc_score(c, p, &t, n, dxy, 0, dout);
        if (!rc)
            e = hipMemcpyAsync(out, dout, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (!rc && e == hipSuccess)
            e = hipStreamSynchronize(c->stream);
    }
    if (e != hipSuccess && !rc)
        rc = hip_fail(c, e, "vtg_correlate_planes");
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(dev);
    t.block = nullptr;
    vtg_pyr_destroy(c, p);
    return rc;
}
