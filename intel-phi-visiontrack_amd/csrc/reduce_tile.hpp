// reduce_tile.hpp: one 64 x TH output tile of the pyramid reduce (img_reduce_ref, src/kern/img_ref.c:98-114), shared by
// the per-level reduce launches (pyramid.hip) and the fused pyramid + LK launch (lk.hip). The 131 x (2 TH + 3) input
// tile (2-pixel halo, coordinates clamped: the replicated border) is read with every load of a thread issued before
// the first is used, then summed in LDS: 5 rows at each even row, then 5 columns at each even column. EXACT: integer
// sums and one (h + 128) >> 8, bit-exact to the reference; else fp32 without rounding (fp16 / bf16 levels). COPY: the
// level-0 job also writes the frame into level 0 from the same tile, so the frame is read once. All 256 threads of the
// workgroup call it together.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "vtg_internal.hpp"

namespace vtg {
namespace rt {

constexpr int TW = 64, IW = 2 * TW + 3;

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Element -> value in Q6 units, and back. Exact storages carry integers; fp16 / bf16 carry grey = Q6 / 64.
template <typename T> struct Q6;
template <> struct Q6<uint8_t> {
    static __device__ __forceinline__ float get(uint8_t v) { return (float)v * 64.0f; }
};
template <> struct Q6<uint16_t> {
    static __device__ __forceinline__ float get(uint16_t v) { return (float)v; }
    static __device__ __forceinline__ uint16_t put(int q) { return (uint16_t)q; }
};
template <> struct Q6<float> {
    static __device__ __forceinline__ float get(float v) { return v; }
    static __device__ __forceinline__ float put(int q) { return (float)q; }
    static __device__ __forceinline__ float putf(float q) { return q; }
};
template <> struct Q6<__half> {
    static __device__ __forceinline__ float get(__half v) { return __half2float(v) * 64.0f; }
    static __device__ __forceinline__ __half putf(float q) { return __float2half(q * (1.0f / 64.0f)); }
};
template <> struct Q6<__hip_bfloat16> {
    static __device__ __forceinline__ float get(__hip_bfloat16 v) { return __bfloat162float(v) * 64.0f; }
    static __device__ __forceinline__ __hip_bfloat16 putf(float q) { return __float2bfloat16(q * (1.0f / 64.0f)); }
};

// One reduce job: a level made from the level below it, possibly writing the frame into level 0 on the way (copy, cp
// elements a row) and, from a u8 frame, its bytes into a second plane as well (copy8, cp8 bytes a row; null: none).
struct RJob {
    const void *src;
    void *dst, *copy;
    int sw, sh, sp, dw, dh, dp, cp;
    int tiles_x, t0; // tiles across the output, first linear tile of the job in the launch
    void *copy8;
    int cp8;
};
constexpr int JMAX = 32;
struct RJobs {
    RJob j[JMAX];
    int t0s[JMAX]; // j[k].t0 again, side by side: a workgroup finds its job from a few wide scalar loads
    int n;
};

template <typename TC, typename TS> struct CopyConv {
    static __device__ __forceinline__ TC get(TS v) { return v; }
};
template <> struct CopyConv<float, uint8_t> {
    static __device__ __forceinline__ float get(uint8_t v) { return (float)v * 64.0f; }
};

template <bool EXACT> struct AccT { using type = float; };
template <> struct AccT<true> { using type = int; };

// Four frame bytes (one aligned word) into four level-0 elements at p, 16-byte aligned for fp32, 4 for bytes.
template <typename TC> __device__ __forceinline__ void store4(TC *p, uint32_t w)
{
#pragma unroll
    for (int b = 0; b < 4; b++)
        p[b] = CopyConv<TC, uint8_t>::get((uint8_t)(w >> (8 * b)));
}
template <> __device__ __forceinline__ void store4<float>(float *p, uint32_t w)
{
    *(float4 *)p = make_float4((float)(w & 255u) * 64.0f, (float)((w >> 8) & 255u) * 64.0f,
                               (float)((w >> 16) & 255u) * 64.0f, (float)(w >> 24) * 64.0f);
}
template <> __device__ __forceinline__ void store4<uint8_t>(uint8_t *p, uint32_t w) { *(uint32_t *)p = w; }

// Reduce for a batch of jobs, one 64 x 16 output tile per workgroup. The 131 x 35 input tile (2-pixel halo,
// coordinates clamped: the replicated border) is read with every load of a thread issued before the first is used,
// then summed in LDS: 5 rows at each even row, then 5 columns at each even column. EXACT: integer sums and one
// (h + 128) >> 8, bit-exact to the reference; else fp32 without rounding (fp16 / bf16 levels).
// LDS a tile needs: the input tile (2 TH + 3 rows) and the vertical sums (TH rows), IW + 1 accumulators a row.
template <int TH> constexpr int tile_lds_elems() { return (2 * TH + 3 + TH) * (IW + 1); }

// lds: tile_lds_elems<TH>() accumulators of LDS owned by the caller (free again when this returns).
template <typename TS, typename TD, typename TC, bool COPY, bool EXACT, int TH>
__device__ __forceinline__ void reduce_tile(const RJob &J, int tile, typename AccT<EXACT>::type *lds)
{
    constexpr int IH = 2 * TH + 3;
    using Acc = typename AccT<EXACT>::type;
    Acc (*tin)[IW + 1] = (Acc(*)[IW + 1])lds;
    Acc (*tv)[IW + 1] = (Acc(*)[IW + 1])(lds + IH * (IW + 1));
    const int tyi = tile / J.tiles_x, txi = tile - tyi * J.tiles_x;
    const int ox0 = txi * TW, oy0 = tyi * TH;
    const int ix0 = 2 * ox0 - 2, iy0 = 2 * oy0 - 2;
    const TS *src = (const TS *)J.src;
    bool fast = false;
    if constexpr (sizeof(TS) == 1) {
        // A u8 source tile whose columns all lie inside the frame, rows and pointers word-aligned: the input read as
        // 34 aligned words a row (columns ix0 - 2 .. ix0 + 133), the level-0 copies written a word (4 pixels) at a
        // time. Rows are still clamped. The other tiles take the element path below; both store the same values.
        fast = txi > 0 && ix0 + IW + 3 <= J.sw && (J.sp & 3) == 0 && ((size_t)J.src & 3) == 0 &&
               (!COPY || (((size_t)J.copy & 15) == 0 && (J.cp * sizeof(TC)) % 16 == 0 &&
                          (!J.copy8 || (((size_t)J.copy8 & 3) == 0 && (J.cp8 & 3) == 0))));
        if (fast) {
            constexpr int NW = IW / 4 + 2, NWI = IH * NW, NWL = (NWI + 255) / 256; // 34 words a row
            const uint8_t *s8 = (const uint8_t *)J.src + ix0 - 2;                  // word-aligned: ix0 = 128 txi - 2
            uint32_t v[NWL];
#pragma unroll
            for (int k = 0; k < NWL; k++) {
                const int i = threadIdx.x + 256 * k, ic = i < NWI ? i : NWI - 1;
                const int r = ic / NW, w = ic - r * NW;
                v[k] = *(const uint32_t *)(s8 + (size_t)clampi(iy0 + r, 0, J.sh - 1) * J.sp + 4 * w);
            }
#pragma unroll
            for (int k = 0; k < NWL; k++) {
                const int i = threadIdx.x + 256 * k;
                if (i < NWI) {
                    const int r = i / NW, w = i - r * NW;
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        const int cc = 4 * w - 2 + b; // byte b of word w is column ix0 + cc
                        if (cc >= 0 && cc < IW)
                            tin[r][cc] = (Acc)Q6<TS>::get((TS)(v[k] >> (8 * b)));
                    }
                    if (COPY) {
                        // words 1 .. 2 TW / 4 of rows 2 .. 2 TH + 1 are the tile's own level-0 pixels
                        const int yy = iy0 + r, xx = ix0 - 2 + 4 * w;
                        if (r >= 2 && r < 2 + 2 * TH && w >= 1 && w <= 2 * TW / 4 && yy < J.sh) {
                            store4<TC>((TC *)J.copy + (size_t)yy * J.cp + xx, v[k]);
                            if (J.copy8)
                                *(uint32_t *)((uint8_t *)J.copy8 + (size_t)yy * J.cp8 + xx) = v[k];
                        }
                    }
                }
            }
        }
    }
    if (!fast) {
        constexpr int NIN = IH * IW, NLD = (NIN + 255) / 256;
        TS raw[NLD];
#pragma unroll
        for (int k = 0; k < NLD; k++) {
            const int i = threadIdx.x + 256 * k, ic = i < NIN ? i : NIN - 1;
            const int r = ic / IW, cc = ic - r * IW;
            raw[k] = src[(size_t)clampi(iy0 + r, 0, J.sh - 1) * J.sp + clampi(ix0 + cc, 0, J.sw - 1)];
        }
#pragma unroll
        for (int k = 0; k < NLD; k++) {
            const int i = threadIdx.x + 256 * k;
            if (i < NIN) {
                const int r = i / IW, cc = i - r * IW;
                tin[r][cc] = (Acc)Q6<TS>::get(raw[k]);
                if (COPY) {
                    const int yy = iy0 + r, xx = ix0 + cc;
                    if (r >= 2 && r < 2 + 2 * TH && cc >= 2 && cc < 2 + 2 * TW && yy < J.sh && xx < J.sw) {
                        ((TC *)J.copy)[(size_t)yy * J.cp + xx] = CopyConv<TC, TS>::get(raw[k]);
                        if constexpr (sizeof(TS) == 1)
                            if (J.copy8)
                                ((uint8_t *)J.copy8)[(size_t)yy * J.cp8 + xx] = (uint8_t)raw[k];
                    }
                }
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < TH * IW; i += 256) {
        const int j = i / IW, cc = i - j * IW, r = 2 * j;
        tv[j][cc] = tin[r][cc] + tin[r + 4][cc] + (Acc)4 * (tin[r + 1][cc] + tin[r + 3][cc]) + (Acc)6 * tin[r + 2][cc];
    }
    __syncthreads();
    TD *dst = (TD *)J.dst;
#pragma unroll
    for (int k = 0; k < TH * TW / 256; k++) {
        const int i = threadIdx.x + 256 * k;
        const int j = i / TW, kk = i - j * TW;
        const int ox = ox0 + kk, oy = oy0 + j;
        if (ox < J.dw && oy < J.dh) {
            const int cc = 2 * kk;
            const Acc h = tv[j][cc] + tv[j][cc + 4] + (Acc)4 * (tv[j][cc + 1] + tv[j][cc + 3]) + (Acc)6 * tv[j][cc + 2];
            if constexpr (EXACT)
                dst[(size_t)oy * J.dp + ox] = Q6<TD>::put((h + 128) >> 8);
            else
                dst[(size_t)oy * J.dp + ox] = Q6<TD>::putf(h * (1.0f / 256.0f));
        }
    }
}


}  // namespace rt
}  // namespace vtg
