// ldnorm.hpp: the contrast stretch of one Q6 value, clamp(rintf((float)((long double)s * g + o)), 0, 16320), with
// the reference's rounding reproduced exactly in integer arithmetic (img_norm_ref, src/kern/img_ref.c:42-55: the x87
// product of two floats is exact, the sum is rounded once to a 64-bit significand, then to float, then to an integer).
// GPUs have no 80-bit type, so the sum is formed exactly as a 128-bit integer times a power of two, rounded to 64
// significant bits (nearest, ties to even), then to 24, and the float is rounded to an integer (ties to even).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vtg {

struct FDec {
    int64_t m; // signed significand (|m| < 2^24), value = m * 2^e
    int e;
};

__host__ __device__ inline FDec fdec(float f)
{
    uint32_t b;
    __builtin_memcpy(&b, &f, 4);
    const int ex = (int)((b >> 23) & 0xffu);
    int64_t m = (int64_t)(b & 0x7fffffu);
    int e;
    if (ex == 0) {
        e = -149;
    } else {
        m |= (int64_t)1 << 23;
        e = ex - 150;
    }
    return FDec{(b >> 31) ? -m : m, e};
}

// v shifted by s bits (left if s >= 0, right otherwise with any lost bit jammed into bit 0)
__host__ __device__ inline unsigned __int128 shift_jam(unsigned __int128 v, int s)
{
    if (s >= 0)
        return v << s;
    const int r = -s;
    if (r >= 127)
        return v != 0 ? 1 : 0;
    const unsigned __int128 lost = v & (((unsigned __int128)1 << r) - 1);
    return (v >> r) | (lost != 0 ? 1 : 0);
}

// Round m (> 0) to `bits` significant bits, nearest, ties to even; the exponent e grows by the bits dropped.
__host__ __device__ inline void round_sig(unsigned __int128 &m, int &e, int bits)
{
    int msb = 127;
    while (msb > 0 && !((m >> msb) & 1))
        msb--;
    if (msb < bits)
        return;
    const int r = msb - bits + 1;
    const unsigned __int128 rem = m & (((unsigned __int128)1 << r) - 1), half = (unsigned __int128)1 << (r - 1);
    unsigned __int128 q = m >> r;
    if (rem > half || (rem == half && (q & 1)))
        q++;
    e += r;
    if (q >> bits) { // carried into a new bit
        q >>= 1;
        e++;
    }
    m = q;
}

// rintf(f) in [0, 16320] for an integer value m * 2^e with |m| < 2^24 and sign neg.
__host__ __device__ inline float clamp_rint(bool neg, uint64_t m, int e)
{
    if (m == 0 || neg)
        return 0.0f;
    // m * 2^e >= 2^15 > 16320 whenever the top bit reaches 2^15
    int msb = 63;
    while (!((m >> msb) & 1))
        msb--;
    if (msb + e >= 15)
        return 16320.0f;
    int64_t v;
    if (e >= 0) {
        v = (int64_t)(m << e);
    } else {
        const int r = -e;
        if (r > 40)
            return 0.0f; // below 2^-16: rounds to 0
        const uint64_t q = m >> r, rem = m & ((1ull << r) - 1), half = 1ull << (r - 1);
        v = (int64_t)(q + ((rem > half || (rem == half && (q & 1))) ? 1 : 0));
    }
    return v > 16320 ? 16320.0f : (float)v;
}

__host__ __device__ inline float norm_exact(float s, float g, float o)
{
    const FDec a = fdec(s), b = fdec(g), c = fdec(o);
    const int64_t pm = a.m * b.m; // exact, |pm| < 2^48
    const int pe = a.e + b.e;
    const bool pn = pm < 0, on = c.m < 0;
    const unsigned __int128 pa = (unsigned __int128)(pn ? -pm : pm), oa = (unsigned __int128)(on ? -c.m : c.m);
    if (pa == 0 && oa == 0)
        return 0.0f;
    // bit 0 of the accumulator weighs 2^base; the larger term's top bit lands at or below bit 110
    const int top = (pa ? pe + 48 : -100000) > (oa ? c.e + 24 : -100000) ? pe + 48 : c.e + 24;
    const int base = top - 110;
    const unsigned __int128 x = pa ? shift_jam(pa, pe - base) : 0, y = oa ? shift_jam(oa, c.e - base) : 0;
    bool neg;
    unsigned __int128 mag;
    if (pn == on) {
        mag = x + y;
        neg = pn;
    } else if (x >= y) {
        mag = x - y;
        neg = pn;
    } else {
        mag = y - x;
        neg = on;
    }
    if (mag == 0)
        return 0.0f;
    int e = base;
    round_sig(mag, e, 64); // the long double sum
    round_sig(mag, e, 24); // (float)
    return clamp_rint(neg, (uint64_t)mag, e);
}

}  // namespace vtg
