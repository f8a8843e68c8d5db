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

// Exact long double arithmetic on non-negative values m * 2^e (m < 2^64), each result rounded as x87 rounds it: the
// still-cell differences and sigmas (src/kern/still.c:99-111, 136).

struct XL {
    uint64_t m;
    int e;
};

__host__ __device__ inline XL xl_add(XL a, float f)
{
    const FDec d = fdec(f); // f >= 0
    if (d.m == 0)
        return a;
    if (a.m == 0)
        return XL{(uint64_t)d.m, d.e};
    const int top = (a.e + 64 > d.e + 24 ? a.e + 64 : d.e + 24), base = top - 126;
    unsigned __int128 mag = shift_jam((unsigned __int128)a.m, a.e - base) + shift_jam((unsigned __int128)d.m, d.e - base);
    int e = base;
    round_sig(mag, e, 64);
    return XL{(uint64_t)mag, e};
}

// a / n (n >= 1), rounded to 64 bits and then to float: (float)(long double / n)
__host__ __device__ inline float xl_div_to_float(XL a, int n)
{
    if (a.m == 0)
        return 0.0f;
    const int lz = __builtin_clzll(a.m); // the top bit to bit 63, so the quotient keeps more than 64 bits
    a.m <<= lz;
    a.e -= lz;
    const unsigned __int128 num = (unsigned __int128)a.m << 60;
    const unsigned __int128 q = num / (unsigned __int128)n, r = num - q * (unsigned __int128)n;
    unsigned __int128 mag = (q << 1) | (r != 0 ? 1 : 0); // a sticky bit below every rounding position
    int e = a.e - 61;
    round_sig(mag, e, 64);
    round_sig(mag, e, 24);
    return ldexpf((float)(uint64_t)mag, e);
}

__host__ __device__ inline uint64_t isqrt128(unsigned __int128 v)
{
    unsigned __int128 r = 0, bit = (unsigned __int128)1 << 126;
    while (bit > v)
        bit >>= 2;
    while (bit) {
        if (v >= r + bit) {
            v -= r + bit;
            r = (r >> 1) + bit;
        } else {
            r >>= 1;
        }
        bit >>= 2;
    }
    return (uint64_t)r;
}

// (float)sqrtl((long double)a + (long double)b) for a, b >= 0: the sum rounded to 64 bits, its root to 64, then to 24
__host__ __device__ inline float sqrt_sum(float a, float b)
{
    if (!(a >= 0.0f && b >= 0.0f)) // a covariance is never negative; keep whatever a negative sum gives
        return (float)sqrt((double)a + (double)b);
    XL s = xl_add(XL{0, 0}, a);
    s = xl_add(s, b);
    if (s.m == 0)
        return 0.0f;
    // s.m * 2^e with an even exponent and the integer part in [2^126, 2^128), so the root has 64 bits
    const int lz = __builtin_clzll(s.m); // s.m < 2^64
    int k = 64 + lz;                      // shift that puts the top bit at 127
    if (((s.e - k) & 1) != 0)
        k--;                              // top bit at 126
    const unsigned __int128 M = (unsigned __int128)s.m << k;
    uint64_t r = isqrt128(M);
    const unsigned __int128 rem = M - (unsigned __int128)r * r;
    unsigned __int128 mag = r;
    int e = (s.e - k) / 2;
    if (rem > (unsigned __int128)r) // the root's fraction is above one half (it is never exactly one half)
        mag += 1;
    round_sig(mag, e, 64);
    round_sig(mag, e, 24);
    return ldexpf((float)(uint64_t)mag, e);
}

}  // namespace vtg

namespace vtg {

// Signed x87 long double values m * 2^e (m < 2^64, 0 for zero) and the operations the LK covariance takes
// (lk_driver.inc:100-104): products and quotients rounded to a 64-bit significand (nearest, ties to even), as x87
// rounds them at its default precision.
struct XS {
    bool neg;
    unsigned __int128 m;
    int e;
};

__host__ __device__ inline XS xs_from_float(float f)
{
    const FDec d = fdec(f);
    return XS{d.m < 0, (unsigned __int128)(d.m < 0 ? -d.m : d.m), d.e};
}

__host__ __device__ inline XS xs_from_int(int64_t v)
{
    return XS{v < 0, (unsigned __int128)(v < 0 ? -v : v), 0};
}

__host__ __device__ inline XS xs_mul(XS a, XS b)
{
    XS r{a.neg != b.neg, a.m * b.m, a.e + b.e}; // both below 2^64
    if (r.m)
        round_sig(r.m, r.e, 64);
    return r;
}

__host__ __device__ inline XS xs_div(XS a, XS b)
{
    // a / b = (a.m 2^s / b.m) 2^(a.e - b.e - s), with a.m and b.m first brought to 2^63 .. 2^64 and s = 64: a quotient
    // of 64 or 65 bits, and a sticky bit below it for the remainder
    if (!a.m)
        return XS{false, 0, 0};
    int la = 0, lb = 0;
    while (!((a.m >> (63 - la)) & 1))
        la++;
    while (!((b.m >> (63 - lb)) & 1))
        lb++;
    const unsigned __int128 am = a.m << la, bm = b.m << lb;
    const unsigned __int128 num = am << 64;
    const unsigned __int128 q = num / bm, rem = num - q * bm;
    XS r{a.neg != b.neg, (q << 1) | (rem != 0 ? 1 : 0), (a.e - la) - (b.e - lb) - 65};
    round_sig(r.m, r.e, 64);
    return r;
}

// gxx gyy - gxy gxy for floats, the exact difference of the two exact products rounded once to 64 bits.
__host__ __device__ inline XS xs_det(float gxx, float gyy, float gxy)
{
    const XS p = xs_mul(xs_from_float(gxx), xs_from_float(gyy)), q = xs_mul(xs_from_float(gxy), xs_from_float(gxy));
    // p and q are exact (48 significant bits at most); align both below bit 110 of one accumulator
    if (!p.m && !q.m)
        return XS{false, 0, 0};
    const int tp = p.m ? p.e + 48 : -100000, tq = q.m ? q.e + 48 : -100000;
    const int base = (tp > tq ? tp : tq) - 110;
    const unsigned __int128 x = p.m ? shift_jam(p.m, p.e - base) : 0, y = q.m ? shift_jam(q.m, q.e - base) : 0;
    // p >= 0 (gxx, gyy are sums of squares) and q >= 0
    XS r;
    if (x >= y)
        r = XS{p.neg, x - y, base};
    else
        r = XS{!p.neg, y - x, base};
    if (r.m)
        round_sig(r.m, r.e, 64);
    return r;
}

__host__ __device__ inline float xs_to_float(XS a)
{
    if (!a.m)
        return a.neg ? -0.0f : 0.0f;
    round_sig(a.m, a.e, 24);
    const float f = ldexpf((float)(uint64_t)a.m, a.e);
    return a.neg ? -f : f;
}

// Whether rounding the double d to float is settled even if d is off by a few of its own ulps: its discarded 29
// significand bits are more than 16 ulps from the midpoint pattern (and d is a normal float's magnitude).
__host__ __device__ inline bool float_settled(double d)
{
    const double a = d < 0 ? -d : d;
    if (a == 0.0)
        return true;
    if (!(a > 0x1p-120 && a < 0x1p120))
        return false;
    uint64_t bits;
    __builtin_memcpy(&bits, &a, 8);
    const int64_t low = (int64_t)(bits & ((1ull << 29) - 1)) - (1ll << 28);
    return low > 16 || low < -16;
}

// The covariance s^2 G^-1 of lk_driver.inc:100-104 in the reference's x87 arithmetic: s2 = r1 area / (area - 2),
// then (float)(s2 gyy / det), (float)(-s2 gxy / det), (float)(s2 gxx / det). The fp64 result is taken where fp64 and
// x87 must round to the same float; elsewhere the exact emulation decides.
__host__ __device__ inline void cov_ld(float r1, int area, float gxx, float gyy, float gxy, float *cxx, float *cxy,
                                       float *cyy)
{
    const double det = (double)gxx * gyy - (double)gxy * gxy;
    const double s2 = (double)r1 * area / (area - 2);
    const double a = s2 * gyy / det, b = -s2 * gxy / det, c = s2 * gxx / det;
    if (float_settled(a) && float_settled(b) && float_settled(c) && det > 0) {
        *cxx = (float)a;
        *cxy = (float)b;
        *cyy = (float)c;
        return;
    }
    const XS d = xs_det(gxx, gyy, gxy);
    XS s = xs_div(xs_mul(xs_from_float(r1), xs_from_int(area)), xs_from_int(area - 2));
    *cxx = xs_to_float(xs_div(xs_mul(s, xs_from_float(gyy)), d));
    s.neg = !s.neg;
    *cxy = xs_to_float(xs_div(xs_mul(s, xs_from_float(gxy)), d));
    s.neg = !s.neg;
    *cyy = xs_to_float(xs_div(xs_mul(s, xs_from_float(gxx)), d));
}

}  // namespace vtg
